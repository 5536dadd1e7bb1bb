"""ctypes mirror of include/crm_hip.h and host-side buffers for one crm() call.

The reference passes Fortran derived types (crm_input_type, crm_state_type, crm_rad_type,
crm_output_type; crm/crm_input_module.F90:10-47, crm/crm_state_module.F90:12-56,
crm/crm_rad_module.F90:17-34, crm/crm_output_module.F90:7-133) whose arrays are
column-major with pcols leading.  `CrmCase` holds numpy arrays with exactly those shapes in
Fortran order and builds the pointer views that the C ABI takes.
"""
import ctypes as C

import numpy as np

c_dp = C.POINTER(C.c_double)


class CrmDims(C.Structure):
    _fields_ = [("ncrms", C.c_int), ("ld", C.c_int), ("plev", C.c_int), ("nx", C.c_int),
                ("ny", C.c_int), ("nz", C.c_int), ("nx_rad", C.c_int), ("ny_rad", C.c_int),
                ("dx", C.c_double), ("dt", C.c_double)]


class CrmColumnInfo(C.Structure):
    _fields_ = [("lchnk", C.c_int), ("gcol", C.POINTER(C.c_int)), ("lat", c_dp),
                ("lon", c_dp), ("igstep", C.c_int)]


INPUT_PLEV = ["zmid", "tl", "ql", "qccl", "qiil", "pmid", "pdel", "ul", "vl"]
INPUT_PLEV1 = ["zint", "pint"]
INPUT_COL = ["ps", "phis", "ocnfrac", "tau00", "wndls", "bflxls", "fluxu00", "fluxv00",
             "fluxt00", "fluxq00"]
INPUT_ORDER = ["zmid", "zint", "tl", "ql", "qccl", "qiil", "ps", "pmid", "pint", "pdel",
               "phis", "ul", "vl", "ocnfrac", "tau00", "wndls", "bflxls", "fluxu00",
               "fluxv00", "fluxt00", "fluxq00"]
STATE_ORDER = ["u_wind", "v_wind", "w_wind", "temperature", "qt", "qp", "qn"]
RAD_ORDER = ["qrad", "temperature", "qv", "qc", "qi", "cld"]
OUT_GRID = ["qcl", "qci", "qpl", "qpi", "tk", "tkh"]
OUT_SFC = ["prec_crm"]
OUT_COL = ["cltot", "clhgh", "clmed", "cllow", "precc", "precl", "precsc", "precsl",
           "jt_crm", "mx_crm", "prectend", "precstend", "taux", "tauy", "z0m",
           "timing_factor"]
OUT_PLEV = ["cldtop", "qc_mean", "qi_mean", "qs_mean", "qg_mean", "qr_mean", "sltend",
            "qltend", "qcltend", "qiltend", "cld", "gicewp", "gliqwp", "mctot", "mcup",
            "mcdn", "mcuup", "mcudn", "mu_crm", "md_crm", "du_crm", "eu_crm", "ed_crm",
            "flux_qt", "fluxsgs_qt", "tkez", "tkesgsz", "tkz", "flux_u", "flux_v",
            "flux_qp", "precflux", "qt_ls", "qt_trans", "qp_trans", "qp_fall", "qp_src",
            "qp_evp", "t_ls"]
OUT_ORDER = OUT_GRID + OUT_SFC + OUT_COL + OUT_PLEV


def _struct(name, order):
    return type(name, (C.Structure,), {"_fields_": [(f, c_dp) for f in order]})


CrmInputView = _struct("CrmInputView", INPUT_ORDER)
CrmStateView = _struct("CrmStateView", STATE_ORDER)
CrmRadView = _struct("CrmRadView", RAD_ORDER)
CrmOutputView = _struct("CrmOutputView", OUT_ORDER)


def _ptr(a):
    if a is None:
        return c_dp()
    if isinstance(a, int):        # a raw device address
        return C.cast(C.c_void_p(a), c_dp)
    assert a.dtype == np.float64 and a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(c_dp)


class CrmCase:
    """All caller-side arrays of one crm() call, in the reference's shapes and Fortran order."""

    def __init__(self, ncrms, plev, nx, ny, nz, dx, dt, nx_rad=None, ny_rad=None, ld=None):
        self.ncrms, self.plev = ncrms, plev
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nx_rad = nx if nx_rad is None else nx_rad
        self.ny_rad = ny if ny_rad is None else ny_rad
        self.dx, self.dt = float(dx), float(dt)
        self.ld = ncrms if ld is None else ld
        z = lambda *s: np.zeros(s, dtype=np.float64, order="F")
        L = self.ld
        self.inp = {f: z(L, plev) for f in INPUT_PLEV}
        self.inp.update({f: z(L, plev + 1) for f in INPUT_PLEV1})
        self.inp.update({f: z(L) for f in INPUT_COL})
        self.state = {f: z(L, nx, ny, nz) for f in STATE_ORDER}
        self.rad = {f: z(L, self.nx_rad, self.ny_rad, nz) for f in RAD_ORDER}
        self.out = {f: z(L, nx, ny, nz) for f in OUT_GRID}
        self.out["prec_crm"] = z(L, nx, ny)
        self.out.update({f: z(L) for f in OUT_COL})
        self.out.update({f: z(L, plev) for f in OUT_PLEV})
        self.gcol = np.arange(1, ncrms + 1, dtype=np.int32)
        self.lat = np.zeros(ncrms)
        self.lon = np.zeros(ncrms)
        self.lchnk = 0
        self.igstep = 2

    def copy(self):
        c = CrmCase.__new__(CrmCase)
        c.__dict__.update(self.__dict__)
        for d in ("inp", "state", "rad", "out"):
            setattr(c, d, {k: v.copy(order="F") for k, v in getattr(self, d).items()})
        for a in ("gcol", "lat", "lon"):
            setattr(c, a, getattr(self, a).copy())
        return c

    def dims(self):
        return CrmDims(self.ncrms, self.ld, self.plev, self.nx, self.ny, self.nz,
                       self.nx_rad, self.ny_rad, self.dx, self.dt)

    def column_info(self):
        return CrmColumnInfo(self.lchnk, self.gcol.ctypes.data_as(C.POINTER(C.c_int)),
                             _ptr(self.lat), _ptr(self.lon), self.igstep)

    def views(self):
        """(input, state, rad, output) views over the host arrays; keep self alive."""
        return (CrmInputView(*[_ptr(self.inp[f]) for f in INPUT_ORDER]),
                CrmStateView(*[_ptr(self.state[f]) for f in STATE_ORDER]),
                CrmRadView(*[_ptr(self.rad[f]) for f in RAD_ORDER]),
                CrmOutputView(*[_ptr(self.out[f]) for f in OUT_ORDER]))


def views_from_addresses(inp, state, rad, out):
    """Views over device arrays given as dicts name -> integer address."""
    return (CrmInputView(*[_ptr(inp[f]) for f in INPUT_ORDER]),
            CrmStateView(*[_ptr(state[f]) for f in STATE_ORDER]),
            CrmRadView(*[_ptr(rad[f]) for f in RAD_ORDER]),
            CrmOutputView(*[_ptr(out[f]) for f in OUT_ORDER]))
