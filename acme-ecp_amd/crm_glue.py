"""Host binding of the GCM-side kernels of libcrm_hip.so (include/crm_glue.h): crm_input from
the physics state, the surface-flux bypass, the tendencies after the call, check_energy_chng
and the CRM side of the radiation, all on device arrays (torch tensors in the physics types'
Fortran layout: a (pcols, pver) array is a contiguous tensor of shape (pver, pcols), a
(pcols, nrad, pver) array one of shape (pver, nrad, pcols))."""
import ctypes as C

import crm_hip

STATE = ["zm", "zi", "t", "qv", "qcl", "qci", "ps", "pmid", "pint", "pdel", "phis", "u", "v"]
SURFACE = ["ocnfrac", "wsx", "wsy", "shf", "lhf"]
SYMBOLS = ["crm_hip_gcm_input", "crm_hip_gcm_flux_bypass", "crm_hip_gcm_tend", "crm_hip_check_energy_chng",
           "crm_hip_rad_columns", "crm_hip_rad_qrad"]
CLOUD = ["cld", "fice", "cicewp", "cliqwp"]
RAD_COLUMN = ["t", "qv", "qc", "qi", "cld", "fice", "cicewp", "cliqwp"]


class GcmStateView(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in STATE]


class GcmSurfaceView(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in SURFACE]


class GcmCloudView(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in CLOUD]


class RadColumnView(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in RAD_COLUMN]


_bound = False


def _lib():
    global _bound
    lib = crm_hip.load()
    if not _bound:
        P, I, D = C.c_void_p, C.c_int, C.c_double
        lib.crm_hip_gcm_input.argtypes = [P, I, I, I, I, I, P, P, P, P, P]
        lib.crm_hip_gcm_flux_bypass.argtypes = [P, I, I, I, P, P, P, P, P]
        lib.crm_hip_gcm_tend.argtypes = [P, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P]
        lib.crm_hip_check_energy_chng.argtypes = [P, I, I, I, P, D] + [P] * 10
        lib.crm_hip_rad_columns.argtypes = [P, I, I, I, I, I, P, P, P, P]
        lib.crm_hip_rad_qrad.argtypes = [P, I, I, I, I, P, P, P]
        _bound = True
    return lib


def _p(t):
    return C.c_void_p(t.data_ptr())


def _check(rc, what):
    if rc:
        raise crm_hip.CrmError(rc, what)


def state_view(state):
    return GcmStateView(*[state[f].data_ptr() for f in STATE])


def gcm_input(hip, ncol, pcols, pver, crm_nz, nrad, state, srf, angle, dcase):
    """crm_input of dcase (a crm_hip.DeviceCase) from the physics state and surface fluxes;
    its qrad is converted from Q dp to a heating rate."""
    vi, _, vr, _ = dcase.views()
    sv, fv = state_view(state), GcmSurfaceView(*[srf[f].data_ptr() for f in SURFACE])
    _check(_lib().crm_hip_gcm_input(hip.ctx, ncol, pcols, pver, crm_nz, nrad, C.byref(sv), C.byref(fv), _p(angle),
                                    C.byref(vi), C.byref(vr)), "crm_hip_gcm_input")


def flux_bypass(hip, ncol, pcols, pver, pdel, shf, cflx, ptend_s, ptend_q):
    _check(_lib().crm_hip_gcm_flux_bypass(hip.ctx, ncol, pcols, pver, _p(pdel), _p(shf), _p(cflx), _p(ptend_s),
                                          _p(ptend_q)), "crm_hip_gcm_flux_bypass")


def gcm_tend(hip, ncol, pcols, pver, crm_nz, nrad, pdel, dcase, qrs, qrl, ptend, sp_rad_flux):
    """ptend: dict s, qv, qcl, qci of device arrays."""
    _, _, vr, vo = dcase.views()
    _check(_lib().crm_hip_gcm_tend(hip.ctx, ncol, pcols, pver, crm_nz, nrad, _p(pdel), C.byref(vo), C.byref(vr),
                                   _p(qrs), _p(qrl), _p(ptend["s"]), _p(ptend["qv"]), _p(ptend["qcl"]),
                                   _p(ptend["qci"]), _p(sp_rad_flux)), "crm_hip_gcm_tend")


def check_energy_chng(hip, ncol, pcols, pver, state, ztodt, flx, cur, tnd, rer):
    """flx: dict vap, cnd, ice, sen; cur, tnd, rer: dicts te, tw (device arrays)."""
    sv = state_view(state)
    _check(_lib().crm_hip_check_energy_chng(hip.ctx, ncol, pcols, pver, C.byref(sv), C.c_double(ztodt),
                                            _p(flx["vap"]), _p(flx["cnd"]), _p(flx["ice"]), _p(flx["sen"]),
                                            _p(cur["te"]), _p(cur["tw"]), _p(tnd["te"]), _p(tnd["tw"]),
                                            _p(rer["te"]), _p(rer["tw"])), "crm_hip_check_energy_chng")


def rad_columns(hip, ncol, pcols, pver, crm_nz, nrad, state, cloud, dcase, out):
    """The radiation's inputs in every CRM radiation column from the physics state (dict of
    device arrays), the GCM cloud fields (dict cld, fice, cicewp, cliqwp) and the crm_rad
    averages of dcase (a crm_hip.DeviceCase); out: dict of (pver, nrad, pcols) tensors."""
    _, _, vr, _ = dcase.views()
    sv = state_view(state)
    cv = GcmCloudView(*[cloud[f].data_ptr() for f in CLOUD])
    ov = RadColumnView(*[out[f].data_ptr() for f in RAD_COLUMN])
    _check(_lib().crm_hip_rad_columns(hip.ctx, ncol, pcols, pver, crm_nz, nrad, C.byref(sv), C.byref(cv),
                                      C.byref(vr), C.byref(ov)), "crm_hip_rad_columns")


def rad_qrad(hip, ncol, pcols, nrad, crm_nz, qrs_crm, qrl_crm, dcase):
    """dcase's crm_rad qrad = qrs_crm + qrl_crm (device arrays in qrad's layout)."""
    _, _, vr, _ = dcase.views()
    _check(_lib().crm_hip_rad_qrad(hip.ctx, ncol, pcols, nrad, crm_nz, _p(qrs_crm), _p(qrl_crm), C.byref(vr)),
           "crm_hip_rad_qrad")
