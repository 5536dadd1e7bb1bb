"""Host binding of the device history of libcrm_hip.so (include/crm_history.h): time-mean
history fields of the CRM (the outfld calls of crm_physics_tend,
components/cam/src/physics/cam/crm_physics.F90:1201-1461) accumulated in HBM.  Arrays are
torch tensors in the Fortran layout of their sources (a (pcols, pver) array is a contiguous
tensor of shape (pver, pcols))."""
import ctypes as C

import crm_hip

SYMBOLS = ["crm_hist_create", "crm_hist_accumulate", "crm_hist_nfields", "crm_hist_name",
           "crm_hist_size", "crm_hist_index", "crm_hist_mean", "crm_hist_reset", "crm_hist_destroy"]


class HistInputs(C.Structure):
    _fields_ = [("ncol", C.c_int), ("state", C.c_void_p), ("out", C.c_void_p), ("qrad", C.c_void_p),
                ("pmid", C.c_void_p), ("pdel", C.c_void_p), ("ptend_s", C.c_void_p),
                ("ptend_qv", C.c_void_p), ("ptend_qcl", C.c_void_p), ("ptend_qci", C.c_void_p),
                ("qrs", C.c_void_p), ("qrl", C.c_void_p)]


_bound = False


def _lib():
    global _bound
    lib = crm_hip.load()
    if not _bound:
        P, I = C.c_void_p, C.c_int
        lib.crm_hist_create.argtypes = [C.POINTER(P), I, I, I, I, I, I, I, I]
        lib.crm_hist_accumulate.argtypes = [P, P, C.POINTER(HistInputs)]
        lib.crm_hist_nfields.argtypes = []
        lib.crm_hist_name.argtypes = [I]
        lib.crm_hist_name.restype = C.c_char_p
        lib.crm_hist_size.argtypes = [P, I]
        lib.crm_hist_size.restype = C.c_longlong
        lib.crm_hist_index.argtypes = [C.c_char_p]
        lib.crm_hist_mean.argtypes = [P, P, I, P, C.POINTER(C.c_longlong)]
        lib.crm_hist_reset.argtypes = [P, P]
        lib.crm_hist_destroy.argtypes = [P]
        _bound = True
    return lib


def _check(rc, what):
    if rc:
        raise crm_hip.CrmError(rc, what)


def field_names():
    lib = _lib()
    return [lib.crm_hist_name(f).decode() for f in range(lib.crm_hist_nfields())]


class CrmHistory:
    """Device sums of every history field for chunks of at most pcols columns."""

    def __init__(self, device, pcols, pver, nx, ny, crm_nz, nx_rad, ny_rad):
        self.lib = _lib()
        self.h = C.c_void_p()
        self.pcols = pcols
        _check(self.lib.crm_hist_create(C.byref(self.h), device, pcols, pver, nx, ny, crm_nz, nx_rad, ny_rad),
               "crm_hist_create")

    def close(self):
        if self.h:
            self.lib.crm_hist_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def accumulate(self, hip, ncol, dcase=None, qrad=None, pmid=None, pdel=None, ptend=None, qrs=None,
                   qrl=None):
        """One outfld pass: crm_state/crm_output of a crm_hip.DeviceCase, and optional physics
        arrays (ptend: dict s, qv, qcl, qci); fields whose inputs are missing are skipped."""
        keep = []
        sv = ov = None
        if dcase is not None:
            _, sv, _, ov = dcase.views()
            keep += [sv, ov]
        pt = ptend or {}

        def p(t):
            return None if t is None else t.data_ptr()
        inp = HistInputs(ncol, C.cast(C.byref(sv), C.c_void_p) if sv is not None else None,
                         C.cast(C.byref(ov), C.c_void_p) if ov is not None else None, p(qrad), p(pmid),
                         p(pdel), p(pt.get("s")), p(pt.get("qv")), p(pt.get("qcl")), p(pt.get("qci")),
                         p(qrs), p(qrl))
        _check(self.lib.crm_hist_accumulate(self.h, hip.ctx, C.byref(inp)), "crm_hist_accumulate")
        del keep

    def mean(self, hip, name):
        """(time mean as a torch tensor of shape (per-column size, pcols), sample count)."""
        import torch
        f = self.lib.crm_hist_index(name.encode())
        if f < 0:
            raise KeyError(name)
        n = self.lib.crm_hist_size(self.h, f)
        dst = torch.empty((n, self.pcols), dtype=torch.float64, device="cuda:0")
        cnt = C.c_longlong()
        _check(self.lib.crm_hist_mean(self.h, hip.ctx, f, C.c_void_p(dst.data_ptr()), C.byref(cnt)),
               "crm_hist_mean")
        return dst, cnt.value

    def reset(self, hip):
        _check(self.lib.crm_hist_reset(self.h, hip.ctx), "crm_hist_reset")
