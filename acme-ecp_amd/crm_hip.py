"""Host binding of libcrm_hip.so (include/crm_hip.h) for Python callers.

The drop-in target is crm() (components/cam/src/physics/crm/crm_module.F90:45-56): one call
advances every CRM instance of a GCM chunk by dt_gl.  `CrmHip.run(case, dt_gl)` does that on
host arrays (a `crm_views.CrmCase`), `CrmHip.run_device(...)` on arrays already resident in
HBM.  Status codes of the library become `CrmError`, with the same meaning as the reference's
aborts (kurant.F90:64-75, crm_module.F90:686-692, damping.F90:39-42).

There is no CPU fallback: without the built library every entry point raises.
"""
import ctypes as C
import os

import numpy as np

from crm_views import (CrmColumnInfo, CrmDims, CrmInputView, CrmOutputView, CrmRadView,
                       CrmStateView, views_from_addresses)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CRM_HIP_LIB") or os.path.join(HERE, "lib", "libcrm_hip.so")

CRM_OK, CRM_ERR_ARG, CRM_ERR_CFL, CRM_ERR_NXRAD, CRM_ERR_DAMPING = 0, 1, 2, 3, 4
CRM_ERR_NAN, CRM_ERR_DEVICE, CRM_ERR_NOMEM, CRM_ERR_UNSUPPORTED, CRM_ERR_ACCEL = 5, 6, 7, 8, 9
CRM_ERR_SCRATCH = 10
NOALLOC_SCRATCH_BYTES = 256 << 20
PRECISION = {"fp64": 0, "mixed": 1}

# Every function include/crm_hip.h declares.
ABI_SYMBOLS = ["crm_hip_init", "crm_hip_run", "crm_hip_run_device", "crm_hip_run_device_noalloc",
               "crm_hip_accel_init", "crm_hip_set_coriolis",
               "crm_hip_run_many", "crm_hip_set_schedule", "crm_hip_get_schedule",
               "crm_hip_set_variants", "crm_hip_get_variants",
               "crm_hip_autotune", "crm_hip_set_graphs", "crm_hip_get_stats", "crm_hip_finalize",
               "crm_hip_errstr"]
# include/crm_hip_debug.h
DEBUG_SYMBOLS = ["crm_hip_debug_run", "crm_hip_debug_get", "crm_hip_fft991_operators",
                 "crm_hip_profile_kernels", "crm_hip_kernel_time", "crm_hip_kernel_name",
                 "crm_hip_graph_launches", "crm_hip_graph_cache_size",
                 "crm_hip_debug_stop_call_at"]
# Kernel classes of one call (CRM_K_* of include/crm_hip_debug.h), in launch order.
KERNELS = ["setup", "init", "setperturb", "kurant", "precip_init", "pre", "sgs", "mom", "press",
           "mpdata", "precip", "diffuse", "micro", "stats", "final"]


class CrmHipConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("max_ncrms", C.c_int)]


class CrmHipStats(C.Structure):
    _fields_ = [("nstop", C.c_int), ("ncycle_max", C.c_int), ("nprec_max", C.c_int),
                ("ms_device", C.c_double), ("accel_ceased", C.c_int)]


class CrmClassDesc(C.Structure):
    _fields_ = [("dims", CrmDims), ("col", CrmColumnInfo), ("inp", CrmInputView),
                ("state", CrmStateView), ("rad", CrmRadView), ("out", CrmOutputView)]


class CrmError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {errstr(code)} (status {code})")


_lib = None


def load():
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP CRM library missing: {LIB_PATH} (build it first)")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 and loads it under a
    # file name that does not match the soname, so if this library pulled in the system copy
    # first, torch would start a second runtime that finds no GPU.  Loading torch first lets
    # the library bind to torch's runtime by soname.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.crm_hip_init.argtypes = [C.POINTER(P), C.POINTER(CrmHipConfig)]
    views = [P, C.POINTER(CrmDims), C.POINTER(CrmColumnInfo), C.c_double,
             C.POINTER(CrmInputView), C.POINTER(CrmStateView), C.POINTER(CrmRadView),
             C.POINTER(CrmOutputView)]
    lib.crm_hip_run.argtypes = views
    lib.crm_hip_run_device.argtypes = views
    lib.crm_hip_run_device_noalloc.argtypes = views
    lib.crm_hip_debug_run.argtypes = views + [C.c_int]
    lib.crm_hip_debug_get.argtypes = [P, C.c_char_p, C.POINTER(C.c_double)]
    lib.crm_hip_get_stats.argtypes = [P, C.POINTER(CrmHipStats)]
    lib.crm_hip_accel_init.argtypes = [P, C.c_int, C.c_double, C.c_int]
    lib.crm_hip_set_coriolis.argtypes = [P, C.c_int]
    IA = C.c_int * len(KERNELS)
    lib.crm_hip_set_schedule.argtypes = [P, IA]
    lib.crm_hip_get_schedule.argtypes = [P, IA]
    lib.crm_hip_set_variants.argtypes = [P, IA]
    lib.crm_hip_get_variants.argtypes = [P, IA]
    lib.crm_hip_autotune.argtypes = views + [IA]
    lib.crm_hip_set_graphs.argtypes = [P, C.c_int]
    lib.crm_hip_debug_stop_call_at.argtypes = [P, C.c_int]
    lib.crm_hip_graph_launches.argtypes = [P]
    lib.crm_hip_graph_launches.restype = C.c_longlong
    lib.crm_hip_graph_cache_size.argtypes = [P]
    lib.crm_hip_run_many.argtypes = [C.c_int, C.POINTER(P), C.POINTER(CrmClassDesc), C.c_double,
                                     C.POINTER(C.c_int)]
    lib.crm_hip_finalize.argtypes = [P]
    lib.crm_hip_errstr.argtypes = [C.c_int]
    lib.crm_hip_errstr.restype = C.c_char_p
    lib.crm_hip_fft991_operators.argtypes = [C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]
    lib.crm_hip_profile_kernels.argtypes = [P, C.c_uint]
    lib.crm_hip_kernel_time.argtypes = [P, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_longlong)]
    lib.crm_hip_kernel_name.argtypes = [C.c_int]
    lib.crm_hip_kernel_name.restype = C.c_char_p
    _lib = lib
    return lib


def errstr(code):
    return load().crm_hip_errstr(code).decode()


def fft991_operators(n):
    """Dense forward (Ft[i][s]) and inverse (Gt[s][i]) matrices of the reference's x-transform
    (crm/fft.F90:5-153) in the library's coefficient-slot order."""
    Ft, Gt = np.zeros((n, n)), np.zeros((n, n))
    dp = C.POINTER(C.c_double)
    rc = load().crm_hip_fft991_operators(n, Ft.ctypes.data_as(dp), Gt.ctypes.data_as(dp))
    if rc:
        raise CrmError(rc, "crm_hip_fft991_operators")
    return Ft, Gt


class CrmHip:
    """One library context: a HIP stream and persistent device buffers on one GPU."""

    def __init__(self, device=0, precision="fp64", max_ncrms=0):
        self.lib = load()
        self.ctx = C.c_void_p()
        cfg = CrmHipConfig(device, PRECISION[precision], max_ncrms)
        rc = self.lib.crm_hip_init(C.byref(self.ctx), C.byref(cfg))
        if rc:
            raise CrmError(rc, "crm_hip_init")

    def close(self):
        if self.ctx:
            self.lib.crm_hip_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, fn, dims, col, dt_gl, views, *extra):
        vi, vs, vr, vo = views
        rc = fn(self.ctx, C.byref(dims), C.byref(col), C.c_double(dt_gl), C.byref(vi),
                C.byref(vs), C.byref(vr), C.byref(vo), *extra)
        return rc

    def run(self, case, dt_gl, check=True):
        """crm() on a CrmCase's host arrays, in place.  Returns the status code."""
        rc = self._call(self.lib.crm_hip_run, case.dims(), case.column_info(), dt_gl,
                        case.views())
        if rc and check:
            raise CrmError(rc, "crm_hip_run")
        return rc

    def run_device(self, dcase, dt_gl, check=True):
        """crm() on a DeviceCase (arrays resident on the GPU), in place."""
        rc = self._call(self.lib.crm_hip_run_device, dcase.case.dims(),
                        dcase.case.column_info(), dt_gl, dcase.views())
        if rc and check:
            raise CrmError(rc, "crm_hip_run_device")
        return rc

    def run_device_noalloc(self, dcase, dt_gl, check=True):
        """crm() on a DeviceCase through the entry point that allocates nothing (every device
        buffer of the call in the library's fixed per-GPU array)."""
        rc = self._call(self.lib.crm_hip_run_device_noalloc, dcase.case.dims(),
                        dcase.case.column_info(), dt_gl, dcase.views())
        if rc and check:
            raise CrmError(rc, "crm_hip_run_device_noalloc")
        return rc

    def accel_init(self, use_crm_accel, crm_accel_factor=0.0, crm_accel_uv=False):
        """Mean-state acceleration for later calls (crm_accel_init, accelerate_crm.F90:30-38,45-75)."""
        rc = self.lib.crm_hip_accel_init(self.ctx, int(bool(use_crm_accel)), float(crm_accel_factor),
                                         int(bool(crm_accel_uv)))
        if rc:
            raise CrmError(rc, "crm_hip_accel_init")

    def set_coriolis(self, on):
        """The reference's docoriolis switch (params.F90:107) for later calls (3D CRMs)."""
        rc = self.lib.crm_hip_set_coriolis(self.ctx, int(bool(on)))
        if rc:
            raise CrmError(rc, "crm_hip_set_coriolis")

    def schedule(self):
        """{kernel class: threads per workgroup} of this context."""
        a = (C.c_int * len(KERNELS))()
        self.lib.crm_hip_get_schedule(self.ctx, a)
        return dict(zip(KERNELS, list(a)))

    def set_schedule(self, sched):
        cur = self.schedule()
        cur.update(sched)
        a = (C.c_int * len(KERNELS))(*[cur[k] for k in KERNELS])
        rc = self.lib.crm_hip_set_schedule(self.ctx, a)
        if rc:
            raise CrmError(rc, "crm_hip_set_schedule")

    def variants(self):
        """{kernel class: launch variant} of this context."""
        a = (C.c_int * len(KERNELS))()
        self.lib.crm_hip_get_variants(self.ctx, a)
        return dict(zip(KERNELS, list(a)))

    def set_variants(self, variants):
        cur = self.variants()
        cur.update(variants)
        a = (C.c_int * len(KERNELS))(*[cur[k] for k in KERNELS])
        rc = self.lib.crm_hip_set_variants(self.ctx, a)
        if rc:
            raise CrmError(rc, "crm_hip_set_variants")

    def autotune(self, dcase, dt_gl):
        """Tune the workgroup size of every kernel class on a DeviceCase (advanced in place by
        the trial calls: pass a scratch copy); returns the chosen schedule."""
        a = (C.c_int * len(KERNELS))()
        rc = self._call(self.lib.crm_hip_autotune, dcase.case.dims(), dcase.case.column_info(),
                        dt_gl, dcase.views(), a)
        if rc:
            raise CrmError(rc, "crm_hip_autotune")
        return dict(zip(KERNELS, list(a)))

    def set_graphs(self, mode):
        """0 (or False): plain launches; 1: graphs per steady substep; 2 (or True, the
        default): one graph per call with the subcycle decision on the device; 3: as 2, and
        run_many over contexts that all have mode 3 captures the classes into one graph."""
        if isinstance(mode, bool):
            mode = 2 if mode else 0
        rc = self.lib.crm_hip_set_graphs(self.ctx, int(mode))
        if rc:
            raise CrmError(rc, "crm_hip_set_graphs")

    def graph_launches(self):
        return int(self.lib.crm_hip_graph_launches(self.ctx))

    def graph_cache_size(self):
        return int(self.lib.crm_hip_graph_cache_size(self.ctx))

    def debug_stop_call_at(self, nstep):
        """Test hook: captured calls stop at substep nstep (0: off) and resume on the host."""
        rc = self.lib.crm_hip_debug_stop_call_at(self.ctx, int(nstep))
        if rc:
            raise CrmError(rc, "crm_hip_debug_stop_call_at")

    def debug_run(self, case, dt_gl, phase_stop):
        return self._call(self.lib.crm_hip_debug_run, case.dims(), case.column_info(), dt_gl,
                          case.views(), C.c_int(phase_stop))

    def debug_get(self, name, size):
        buf = np.zeros(size)
        n = self.lib.crm_hip_debug_get(self.ctx, name.encode(),
                                       buf.ctypes.data_as(C.POINTER(C.c_double)))
        if n < 0:
            raise KeyError(name)
        return buf[:n]

    def profile_kernels(self, names):
        """Time every launch of the named kernel classes with HIP events on the library's
        stream from now on (resets the totals); an empty list turns timing off."""
        mask = 0
        for n in names:
            mask |= 1 << KERNELS.index(n)
        rc = self.lib.crm_hip_profile_kernels(self.ctx, mask)
        if rc:
            raise CrmError(rc, "crm_hip_profile_kernels")

    def kernel_times(self):
        """{kernel class: (total ms, launches)} since the last profile_kernels()."""
        out = {}
        for kid, n in enumerate(KERNELS):
            ms, cnt = C.c_double(), C.c_longlong()
            self.lib.crm_hip_kernel_time(self.ctx, kid, C.byref(ms), C.byref(cnt))
            if cnt.value:
                out[n] = (ms.value, cnt.value)
        return out

    def stats(self):
        s = CrmHipStats()
        rc = self.lib.crm_hip_get_stats(self.ctx, C.byref(s))
        if rc:
            raise CrmError(rc, "crm_hip_get_stats")
        return s


class DeviceCase:
    """Device-resident copies (torch tensors in HBM) of a CrmCase's arrays.

    Each array keeps the caller's column-major layout: a Fortran (ld, a, b, c) array becomes a
    contiguous torch tensor of shape (c, b, a, ld)."""

    def __init__(self, case, device="cuda:0"):
        import torch
        self.case = case
        self.torch = torch
        self.arrays = {}
        for group in ("inp", "state", "rad", "out"):
            self.arrays[group] = {k: torch.from_numpy(np.ascontiguousarray(v.T)).to(device)
                                  for k, v in getattr(case, group).items()}

    def views(self):
        addr = {g: {k: t.data_ptr() for k, t in d.items()} for g, d in self.arrays.items()}
        return views_from_addresses(addr["inp"], addr["state"], addr["rad"], addr["out"])

    def download(self, case=None):
        """Copy state, rad and output arrays back into case (default: the source case)."""
        case = case or self.case
        for group in ("state", "rad", "out"):
            dst = getattr(case, group)
            for k, t in self.arrays[group].items():
                dst[k][...] = t.cpu().numpy().T
        return case


def run_many(hips, dcases, dt_gl, check=True):
    """Advance several DeviceCases (one per CrmHip context, uneven size classes) by dt_gl
    concurrently (crm_hip_run_many); returns the per-class status codes."""
    lib = load()
    n = len(hips)
    descs = (CrmClassDesc * n)()
    keep = []
    for j, (h, dc) in enumerate(zip(hips, dcases)):
        dims, col = dc.case.dims(), dc.case.column_info()
        vi, vs, vr, vo = dc.views()
        keep.append((dims, col, vi, vs, vr, vo))
        descs[j].dims, descs[j].col = dims, col
        descs[j].inp, descs[j].state, descs[j].rad, descs[j].out = vi, vs, vr, vo
    ctxs = (C.c_void_p * n)(*[h.ctx.value for h in hips])
    status = (C.c_int * n)()
    rc = lib.crm_hip_run_many(n, ctxs, descs, C.c_double(dt_gl), status)
    if rc and check:
        raise CrmError(rc, "crm_hip_run_many")
    return list(status)


def _variant_count(kind, dims):
    """Launch variants of a kernel class for a CRM shape (variants_of, crm_driver.hpp): the
    pressure solve has five at 64 columns in 2D and one elsewhere; MPDATA has six."""
    if kind == "press":
        return 5 if dims.nx == 64 and dims.ny == 1 else 1
    if kind == "mpdata":
        return 6
    return 1


def tune_many(hips, dcases, dt_gl, kinds=("press",), margin=0.02):
    """Choose launch variants for size classes advanced together by run_many, timing the whole
    batch instead of each class alone (CrmHip.autotune).  Run concurrently, the classes share
    the CUs, and a variant that wins alone with a larger workgroup (the eight-wave pressure
    workgroup, which needs eight free wave slots on one CU at once) can lose in the batch.

    Coordinate descent over the classes in order and, per class, over each kernel class in
    kinds: one trial per candidate variant, each a batch call timed after a call that captures
    the changed graphs, every call from the same saved state.  A candidate replaces the current
    variant only when the batch is more than margin faster with it.  The trials advance the
    DeviceCases, which are then restored to the state they had on entry.  The default searches
    the pressure variants only: MPDATA's variants measured within 1% of each other in a batch.
    Returns the chosen {kernel class: variant} of each context and the number of trials."""
    import time

    import torch
    saved = [{g: {k: t.clone() for k, t in d.items()} for g, d in dc.arrays.items() if g != "inp"}
             for dc in dcases]

    def restore():
        for dc, s in zip(dcases, saved):
            for g, d in s.items():
                for k, t in d.items():
                    dc.arrays[g][k].copy_(t)
        torch.cuda.synchronize()

    def trial():
        restore()
        run_many(hips, dcases, dt_gl)
        restore()
        t0 = time.perf_counter()
        run_many(hips, dcases, dt_gl)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    trials = 1
    best = trial()
    for h, dc in zip(hips, dcases):
        dims = dc.case.dims()
        for kind in kinds:
            cur = h.variants()[kind]
            for v in range(_variant_count(kind, dims)):
                if v == cur:
                    continue
                keep = h.variants()
                h.set_variants({kind: v})
                t = trial()
                trials += 1
                if t < best * (1.0 - margin):
                    best, cur = t, v
                else:
                    h.set_variants(keep)
    restore()
    return [h.variants() for h in hips], trials
