"""Host binding of the microphysics network of libcrm_hip.so (include/crm_mlp.h) and the
training driver that fits it to the GPU CRM's own microphysics.

The network maps the state of one grid point at the entry of micro_proc
(MICRO_SAM1MOM/microphysics.F90:292-357) -- T, p, rho, qv, qcl, qci, qpl, qpi, w -- to the
rates of change of qv, qcl, qci, qpl, qpi and T over that call.  Training pairs are recorded by
the CRM itself (crm_hip_record_micro_samples) while it runs the plain scheme.  Everything
runs on the GPU: sampling, the forward and backward passes and the Adam updates (BF16 matrix
cores, FP32 accumulation).
"""
import ctypes as C

import numpy as np

import crm_hip

NIN, NH, NOUT = 9, 128, 6
NPARAM = NIN * NH + NH + 2 * (NH * NH + NH) + NH * NOUT + NOUT
FEATURES = ["T", "p", "rho", "qv", "qcl", "qci", "qpl", "qpi", "w"]
TARGETS = ["dqv/dt", "dqcl/dt", "dqci/dt", "dqpl/dt", "dqpi/dt", "dT/dt"]
SYMBOLS = ["crm_mlp_create", "crm_mlp_destroy", "crm_mlp_set_params", "crm_mlp_get_params",
           "crm_mlp_set_norm", "crm_mlp_forward", "crm_mlp_grad", "crm_mlp_train_step",
           "crm_hip_set_micro_mlp", "crm_hip_record_micro_samples", "crm_hip_micro_samples_count"]

_bound = False


def _lib():
    global _bound
    lib = crm_hip.load()
    if not _bound:
        P, F = C.c_void_p, C.POINTER(C.c_float)
        lib.crm_mlp_create.argtypes = [C.POINTER(P), C.c_int]
        lib.crm_mlp_destroy.argtypes = [P]
        lib.crm_mlp_set_params.argtypes = [P, F]
        lib.crm_mlp_get_params.argtypes = [P, F]
        lib.crm_mlp_set_norm.argtypes = [P, F, F, F, F, F]
        lib.crm_mlp_forward.argtypes = [P, P, C.c_int, P]
        lib.crm_mlp_grad.argtypes = [P, P, P, C.c_int, F, F]
        lib.crm_mlp_train_step.argtypes = [P, P, P, C.c_int, C.c_float, F]
        lib.crm_hip_set_micro_mlp.argtypes = [P, P]
        lib.crm_hip_record_micro_samples.argtypes = [P, P, P, C.c_longlong, C.c_int]
        lib.crm_hip_micro_samples_count.argtypes = [P]
        lib.crm_hip_micro_samples_count.restype = C.c_longlong
        _bound = True
    return lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def unpack(p):
    """Packed parameters -> [W1, b1, W2, b2, W3, b3, W4, b4] (row-major, y = x W + b)."""
    shapes = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(np.asarray(p[o:o + n], dtype=np.float32).reshape(sh))
        o += n
    return out


def pack(mats):
    return np.concatenate([np.asarray(m, dtype=np.float32).ravel() for m in mats])


def init_params(seed=7):
    """He-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    mats = []
    for fan_in, fan_out in ((NIN, NH), (NH, NH), (NH, NH), (NH, NOUT)):
        lim = np.sqrt(6.0 / fan_in)
        mats += [rng.uniform(-lim, lim, (fan_in, fan_out)), np.zeros(fan_out)]
    return pack(mats)


class CrmMlp:
    """The network on one GPU (torch tensors for device arrays)."""

    def __init__(self, device=0):
        self.lib = _lib()
        self.ptr = C.c_void_p()
        rc = self.lib.crm_mlp_create(C.byref(self.ptr), device)
        if rc:
            raise crm_hip.CrmError(rc, "crm_mlp_create")
        self.norm = None

    def close(self):
        if self.ptr:
            self.lib.crm_mlp_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAM
        self._check(self.lib.crm_mlp_set_params(self.ptr, _fp(p)), "set_params")

    def params(self):
        p = np.zeros(NPARAM, dtype=np.float32)
        self._check(self.lib.crm_mlp_get_params(self.ptr, _fp(p)), "get_params")
        return p

    def set_norm(self, mean, scale, oscale, lo=None, hi=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (mean, scale, oscale)]
        lo = np.full(NIN, -np.inf, np.float32) if lo is None else np.ascontiguousarray(lo, np.float32)
        hi = np.full(NIN, np.inf, np.float32) if hi is None else np.ascontiguousarray(hi, np.float32)
        self.norm = dict(mean=arrs[0], scale=arrs[1], oscale=arrs[2], lo=lo, hi=hi)
        self._check(self.lib.crm_mlp_set_norm(self.ptr, *[_fp(a) for a in arrs], _fp(lo), _fp(hi)),
                    "set_norm")

    def forward(self, X):
        """X: torch float32 [n, 9] on the GPU -> [n, 6]."""
        import torch
        Y = torch.empty((X.shape[0], NOUT), dtype=torch.float32, device=X.device)
        self._check(self.lib.crm_mlp_forward(self.ptr, X.data_ptr(), X.shape[0], Y.data_ptr()),
                    "forward")
        return Y

    def grad(self, X, T):
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.lib.crm_mlp_grad(self.ptr, X.data_ptr(), T.data_ptr(), X.shape[0], _fp(g),
                                          _fp(loss)), "grad")
        return g, float(loss[0])

    def train_step(self, X, T, lr):
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.lib.crm_mlp_train_step(self.ptr, X.data_ptr(), T.data_ptr(), X.shape[0],
                                                C.c_float(lr), _fp(loss)), "train_step")
        return float(loss[0])

    def _check(self, rc, what):
        if rc:
            raise crm_hip.CrmError(rc, "crm_mlp_" + what)


def attach(hip, mlp):
    """Use the network for the microphysics of later calls of a CrmHip context (None: plain)."""
    rc = _lib().crm_hip_set_micro_mlp(hip.ctx, mlp.ptr if mlp is not None else None)
    if rc:
        raise crm_hip.CrmError(rc, "crm_hip_set_micro_mlp")


def record(hip, cap, stride, device="cuda:0"):
    """Start recording (input, target) pairs of the plain scheme into new device buffers;
    returns (X, T) torch tensors [cap, 9], [cap, 6]."""
    import torch
    X = torch.zeros((cap, NIN), dtype=torch.float32, device=device)
    T = torch.zeros((cap, NOUT), dtype=torch.float32, device=device)
    rc = _lib().crm_hip_record_micro_samples(hip.ctx, X.data_ptr(), T.data_ptr(), cap, stride)
    if rc:
        raise crm_hip.CrmError(rc, "crm_hip_record_micro_samples")
    return X, T


def stop_recording(hip):
    n = int(_lib().crm_hip_micro_samples_count(hip.ctx))
    _lib().crm_hip_record_micro_samples(hip.ctx, None, None, 0, 1)
    return n


def fit_norm(X, T):
    """Normalisation and training envelope from recorded pairs (torch tensors)."""
    x = X.double()
    t = T.double()
    mean = x.mean(0).float().cpu().numpy()
    scale = x.std(0).clamp_min(1e-12).float().cpu().numpy()
    oscale = t.pow(2).mean(0).sqrt().clamp_min(1e-20).float().cpu().numpy()
    lo = X.min(0).values.cpu().numpy()
    hi = X.max(0).values.cpu().numpy()
    return mean, scale, oscale, lo, hi


def train(mlp, X, T, steps, batch=16384, lr=1e-3, seed=7, log=None, cosine=True):
    """Adam on random mini-batches of the recorded pairs, learning rate on a cosine decay to
    zero (or constant); returns the loss history."""
    import math

    import torch
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    hist = []
    n = X.shape[0]
    for s in range(steps):
        idx = torch.randint(0, n, (min(batch, n),), device=X.device, generator=g)
        xb = X.index_select(0, idx).contiguous()
        tb = T.index_select(0, idx).contiguous()
        lr_s = lr * 0.5 * (1.0 + math.cos(math.pi * s / steps)) if cosine else lr
        hist.append(mlp.train_step(xb, tb, max(lr_s, 1e-7)))
        if log and (s + 1) % max(1, steps // 10) == 0:
            log(f"step {s + 1}/{steps} loss {hist[-1]:.4e}")
    return hist
