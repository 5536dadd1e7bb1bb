"""CPU reference of the microphysics network (include/crm_mlp.h), the checker of its GPU
forward and backward passes.  TEST INFRASTRUCTURE ONLY.

rounding="bf16" rounds exactly where the GPU stores BF16 (normalised inputs, weights,
activations, output and hidden gradients; round to nearest, ties to even) and accumulates in
FP32 (in another order than the matrix cores, so agreement is to FP32 accumulation error);
rounding=None is the exact-arithmetic (FP64) network, whose gradient is checked against
finite differences."""
import numpy as np

NIN, NH, NOUT = 9, 128, 6


def bf16(x):
    """Round to BF16 (nearest, ties to even), returned as float32."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    shapes = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(np.asarray(p[o:o + n]).reshape(sh))
        o += n
    return out


def _r(x, rounding):
    return bf16(x) if rounding == "bf16" else x


def forward(p, X, mean, scale, oscale, rounding="bf16"):
    """Raw inputs X[n,9] -> (de-normalised outputs [n,6], cache for backward)."""
    dt = np.float32 if rounding == "bf16" else np.float64
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(a, dtype=dt) for a in unpack(p)]
    rs = (dt(1) / np.asarray(scale, dt)).astype(dt)  # the reciprocal, rounded once
    Xn = _r(((np.asarray(X, dt) - np.asarray(mean, dt)) * rs).astype(dt), rounding)
    Ws = [_r(W, rounding) for W in (W1, W2, W3, W4)]
    H1 = _r(np.maximum(Xn @ Ws[0] + b1, 0), rounding)
    H2 = _r(np.maximum(H1 @ Ws[1] + b2, 0), rounding)
    H3 = _r(np.maximum(H2 @ Ws[2] + b3, 0), rounding)
    Yn = H3 @ Ws[3] + b4
    return Yn * oscale, (Xn, H1, H2, H3, Yn, Ws)


def loss_grad(p, X, T, mean, scale, oscale, rounding="bf16"):
    """L = mean over n x 6 of (y' - T/oscale)^2 and dL/dp (packed)."""
    n = X.shape[0]
    _, (Xn, H1, H2, H3, Yn, Ws) = forward(p, X, mean, scale, oscale, rounding)
    d = Yn - T / oscale
    L = float((d * d).sum() / (6 * n))
    dY = _r(2 * d / (6 * n), rounding)
    gW4, gb4 = H3.T @ dY, dY.sum(0)
    dH3 = _r((dY @ Ws[3].T) * (H3 > 0), rounding)
    gW3, gb3 = H2.T @ dH3, dH3.sum(0)
    dH2 = _r((dH3 @ Ws[2].T) * (H2 > 0), rounding)
    gW2, gb2 = H1.T @ dH2, dH2.sum(0)
    dH1 = _r((dH2 @ Ws[1].T) * (H1 > 0), rounding)
    gW1, gb1 = Xn.T @ dH1, dH1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return L, g
