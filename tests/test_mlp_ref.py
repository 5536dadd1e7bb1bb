"""The CPU reference of the microphysics network, pinned before it checks the GPU: its
analytic gradient (exact arithmetic) against central finite differences of its loss, and its
BF16 rounding against the definition of round-to-nearest-even."""
import numpy as np

import mlp_ref


def _setup(n=40, seed=3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp_ref.NIN * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 6 + 6)
    X = rng.normal(0, 1, (n, 9))
    T = rng.normal(0, 1, (n, 6))
    mean, scale, oscale = rng.normal(0, 0.1, 9), rng.uniform(0.5, 2, 9), rng.uniform(0.5, 2, 6)
    return p, X, T, mean, scale, oscale


def test_gradient_matches_finite_differences():
    p, X, T, mean, scale, oscale = _setup()
    L, g = mlp_ref.loss_grad(p, X, T, mean, scale, oscale, rounding=None)
    rng = np.random.default_rng(5)
    # a sample of parameters from every block (weights and biases of all four layers)
    offs = np.cumsum([0, 9 * 128, 128, 128 * 128, 128, 128 * 128, 128, 128 * 6, 6])
    idx = np.concatenate([rng.integers(offs[b], offs[b + 1], 6) for b in range(8)])
    h = 1e-6
    for j in idx:
        pp, pm = p.copy(), p.copy()
        pp[j] += h
        pm[j] -= h
        fd = (mlp_ref.loss_grad(pp, X, T, mean, scale, oscale, None)[0] -
              mlp_ref.loss_grad(pm, X, T, mean, scale, oscale, None)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(g[j])), (j, fd, g[j])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.1415926, 65504.0, 1e-20], np.float32)
    r = mlp_ref.bf16(x)
    # representable values stay; halfway cases go to the even neighbour; others to nearest
    assert r[0] == 1.0 and r[3] == 1.0078125
    assert r[1] == 1.0                        # halfway between 1 and 1+2^-7: even is 1
    assert r[2] == 1.0078125                  # above halfway
    assert abs(r[4] - x[4]) <= abs(x[4]) * 2.0 ** -8
    bits = r.view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)
