"""Mean-state acceleration on the GPU (accelerate_crm.F90:109-366; crm_module.F90:718-720,
872-876) against the oracle's restatement (pinned by tests/test_oracle_accel.py) at the
parity bars of the plain CRM, and on the 3D CRM by the same identities the oracle is pinned
with."""
import numpy as np
import pytest

import oracle_lib
from crm_hip import CRM_ERR_ACCEL, CrmHip
from crm_inputs import make_case
from parity import compare_case, convective_case, relerr

pytestmark = pytest.mark.gpu

TOL_1, TOL_N = 1e-10, 1e-8


@pytest.fixture(scope="module")
def hip():
    h = CrmHip(0)
    yield h
    h.close()


@pytest.fixture
def accel(hip):
    def set_(use, factor=0.0, uv=False):
        oracle_lib.oracle_accel_init(use, factor, uv)
        hip.accel_init(use, factor, uv)
    yield set_
    oracle_lib.oracle_accel_init(0, 0.0, False)
    hip.accel_init(0, 0.0, False)


def _worst(errs):
    return sorted(errs.items(), key=lambda kv: -kv[1])[:4]


SHAPES = [(3, 32, 28, 4000.0, 20.0), (2, 64, 58, 1000.0, 5.0), (5, 16, 28, 4000.0, 20.0)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}")
@pytest.mark.parametrize("uv", [False, True])
@pytest.mark.parametrize("nsub", [1, 20])
def test_accelerated_parity(hip, accel, shape, uv, nsub):
    n, nx, nz, dx, dt = shape
    accel(1, 2.0, uv)
    case = make_case(n, nx, nz, dx, dt, randomize=True, seed_base=70 + nsub)
    dt_gl = 3 * dt * nsub
    want, got = case.copy(), case.copy()
    assert oracle_lib.oracle().run(want, dt_gl) == 0
    hip.run(got, dt_gl)
    st = hip.stats()
    assert st.nstop == nsub and st.accel_ceased == 0
    errs = compare_case(got, want, dt_gl=dt_gl)
    print(shape, uv, nsub, _worst(errs))
    assert max(errs.values()) <= (TOL_1 if nsub == 1 else TOL_N), _worst(errs)


def test_cease_parity(hip, accel):
    """A radiative heating of 1 K/s changes the level means by 20 K in the first substep: the
    acceleration stops and nstop returns to dt_gl / dt (accelerate_crm.F90:255-278)."""
    accel(1, 2.0)
    case = convective_case(4, seed=71)
    case.rad["qrad"][...] = 1.0
    want, got = case.copy(), case.copy()
    assert oracle_lib.oracle().run(want, 120.0) == 0
    hip.run(got, 120.0)
    st = hip.stats()
    assert st.accel_ceased == 1 and st.nstop == 6
    errs = compare_case(got, want, dt_gl=120.0)
    assert max(errs.values()) <= TOL_N, _worst(errs)


def test_nstop_not_divisible(hip, accel):
    accel(1, 2.0)
    case = make_case(2, 32, 28, 4000.0, 20.0)
    assert hip.run(case, 80.0, check=False) == CRM_ERR_ACCEL


def test_accelerated_mixed_precision(accel):
    h = CrmHip(0, "mixed")
    try:
        oracle_lib.oracle_accel_init(1, 2.0, True)
        h.accel_init(1, 2.0, True)
        case = make_case(4, 32, 28, 4000.0, 20.0, randomize=True, seed_base=72)
        want, got = case.copy(), case.copy()
        assert oracle_lib.oracle().run(want, 60.0) == 0
        h.run(got, 60.0)
        errs = compare_case(got, want, groups=("state",), dt_gl=60.0)
        # w of this quiet case is noise-sized (mm/s), so its normwise error is held to an
        # absolute bound instead
        dw = np.abs(got.state["w_wind"][:4] - want.state["w_wind"][:4]).max()
        assert dw < 1e-6, dw
        errs.pop("state.w_wind")
        assert max(errs.values()) < 1e-5, _worst(errs)
    finally:
        h.close()


@pytest.mark.parametrize("uv", [False, True])
def test_3d_level_means_amplified(hip, accel, uv):
    """3D (no C restatement): the identities of the oracle's pinning test, on the GPU."""
    factor = 2.0
    accel(1, factor, uv)
    case = make_case(3, 8, 28, 4000.0, 20.0, ny=8, randomize=True, seed_base=73)
    n, nz, ncol = 3, 28, 64
    size = n * (nz + 1) * ncol
    assert hip.debug_run(case.copy(), 1200.0, 19) == 0
    g = lambda name, m=size: hip.debug_get(name, m)  # noqa: E731
    t19, qt19 = g("t").reshape(n, nz, ncol), g("qt").reshape(n, nz, ncol)
    u19, v19 = g("u").reshape(n, nz, ncol), g("v").reshape(n, nz, ncol)
    qsum19 = (g("qv") + g("qcl") + g("qci")).reshape(n, nz, ncol)
    t0, q0 = g("t0", n * nz).reshape(n, nz), g("q0", n * nz).reshape(n, nz)
    u0, v0 = g("u0", n * nz).reshape(n, nz), g("v0", n * nz).reshape(n, nz)
    assert hip.debug_run(case.copy(), 1200.0, 20) == 0
    t20, qt20 = g("t").reshape(n, nz, ncol), g("qt").reshape(n, nz, ncol)
    u20, v20 = g("u").reshape(n, nz, ncol), g("v").reshape(n, nz, ncol)
    dtm = t19.mean(axis=2) - t0
    dqm = qsum19.mean(axis=2) - q0
    assert np.abs(dtm).max() > 1e-4
    assert np.allclose(t20 - t19, factor * dtm[:, :, None], rtol=0, atol=1e-12)
    ok = ~(qt19 + factor * dqm[:, :, None] < 0).any(axis=2)
    dq = (qt20 - qt19)[ok] - np.broadcast_to(factor * dqm[:, :, None], qt19.shape)[ok]
    assert np.abs(dq).max() <= 1e-14 * np.abs(qt19).max()
    if uv:
        assert np.allclose(u20 - u19, factor * (u19.mean(axis=2) - u0)[:, :, None], rtol=0, atol=1e-13)
        assert np.allclose(v20 - v19, factor * (v19.mean(axis=2) - v0)[:, :, None], rtol=0, atol=1e-13)
    else:
        assert np.array_equal(u20, u19) and np.array_equal(v20, v19)


def test_3d_cease_is_the_plain_call(hip, accel):
    case = convective_case(3, nx=8, ny=8, seed=74)
    case.rad["qrad"][...] = 1.0
    plain = case.copy()
    hip.run(plain, 60.0)
    accel(1, 2.0)
    acc = case.copy()
    hip.run(acc, 60.0)
    assert hip.stats().accel_ceased == 1
    for grp in ("state", "out", "rad"):
        for k, v in getattr(plain, grp).items():
            assert np.array_equal(getattr(acc, grp)[k], v), (grp, k)
