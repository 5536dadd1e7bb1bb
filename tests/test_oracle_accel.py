"""Mean-state acceleration in the oracle (accelerate_crm.F90:109-366; crm_module.F90:718-720,
872-876).  accelerate_crm.F90 cannot be compiled here (it needs CAM's phys_control,
cam_logfile, spmd_utils and cam_abortutils), so the C restatement is pinned by what the
scheme fixes exactly:
  * after the call, each level mean of t equals its value after micro_proc plus
    crm_accel_factor times its change since the previous diagnose (phase 19 -> 20);
  * level totals of qt move by nx * factor * (its mean change) where nothing clips, and qv,
    qcl, qci re-add to qt;
  * a 5 K mean change stops the acceleration before it acts: the call is then the plain,
    unaccelerated call, bit for bit, with nstop restored;
  * factor 0 is the plain call up to the round-off of the qv re-split;
  * nstop not divisible by 1 + factor is the reference's abort.
"""
import numpy as np
import pytest

import oracle_lib
from crm_inputs import make_case
from parity import convective_case, relerr

CRM_ERR_ACCEL = 9


@pytest.fixture
def accel():
    def set_(use, factor=0.0, uv=False):
        oracle_lib.oracle_accel_init(use, factor, uv)
    yield set_
    oracle_lib.oracle_accel_init(0, 0.0, False)


def _run(case, dt_gl, phase=0):
    o = oracle_lib.oracle()
    c = case.copy()
    rc = o.run(c, dt_gl, phase)
    return rc, c, o


@pytest.mark.parametrize("uv", [False, True])
def test_level_means_amplified(accel, uv):
    factor = 2.0
    accel(1, factor, uv)
    # a quiet case: the convective test plume changes level means by more than 5 K in its
    # first substep, which stops the acceleration at once (see the next test)
    case = make_case(3, 32, 28, 4000.0, 20.0, randomize=True, seed_base=61)
    n, nx, nz = case.ncrms, case.nx, case.nz
    size = n * (nz + 1) * nx
    # a GCM step of 20 accelerated substeps, stopped in the last one after micro_proc (19)
    # and after the acceleration and diagnose (20)
    rc, _, o19 = _run(case, 1200.0, 19)
    assert rc == 0
    t19 = o19.field("t", size).reshape(n, nz, nx)
    qt19 = o19.field("qt", size).reshape(n, nz, nx)
    u19 = o19.field("u", size).reshape(n, nz, nx)
    t0 = o19.field("t0", n * nz).reshape(n, nz)   # from the diagnose before this substep
    q0 = o19.field("q0", n * nz).reshape(n, nz)
    u0 = o19.field("u0", n * nz).reshape(n, nz)
    qv19 = o19.field("qv", size).reshape(n, nz, nx)
    qcl19 = o19.field("qcl", size).reshape(n, nz, nx)
    qci19 = o19.field("qci", size).reshape(n, nz, nx)
    rc, _, o20 = _run(case, 1200.0, 20)
    assert rc == 0
    t20 = o20.field("t", size).reshape(n, nz, nx)
    qt20 = o20.field("qt", size).reshape(n, nz, nx)
    u20 = o20.field("u", size).reshape(n, nz, nx)
    dt_mean = t19.mean(axis=2) - t0
    dq_mean = (qv19 + qcl19 + qci19).mean(axis=2) - q0
    # t: every column shifted by factor * (mean change)
    assert np.allclose(t20 - t19, factor * dt_mean[:, :, None], rtol=0, atol=1e-12)
    assert np.abs(dt_mean).max() > 1e-4  # the test exercises a real mean change
    # qt: where no level clips, the shift is uniform too
    clip = (qt19 + factor * dq_mean[:, :, None] < 0).any(axis=2)
    ok = ~clip
    assert ok.sum() > 0
    dqt = (qt20 - qt19)[ok] - np.broadcast_to(factor * dq_mean[:, :, None], qt19.shape)[ok]
    assert np.abs(dqt).max() <= 1e-14 * np.abs(qt19).max(), np.abs(dqt).max()
    # u only with crm_accel_uv
    du = u20 - u19
    if uv:
        assert np.allclose(du, factor * (u19.mean(axis=2) - u0)[:, :, None], rtol=0, atol=1e-13)
    else:
        assert np.array_equal(u20, u19)
    # qv + qcl + qci == qt after the re-split (qv = qt - qcl - qci where that is >= 0)
    qv20 = o20.field("qv", size).reshape(n, nz, nx)
    qcl20 = o20.field("qcl", size).reshape(n, nz, nx)
    qci20 = o20.field("qci", size).reshape(n, nz, nx)
    assert np.allclose(qv20 + qcl20 + qci20, qt20, rtol=1e-14, atol=1e-20)


def test_cease_restores_the_plain_call(accel):
    """A mean heating of 20 K per substep trips the 5 K test at the first substep: nothing is
    applied, nstop goes back to dt_gl / dt, and the call equals the unaccelerated one."""
    case = convective_case(3, seed=62)
    case.rad["qrad"][...] = 1.0  # K/s
    rc, plain, _ = _run(case, 3 * case.dt)
    assert rc == 0
    accel(1, 2.0)
    rc, acc, _ = _run(case, 3 * case.dt)
    assert rc == 0
    for g in ("state", "out", "rad"):
        for k, v in getattr(plain, g).items():
            assert np.array_equal(getattr(acc, g)[k], v), (g, k)


def test_factor_zero_is_the_plain_call(accel):
    case = convective_case(3, seed=63)
    rc, plain, _ = _run(case, 10 * case.dt)
    assert rc == 0
    accel(1, 0.0)
    rc, acc, _ = _run(case, 10 * case.dt)
    assert rc == 0
    for k, v in plain.state.items():
        assert relerr(acc.state[k][:3], v[:3]) < 1e-12, k


def test_nstop_not_divisible(accel):
    accel(1, 2.0)
    case = convective_case(2, seed=64)
    rc, _, _ = _run(case, 4 * case.dt)  # nstop 4, 1 + factor = 3
    assert rc == CRM_ERR_ACCEL


def test_accelerated_call_keeps_budgets(accel):
    """Over an accelerated GCM step the outputs stay finite and the call integrates
    nstop / (1 + factor) substeps; the tendencies see the whole dt_gl."""
    accel(1, 2.0)
    case = convective_case(3, seed=65)
    rc, c, _ = _run(case, 1200.0)
    assert rc == 0
    for k, v in c.out.items():
        assert np.all(np.isfinite(v)), k
