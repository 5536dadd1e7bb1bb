"""Behaviour of the C ABI on an MI355X: reproducibility, the device-resident entry point,
caller leading dimensions, status codes (the reference's aborts), context reuse."""
import numpy as np
import pytest

import oracle_lib
from crm_hip import (CRM_ERR_ARG, CRM_ERR_CFL, CRM_ERR_DAMPING, CRM_ERR_NAN, CRM_ERR_NXRAD,
                     CRM_ERR_UNSUPPORTED, CrmError, CrmHip, DeviceCase)
from crm_inputs import make_case
from parity import compare_case, convective_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    h = CrmHip(0)
    yield h
    h.close()


def _equal(a, b):
    return all(np.array_equal(getattr(a, g)[k], getattr(b, g)[k])
               for g in ("state", "rad", "out") for k in getattr(a, g))


def test_bitwise_reproducible(hip):
    """Fixed-order wave reductions: two runs give identical bits."""
    case = convective_case(7, seed=2)
    a, b = case.copy(), case.copy()
    hip.run(a, 200.0)
    hip.run(b, 200.0)
    assert _equal(a, b)


def test_device_entry_equals_host_entry(hip):
    case = convective_case(6, nx=64, nz=58, dx=1000.0, dt=5.0, seed=3)
    a = case.copy()
    hip.run(a, 50.0)
    dc = DeviceCase(case.copy(), "cuda:0")
    hip.run_device(dc, 50.0)
    b = dc.download()
    assert _equal(a, b)


def test_leading_dimension(hip):
    """Arrays dimensioned (pcols, ...) with ncrms < pcols: only the first ncrms rows are read
    or written (crm_module.F90:268-281)."""
    base = convective_case(3, seed=4)
    padded = convective_case(3, seed=4, ld=8)
    for g in ("state", "rad", "out"):
        for k, v in getattr(padded, g).items():
            v[3:] = 12345.0
    hip.run(base, 60.0)
    hip.run(padded, 60.0)
    for g in ("state", "rad", "out"):
        for k, v in getattr(padded, g).items():
            assert np.array_equal(v[:3], getattr(base, g)[k][:3]), (g, k)
            assert np.all(v[3:] == 12345.0), (g, k)


def test_context_grows_and_shrinks(hip):
    """One context serves calls of different batch sizes and size classes in turn."""
    for n, nx, nz, dx, dt in [(2, 32, 28, 4000.0, 20.0), (70, 32, 28, 4000.0, 20.0),
                              (3, 64, 58, 1000.0, 5.0), (2, 128, 58, 1000.0, 5.0),
                              (2, 32, 28, 4000.0, 20.0)]:
        case = convective_case(n, nx=nx, nz=nz, dx=dx, dt=dt, seed=n)
        want, got = case.copy(), case.copy()
        assert oracle_lib.oracle().run(want, dt) == 0
        hip.run(got, dt)
        assert max(compare_case(got, want, dt_gl=dt).values()) <= 1e-10


def test_cfl_violation_is_reported(hip):
    """kurant.F90:64-75 aborts when more than 16 subcycles are needed."""
    case = convective_case(2, seed=5)
    case.state["w_wind"][1, 3, 0, 2] = 120.0
    assert oracle_lib.oracle().run(case.copy(), 20.0) == CRM_ERR_CFL
    with pytest.raises(CrmError) as e:
        hip.run(case.copy(), 20.0)
    assert e.value.code == CRM_ERR_CFL


@pytest.mark.parametrize("kw,code", [
    (dict(nx_rad=3), CRM_ERR_NXRAD),          # crm_module.F90:686-692
    (dict(dt=40.0), CRM_ERR_DAMPING),         # damping.F90:39-42
    (dict(nx=48), CRM_ERR_UNSUPPORTED),       # column count without a compiled kernel
    (dict(nx=256), CRM_ERR_UNSUPPORTED),
    (dict(nz=72), CRM_ERR_UNSUPPORTED),       # more levels than lanes of a wave
    (dict(nx=16, ny=16), CRM_ERR_UNSUPPORTED),  # 3D is compiled for 8 x 8 and 4 x 4 columns
    (dict(nx=8, ny=8, nx_rad=3, ny_rad=8), CRM_ERR_NXRAD),
])
def test_status_codes(hip, kw, code):
    args = dict(ncrms=2, nx=32, nz=28, dx=4000.0, dt=20.0)
    args.update(kw)
    case = make_case(args.pop("ncrms"), args.pop("nx"), args.pop("nz"), args.pop("dx"),
                     args.pop("dt"), **args)
    assert hip.run(case, 1200.0, check=False) == code


def test_non_finite_state_is_reported(hip):
    """A NaN in the state propagates silently through the reference; here every loop with a
    data-dependent trip count is bounded and the write-back flags non-finite values."""
    case = convective_case(2, seed=7)
    case.state["temperature"][1, 5, 0, 10] = np.nan
    assert hip.run(case, 60.0, check=False) == CRM_ERR_NAN


def test_bad_time_step_rejected(hip):
    case = make_case(2, 32, 28, 4000.0, 20.0)
    assert hip.run(case, 10.0, check=False) == CRM_ERR_ARG  # dt_gl < crm_dt: no substep


def test_bad_dims_rejected(hip):
    case = make_case(2, 32, 28, 4000.0, 20.0)
    case.ld = 1  # ld < ncrms
    assert hip.run(case, 20.0, check=False) == CRM_ERR_ARG


def test_context_usable_after_errors(hip):
    case = convective_case(2, seed=6)
    want, got = case.copy(), case.copy()
    assert oracle_lib.oracle().run(want, 20.0) == 0
    hip.run(got, 20.0)
    assert max(compare_case(got, want, dt_gl=20.0).values()) <= 1e-10


@pytest.mark.parametrize("graphs", [0, 2])
def test_stats_report_step_counts(hip, graphs):
    """crm_hip_stats after a call: nstop and the largest sedimentation sub-step count nprec of
    the call, as the oracle counts it (microphysics.F90:602-662), on the captured and the
    plain-launch path."""
    case = convective_case(6, seed=3)
    o = oracle_lib.oracle()
    assert o.run(case.copy(), 200.0) == 0
    want = int(o.field("nprec_max", 1)[0])
    hip.set_graphs(graphs)
    try:
        hip.run(case.copy(), 200.0)
    finally:
        hip.set_graphs(2)
    st = hip.stats()
    assert st.nstop == 10
    assert st.ncycle_max >= 1
    assert want > 1 and st.nprec_max == want
