"""Total-water and static-energy budgets of the HIP CRM close to FP64 round-off over a GCM
step (same budgets as for the oracle; the prognostic static energy at the end of the call is
read through the library's inspection entry)."""
import pytest

from crm_hip import CrmHip
from crm_inputs import make_case
from parity import budget_residuals, convective_case

pytestmark = pytest.mark.gpu

CASES = {
    "convective_32x28": (lambda: convective_case(3, seed=2), 1200.0),
    "convective_64x58": (lambda: convective_case(2, nx=64, nz=58, dx=1000.0, dt=5.0, seed=3), 200.0),
    "convective_128x58": (lambda: convective_case(2, nx=128, nz=58, dx=1000.0, dt=5.0, seed=4), 200.0),
    "quiet_32x28": (lambda: make_case(3, 32, 28, 4000.0, 20.0, randomize=True), 1200.0),
    "batch_866x32x28": (lambda: make_case(866, 32, 28, 4000.0, 20.0, randomize=True), 1200.0),
}


@pytest.fixture(scope="module")
def hip():
    h = CrmHip(0)
    yield h
    h.close()


@pytest.mark.parametrize("name", CASES)
def test_budgets_close(hip, name):
    build, dt_gl = CASES[name]
    before = build()
    after = before.copy()
    assert hip.debug_run(after, dt_gl, 0) == 0
    rw, rh = budget_residuals(before, after, dt_gl, hip.debug_get)
    print(name, "water %.2e energy %.2e" % (rw, rh))
    assert rw < 1e-14, rw
    assert rh < 1e-13, rh
