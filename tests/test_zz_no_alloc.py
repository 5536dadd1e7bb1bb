"""The entry point that allocates nothing (crm_hip_run_device_noalloc): every device buffer of a
call comes from the library's fixed per-GPU array.  Its results must be bitwise those of
crm_hip_run_device on the same inputs, call after call, including a call that follows a larger
one (the array keeps the previous call's contents), in both precisions; a batch beyond the
array is rejected before anything runs."""
import numpy as np
import pytest

import crm_mlp
from crm_hip import (CRM_ERR_SCRATCH, CRM_ERR_UNSUPPORTED, NOALLOC_SCRATCH_BYTES, CrmHip,
                     DeviceCase)
from crm_inputs import make_case
from parity import convective_case

pytestmark = pytest.mark.gpu

# (ncrms, nx, ny, nz, dx, dt, nx_rad, dt_gl, igstep): each call smaller than the one before
CALLS = [
    (12, 64, 1, 58, 1000.0, 5.0, 16, 40.0, 2),
    (5, 64, 1, 58, 1000.0, 5.0, 64, 25.0, 1),   # first GCM step: the seeded perturbation
    (7, 32, 1, 28, 4000.0, 20.0, 32, 200.0, 2),
    (3, 8, 8, 28, 4000.0, 20.0, 8, 60.0, 2),     # 3D
]


def _same(a, b):
    for g in ("state", "rad", "out"):
        for k, t in a.arrays[g].items():
            u = b.arrays[g][k]
            assert t.shape == u.shape and bool((t == u).all()), (g, k)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_noalloc_equals_run_device(prec):
    ref, na = CrmHip(0, prec), CrmHip(0, prec)
    try:
        for j, (n, nx, ny, nz, dx, dt, nxr, gl, ig) in enumerate(CALLS):
            print(f"[noalloc {prec}] call {j}: {n} x {nx}x{ny}x{nz}, nx_rad {nxr}, {gl:g} s", flush=True)
            case = convective_case(n, nx=nx, ny=ny, nz=nz, dx=dx, dt=dt, seed=60 + j, nx_rad=nxr,
                                   ny_rad=(ny if ny > 1 else 1))
            case.igstep = ig
            case.rad["qrad"][...] = np.random.default_rng(j).uniform(-2e-5, 2e-5, case.rad["qrad"].shape)
            a, b = DeviceCase(case.copy(), "cuda:0"), DeviceCase(case.copy(), "cuda:0")
            ref.run_device(a, gl)
            na.run_device_noalloc(b, gl)
            _same(a, b)
            sa, sb = ref.stats(), na.stats()
            assert (sa.nstop, sa.ncycle_max, sa.nprec_max) == (sb.nstop, sb.ncycle_max, sb.nprec_max)
        assert na.graph_launches() == 0  # plain launches only
    finally:
        ref.close()
        na.close()


def test_noalloc_rejects_batch_beyond_its_array():
    """At about 0.85 MB per FP64 64 x 58 instance, 1000 instances need more than the array's
    256 MiB: CRM_ERR_SCRATCH, returned before the views are touched."""
    case = make_case(2, 64, 58, 1000.0, 5.0)
    dc = DeviceCase(case, "cuda:0")
    dc.case = case.copy()
    dc.case.ncrms = dc.case.ld = 1000
    dc.case.gcol = np.arange(1, 1001, dtype=np.int32)
    dc.case.lat = np.zeros(1000)
    dc.case.lon = np.zeros(1000)
    assert NOALLOC_SCRATCH_BYTES == 256 << 20
    with CrmHip(0) as h:
        assert h.run_device_noalloc(dc, 5.0, check=False) == CRM_ERR_SCRATCH


def test_noalloc_rejects_network():
    case = make_case(2, 32, 28, 4000.0, 20.0)
    dc = DeviceCase(case, "cuda:0")
    h, m = CrmHip(0), crm_mlp.CrmMlp(0)
    try:
        m.set_params(crm_mlp.init_params(1))
        m.set_norm(np.zeros(9), np.ones(9), np.ones(6))
        crm_mlp.attach(h, m)
        assert h.run_device_noalloc(dc, 20.0, check=False) == CRM_ERR_UNSUPPORTED
        crm_mlp.attach(h, None)
        assert h.run_device_noalloc(dc, 20.0, check=False) == 0
    finally:
        h.close()
        m.close()
