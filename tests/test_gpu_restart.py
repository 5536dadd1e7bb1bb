"""Exact restart.  The CRM state that E3SM carries between GCM steps is the caller's own arrays
(crm_state and crm_rad, physics-buffer fields of cam/crm_physics.F90:126-164 that a restart
file holds); the library keeps nothing between calls that changes a result.  Two GCM steps on
one context must therefore equal, bit for bit, a first step whose state is written to a file,
read back and advanced by a fresh context (an ERS-style test), for the FP64 and the mixed
engine, a 2D and the 3D shape, with the first-step perturbation (igstep 1) on the first call
only."""
import numpy as np
import pytest

from crm_hip import CrmHip
from parity import convective_case

pytestmark = pytest.mark.gpu

SHAPES = [dict(ncrms=5, nx=32, nz=28, dx=4000.0, dt=20.0),
          dict(ncrms=3, nx=64, nz=58, dx=1000.0, dt=5.0),
          dict(ncrms=2, nx=8, ny=8, nz=28, dx=4000.0, dt=20.0)]


def _save(case, path):
    arrays = {f"state.{k}": v for k, v in case.state.items()}
    arrays.update({f"rad.{k}": v for k, v in case.rad.items()})
    np.savez(path, **arrays)


def _load(case, path):
    with np.load(path, allow_pickle=False) as z:
        for k in z.files:
            group, name = k.split(".", 1)
            getattr(case, group)[name][...] = z[k]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s['nx']}x{s.get('ny', 1)}x{s['nz']}")
def test_restart_is_bitwise(tmp_path, shape, prec):
    shape = dict(shape)
    n = shape.pop("ncrms")
    dt_gl = 60 * shape["dt"]
    case = convective_case(n, seed=21, **shape)
    case.igstep = 1
    # one run of two GCM steps
    hip = CrmHip(0, prec)
    try:
        a = case.copy()
        hip.run(a, dt_gl)
        a.igstep = 2
        hip.run(a, dt_gl)
    finally:
        hip.close()
    # first step, restart file, second step on a fresh context
    path = tmp_path / "crm_restart.npz"
    hip = CrmHip(0, prec)
    try:
        b = case.copy()
        hip.run(b, dt_gl)
        _save(b, path)
    finally:
        hip.close()
    c = case.copy()
    for grp in ("state", "rad"):
        for v in getattr(c, grp).values():
            v[...] = np.nan  # nothing survives but the file
    _load(c, path)
    # output buffers are not restart state; start them as the one-run case has them
    for k, v in b.out.items():
        c.out[k][...] = v
    c.igstep = 2
    hip = CrmHip(0, prec)
    try:
        hip.run(c, dt_gl)
    finally:
        hip.close()
    for grp in ("state", "rad", "out"):
        for k, v in getattr(a, grp).items():
            assert np.array_equal(getattr(c, grp)[k], v), (grp, k)
