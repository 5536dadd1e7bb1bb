"""Write the golden fixtures of the CRM from the reference phase routines.

The reference ships no fixtures or known-answer tests for crm().  These are produced by the
reference CRM itself: the reference Fortran sources compiled by oracle/ref/build_ref.sh
(oracle/_ref/libcrmref_<class>.so) and driven through oracle/ref/crm_ref_harness.F90.  Each
fixture holds the full input case (every crm_input / crm_state / crm_rad / crm_output array)
and every state, rad and output array after the call, so the oracle can be checked against
the reference where the reference cannot be built (the GPU box has no reference tree).

Run from the repository root after building the reference classes:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "acme-ecp_amd"), os.path.join(ROOT, "oracle"),
                os.path.join(ROOT, "tests")]

import oracle_lib  # noqa: E402
from crm_inputs import make_case  # noqa: E402
from parity import convective_case  # noqa: E402

# name -> (case builder, dt_gl)
CASES = {
    # one 2D CRM, 32x28, dx 4 km, dt 20 s, tropical sounding, full 1200 s GCM step
    "single_32x28_gcmstep": (lambda: make_case(1, 32, 28, 4000.0, 20.0), 1200.0),
    # three convective instances, one substep
    "convective_32x28_1sub": (lambda: convective_case(3, seed=3), 20.0),
    # first GCM step (initial noise from the Mersenne twister), 5 substeps
    "perturb_32x28_5sub": (lambda: _first_step(make_case(4, 32, 28, 4000.0, 20.0,
                                                         randomize=True)), 100.0),
    # two instances of the 64x58 benchmark class, dx 1 km, dt 5 s, 4 substeps
    "convective_64x58_4sub": (lambda: convective_case(2, nx=64, nz=58, dx=1000.0, dt=5.0,
                                                      seed=5), 20.0),
}


def _first_step(c):
    c.igstep = 1
    c.gcol = np.array([17, 256, 3001, 42], dtype=np.int32)[: c.ncrms]
    return c


def case_to_arrays(case, prefix):
    out = {}
    for group in ("inp", "state", "rad", "out"):
        for k, v in getattr(case, group).items():
            out[f"{prefix}{group}.{k}"] = v
    return out


def save(name):
    build, dt_gl = CASES[name]
    case = build()
    ref = oracle_lib.reference(case)
    after = case.copy()
    rc = ref.run(after, dt_gl)
    assert rc == 0, (name, rc)
    meta = dict(ncrms=case.ncrms, plev=case.plev, nx=case.nx, ny=case.ny, nz=case.nz,
                dx=case.dx, dt=case.dt, nx_rad=case.nx_rad, ny_rad=case.ny_rad, ld=case.ld,
                igstep=case.igstep, dt_gl=dt_gl)
    arrays = {"meta." + k: np.array(v) for k, v in meta.items()}
    arrays["gcol"] = case.gcol
    arrays.update(case_to_arrays(case, "in:"))
    for group in ("state", "rad", "out"):
        for k, v in getattr(after, group).items():
            arrays[f"ref:{group}.{k}"] = v
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print(path, os.path.getsize(path))


if __name__ == "__main__":
    for n in (sys.argv[1:] or CASES):
        save(n)
