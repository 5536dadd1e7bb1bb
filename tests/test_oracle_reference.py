"""The oracle against the reference phase routines compiled from the reference tree
(oracle/_ref, built by oracle/ref/build_ref.sh when the reference sources are present).
Every internal field after every phase of a substep, and whole calls, must agree bit for bit.
Skipped where the reference library of a size class has not been built."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib
from crm_inputs import make_case
from parity import GRID, PROF, convective_case

INTERNAL = GRID + PROF + ["p", "misc", "mkwle1", "mkwsb1", "mkadv1", "mkdiff1", "qpsrc",
                          "qpevp", "qpfall", "precflux", "uwle", "uwsb"]


def _need(case):
    path = oracle_lib.ref_so_path(case)
    if not os.path.exists(path):
        pytest.skip(f"reference library not built: {os.path.basename(path)}")
    return oracle_lib.reference(case)


def _run_both(case, dt_gl, phase=0):
    ref = _need(case)
    ora = oracle_lib.oracle()
    cr, co = case.copy(), case.copy()
    assert ref.run(cr, dt_gl, phase) == 0
    assert ora.run(co, dt_gl, phase) == 0
    return ref, ora, cr, co


def _assert_cases_equal(cr, co):
    bad = []
    for group in ("state", "rad", "out"):
        for k, v in getattr(cr, group).items():
            if not np.array_equal(v, getattr(co, group)[k]):
                bad.append(f"{group}.{k}")
    assert not bad, bad


@pytest.mark.parametrize("phase", range(1, 21))
def test_every_phase_bitexact(phase):
    """Phases 1..20 of crm_module.F90:751-880 (zero, buoyancy, forcing, qrad, damping,
    ice_fall, boundaries+surface, sgs_proc, boundaries, advect_mom, sgs_mom, pressure, adams,
    boundaries, advect_all_scalars, uvw, boundaries, sgs_scalars, micro_proc, diagnose)."""
    case = convective_case(3, seed=11)
    ref, ora, _, _ = _run_both(case, 2 * case.dt, phase)
    n = case.ncrms * (case.nz + 1) * (case.nx + 8)
    bad = []
    for name in INTERNAL:
        a, b = ref.field(name, n), ora.field(name, n)
        if a.shape != b.shape or not np.array_equal(a, b):
            bad.append(name)
    assert not bad, bad


def test_gcm_step_bitexact():
    case = make_case(2, 32, 28, 4000.0, 20.0, randomize=True)
    _, _, cr, co = _run_both(case, 1200.0)
    _assert_cases_equal(cr, co)


def test_first_step_perturbation_bitexact():
    case = make_case(3, 32, 28, 4000.0, 20.0, randomize=True, seed_base=9)
    case.igstep = 1
    case.gcol = np.array([5, 77, 1234], dtype=np.int32)
    _, _, cr, co = _run_both(case, 200.0)
    _assert_cases_equal(cr, co)


def test_bench_class_bitexact():
    case = convective_case(2, nx=64, nz=58, dx=1000.0, dt=5.0, seed=2)
    _, _, cr, co = _run_both(case, 50.0)
    _assert_cases_equal(cr, co)


def test_128_column_class_bitexact():
    """The widest 2D class (two waves per instance on the GPU)."""
    case = convective_case(2, nx=128, nz=58, dx=1000.0, dt=5.0, seed=3)
    _, _, cr, co = _run_both(case, 50.0)
    _assert_cases_equal(cr, co)


def test_other_level_count_bitexact():
    """A level count other than the SP configurations' 28 and 58 (crm_nz is a build option of
    the reference, crmdims.F90:13-22): 40 levels, 2 km, 10 s, a few substeps and a GCM step."""
    case = convective_case(2, nx=64, nz=40, dx=2000.0, dt=10.0, seed=2)
    _, _, cr, co = _run_both(case, 100.0)
    _assert_cases_equal(cr, co)
    _, _, cr, co = _run_both(case, 1200.0)
    _assert_cases_equal(cr, co)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("isign", [-1, 1])
def test_fft991_bitexact(n, isign):
    """The restated x-transform against the reference's fft991_crm (fft.F90:5-153)."""
    case = make_case(1, 32, 28, 4000.0, 20.0)
    ref = _need(case).lib
    ora = oracle_lib.oracle().lib
    x = np.random.default_rng(n).standard_normal(n + 2)
    a, b = x.copy(), x.copy()
    dp = C.POINTER(C.c_double)
    ref.ref_fft991(a.ctypes.data_as(dp), C.c_int(n), C.c_int(isign))
    ora.oracle_fft991(b.ctypes.data_as(dp), C.c_int(n), C.c_int(isign))
    assert np.array_equal(a, b)
