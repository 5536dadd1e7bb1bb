"""Loaders for the CPU oracles.  TEST INFRASTRUCTURE ONLY.

Two CPU implementations of crm() (crm/crm_module.F90:45-1718) serve as the checker:
  * the C restatement in oracle/crm_oracle.c  -> oracle/_build/libcrm_oracle.so
    (sizes are run-time arguments; one library for every size class);
  * the reference phase routines compiled from the reference tree plus the Fortran
    harness oracle/ref/crm_ref_harness.F90 -> oracle/_ref/libcrmref_<class>.so
    (sizes are compile-time constants in the reference, so one library per class).
Both take the views of include/crm_hip.h.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "acme-ecp_amd"))
from crm_views import CrmDims, CrmColumnInfo  # noqa: E402

ORACLE_SO = os.path.join(HERE, "_build", "libcrm_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def ref_class_name(nx, ny, nz, dx, dt, nx_rad=None, ny_rad=None):
    name = f"{nx}x{ny}x{nz}_dx{int(dx)}_dt{int(dt)}"
    if (nx_rad or nx) != nx or (ny_rad or ny) != ny:
        name += f"_r{nx_rad}x{ny_rad}"
    return name


def ref_so_path(case):
    return os.path.join(REF_DIR, "libcrmref_" + ref_class_name(
        case.nx, case.ny, case.nz, case.dx, case.dt, case.nx_rad, case.ny_rad) + ".so")


class _Runner:
    def __init__(self, path, entry, getter):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.fn = getattr(self.lib, entry)
        self.fn.restype = C.c_int
        self.get = getattr(self.lib, getter)
        self.get.restype = C.c_int
        self.get.argtypes = [C.c_char_p, C.POINTER(C.c_double)]

    def run(self, case, dt_gl, phase_stop=0):
        """Run one crm() call in place on case's arrays; returns the status code."""
        dims = case.dims()
        col = case.column_info()
        vi, vs, vr, vo = case.views()
        return self.fn(C.byref(dims), C.byref(col), C.c_double(dt_gl), C.byref(vi),
                       C.byref(vs), C.byref(vr), C.byref(vo), C.c_int(phase_stop))

    def field(self, name, size):
        buf = np.zeros(size)
        n = self.get(name.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)))
        if n < 0:
            raise KeyError(name)
        return buf[:n]


def reference(case):
    """Reference CRM (compiled reference phases + harness) for case's size class."""
    return _Runner(ref_so_path(case), "ref_crm", "ref_get")


def reference_set_coriolis(case, on):
    """The reference's docoriolis switch (params.F90:107) in the build for case's size class
    (process-wide, like the module variable it sets)."""
    lib = C.CDLL(ref_so_path(case))
    lib.ref_set_coriolis.argtypes = [C.c_int]
    lib.ref_set_coriolis(int(bool(on)))


def oracle():
    """C restatement of the CRM."""
    return _Runner(ORACLE_SO, "oracle_crm", "oracle_get")


def oracle_accel_init(use_crm_accel, crm_accel_factor=0.0, crm_accel_uv=False):
    """Mean-state acceleration settings of the C restatement (process-wide, like the
    reference's module variables set by crm_accel_init, accelerate_crm.F90:30-38,45-75)."""
    lib = C.CDLL(ORACLE_SO)
    lib.oracle_accel_init.argtypes = [C.c_int, C.c_double, C.c_int]
    lib.oracle_accel_init(int(bool(use_crm_accel)), float(crm_accel_factor), int(bool(crm_accel_uv)))


class StatsProbe:
    """The oracle's per-substep statistics and end-of-call outputs (crm_module.F90:896-1659)
    run on a constructed state: set_up() runs the set-up of a call on `case` (reference
    profiles from its crm_input), set() overwrites internal fields, run() evaluates the
    statistics into case.rad / case.out.  mutate(m) plants one of the oracle's numbered
    mistakes (0 removes it), so that a test can show its checks notice it."""

    def __init__(self):
        self.lib = C.CDLL(ORACLE_SO)
        self.lib.oracle_set.restype = C.c_int
        self.lib.oracle_set.argtypes = [C.c_char_p, C.POINTER(C.c_double)]
        self.lib.oracle_stats_probe.restype = C.c_int
        self.lib.oracle_mutate.argtypes = [C.c_int]
        self.runner = oracle()

    def set_up(self, case):
        work = case.copy()
        rc = self.runner.run(work, case.dt, phase_stop=20)
        if rc != 0:
            raise RuntimeError(f"oracle set-up failed: {rc}")

    def get(self, name, size):
        return self.runner.field(name, size)

    def set(self, name, values):
        buf = np.ascontiguousarray(values, dtype=np.float64).ravel()
        n = self.lib.oracle_set(name.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)))
        if n != buf.size:
            raise ValueError(f"oracle_set({name}): {n} values taken, {buf.size} given")

    def run(self, case, dt_gl, dtn, nsteps):
        dims = case.dims()
        vi, _, vr, vo = case.views()
        return self.lib.oracle_stats_probe(C.byref(dims), C.c_double(dt_gl), C.c_double(dtn),
                                           C.c_int(nsteps), C.byref(vi), C.byref(vr), C.byref(vo))

    def mutate(self, m):
        self.lib.oracle_mutate(int(m))


# Planted mistakes of oracle_mutate (the enum in crm_oracle.c).
MUTATIONS = {"eu_du_swapped": 1, "cloud_layers_700_400_swapped": 2, "cld_not_time_mean": 3,
             "mc_threshold_wmin": 4, "jt_mx_swapped": 5, "water_path_pdel_level": 6,
             "cldtop_every_level": 7, "tkez_no_half": 8, "scalar_flux_dt_not_dtn": 9,
             "updraft_cloud_level_above": 10, "rad_qv_unclipped": 11, "prec_split_per_hour": 12}
