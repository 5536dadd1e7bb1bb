"""Time-mean history fields accumulated on the device (include/crm_history.h) against the
restatement of crm_physics.F90:1201-1461 (oracle/history_oracle.py), bitwise: three GCM steps of
a chunk of convecting CRMs with fewer columns than pcols, together with physics arrays."""
import numpy as np
import pytest

import history_oracle
from crm_hip import CrmHip, DeviceCase
from crm_history import CrmHistory, field_names
from parity import convective_case

pytestmark = pytest.mark.gpu


def _column_major(a, ncol):
    """(ncol, per-column size) of a Fortran-layout array's first ncol columns."""
    a = np.asarray(a)
    return np.stack([a[i].reshape(-1, order="F") for i in range(ncol)]) if a.ndim > 1 else a[:ncol, None]


@pytest.mark.parametrize("shape", [(32, 1, 28, 4000.0, 20.0, 8), (8, 8, 28, 4000.0, 20.0, 8)],
                         ids=["32x28_rad8", "8x8x28"])
def test_history_means_match_restatement(shape):
    import torch
    nx, ny, nz, dx, dt, nxr = shape
    ncol, pcols = 5, 8
    case = convective_case(ncol, nx=nx, ny=ny, nz=nz, dx=dx, dt=dt, seed=31, ld=pcols, nx_rad=nxr,
                           ny_rad=(nxr if ny > 1 else 1))
    plev = case.plev
    dc = DeviceCase(case, "cuda:0")
    hip = CrmHip(0)
    hist = CrmHistory(0, pcols, plev, nx, ny, nz, case.nx_rad, case.ny_rad)
    ref = history_oracle.History()
    rng = np.random.default_rng(5)
    try:
        for step in range(3):
            print(f"[history {nx}x{ny}x{nz}] GCM step {step}", flush=True)
            hip.run_device(dc, 60.0)
            phys = {k: np.asfortranarray(rng.normal(1.0, 0.5, (pcols, plev)))
                    for k in ("pmid", "pdel", "s", "qv", "qcl", "qci", "qrs", "qrl")}
            dev = {k: torch.from_numpy(np.ascontiguousarray(v.T)).cuda() for k, v in phys.items()}
            ptend = {k: dev[k] for k in ("s", "qv", "qcl", "qci")}
            hist.accumulate(hip, ncol, dcase=dc, qrad=dc.arrays["rad"]["qrad"], pmid=dev["pmid"],
                            pdel=dev["pdel"], ptend=ptend, qrs=dev["qrs"], qrl=dev["qrl"])
            host = dc.download(case.copy())
            ref.accumulate(ncol, state=host.state, out=host.out, qrad=host.rad["qrad"],
                           pmid=phys["pmid"], pdel=phys["pdel"],
                           ptend={k: phys[k] for k in ("s", "qv", "qcl", "qci")}, qrs=phys["qrs"],
                           qrl=phys["qrl"])
        # a pass with only the heating rate advances only CRM_QRAD
        hist.accumulate(hip, ncol, qrad=dc.arrays["rad"]["qrad"])
        ref.accumulate(ncol, qrad=dc.download(case.copy()).rad["qrad"])
        names = field_names()
        assert len(names) == 65 and set(names) == set(ref.sums)
        for name in names:
            got, n = hist.mean(hip, name)
            got = got.cpu().numpy()  # (per-column size, pcols)
            want, nw = ref.mean(name)
            assert n == nw == (4 if name == "CRM_QRAD" else 3), name
            w = _column_major(want, ncol)
            assert np.array_equal(got[:, :ncol].T, w), (name, np.abs(got[:, :ncol].T - w).max())
            assert np.all(got[:, ncol:] == 0.0), name  # columns beyond the chunk untouched
        # a new interval starts from zero
        hist.reset(hip)
        got, n = hist.mean(hip, "CRM_T")
        assert n == 0 and float(got.abs().max()) == 0.0
        # the CRM state is what the history saw: a 3D CRM's v is averaged too
        if ny > 1:
            assert float(dc.arrays["state"]["v_wind"].abs().max()) > 0
    finally:
        hist.close()
        hip.close()
