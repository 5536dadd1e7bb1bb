"""The GCM-side kernels of the library (include/crm_glue.h) against the CPU restatement
(oracle/glue_oracle.py, pinned by tests/test_oracle_glue.py): bitwise, except the rotated input
wind, whose cos/sin come from the device's and the host's math libraries (within 2 ulp of the
wind speed).  Columns past ncol must stay untouched.  A last test runs the whole step on the
device: physics state -> crm_input -> crm() -> tendencies, against the same chain on the host
with the oracle CRM."""
import numpy as np
import pytest

import crm_glue
import glue_cases
import glue_oracle as G
import oracle_lib
from crm_hip import CrmHip, DeviceCase
from crm_inputs import make_case
from parity import compare_case

pytestmark = pytest.mark.gpu

NCOL, PCOLS, CRM_NZ = 5, 7, 28
PVER = CRM_NZ + 2


@pytest.fixture(scope="module")
def hip():
    h = CrmHip(0)
    yield h
    h.close()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).cuda()


def _host(t):
    return np.asfortranarray(t.cpu().numpy().T)


def _case():
    case = make_case(NCOL, 32, CRM_NZ, 4000.0, 20.0, randomize=True, ld=PCOLS, nx_rad=8)
    case.rad["qrad"][...] = np.random.default_rng(7).normal(0, 1e-2, case.rad["qrad"].shape)
    return case


def test_gcm_input(hip):
    st = glue_cases.physics_state(NCOL, PCOLS, PVER, seed=11)
    srf = glue_cases.surface(PCOLS)
    angle = np.random.default_rng(3).uniform(0, 2 * np.pi, PCOLS)
    case = _case()
    want = case.copy()
    G.gcm_input(NCOL, CRM_NZ, st, srf, angle, want.inp, want.rad["qrad"])
    dc = DeviceCase(case, "cuda:0")
    crm_glue.gcm_input(hip, NCOL, PCOLS, PVER, CRM_NZ, 8, {k: _dev(v) for k, v in st.items()},
                       {k: _dev(v) for k, v in srf.items()}, _dev(angle), dc)
    got_inp = {k: _host(t) for k, t in dc.arrays["inp"].items()}
    for k, v in want.inp.items():
        if k in ("ul", "vl"):
            speed = np.hypot(st["u"], st["v"])
            assert np.all(np.abs(got_inp[k] - v) <= 4e-16 * speed), k
        else:
            assert np.array_equal(got_inp[k], v), k
    assert np.array_equal(_host(dc.arrays["rad"]["qrad"]), want.rad["qrad"])


def test_flux_bypass(hip):
    st = glue_cases.physics_state(NCOL, PCOLS, PVER, seed=12)
    srf = glue_cases.surface(PCOLS)
    ps_, pq = np.full((PCOLS, PVER), 7.0, order="F"), np.full((PCOLS, PVER), 7.0, order="F")
    G.flux_bypass(NCOL, st["pdel"], srf["shf"], srf["cflx"], ps_, pq)
    gs, gq = _dev(np.full((PCOLS, PVER), 7.0)), _dev(np.full((PCOLS, PVER), 7.0))
    crm_glue.flux_bypass(hip, NCOL, PCOLS, PVER, _dev(st["pdel"]), _dev(srf["shf"]), _dev(srf["cflx"]), gs, gq)
    assert np.array_equal(_host(gs), ps_) and np.array_equal(_host(gq), pq)


def test_gcm_tend(hip):
    st = glue_cases.physics_state(NCOL, PCOLS, PVER, seed=13)
    rng = np.random.default_rng(4)
    case = _case()
    for k in ("qltend", "qcltend", "qiltend", "sltend"):
        case.out[k][...] = rng.normal(0, 1e-4, case.out[k].shape)
    for k in ("precc", "precl", "precsc", "precsl"):
        case.out[k][...] = rng.uniform(0, 1e-7, case.out[k].shape)
    qrs = np.asfortranarray(rng.normal(0, 1e-3, (PCOLS, PVER)))
    qrl = np.asfortranarray(rng.normal(0, 1e-3, (PCOLS, PVER)))
    want = case.copy()
    wqrs, wqrl = qrs.copy(order="F"), qrl.copy(order="F")
    wpt = {k: np.full((PCOLS, PVER), 3.0, order="F") for k in ("s", "qv", "qcl", "qci")}
    f = G.gcm_tend(NCOL, CRM_NZ, st["pdel"], want.out, want.rad["qrad"], wqrs, wqrl, wpt)
    dc = DeviceCase(case, "cuda:0")
    gqrs, gqrl = _dev(qrs), _dev(qrl)
    gpt = {k: _dev(np.full((PCOLS, PVER), 3.0)) for k in ("s", "qv", "qcl", "qci")}
    gf = _dev(np.zeros(PCOLS))
    crm_glue.gcm_tend(hip, NCOL, PCOLS, PVER, CRM_NZ, 8, _dev(st["pdel"]), dc, gqrs, gqrl, gpt, gf)
    got = dc.download(case.copy())
    assert np.array_equal(_host(gf)[:NCOL], f)
    assert np.array_equal(_host(gqrs), wqrs) and np.array_equal(_host(gqrl), wqrl)
    for k in gpt:
        assert np.array_equal(_host(gpt[k]), wpt[k]), k
    for k in ("precc", "precl", "precsc", "precsl"):
        assert np.array_equal(got.out[k], want.out[k]), k
    assert np.array_equal(got.rad["qrad"], want.rad["qrad"])


def test_check_energy_chng(hip):
    st = glue_cases.physics_state(NCOL, PCOLS, PVER, seed=14)
    rng = np.random.default_rng(6)
    flx = {k: rng.uniform(0, 1e-4, PCOLS) for k in ("vap", "cnd", "ice", "sen")}
    cur = {"te": rng.uniform(2e9, 3e9, PCOLS), "tw": rng.uniform(10, 50, PCOLS)}
    cur["tw"][1] = 0.0  # a dry column: no relative water error
    tnd = {k: rng.normal(0, 1, PCOLS) for k in ("te", "tw")}
    w_cur = {k: v.copy() for k, v in cur.items()}
    w_tnd = {k: v.copy() for k, v in tnd.items()}
    te_rer, tw_rer = G.check_energy_chng(NCOL, st, 1200.0, flx["vap"], flx["cnd"], flx["ice"], flx["sen"],
                                         w_cur["te"], w_cur["tw"], w_tnd["te"], w_tnd["tw"])
    g_cur = {k: _dev(v) for k, v in cur.items()}
    g_tnd = {k: _dev(v) for k, v in tnd.items()}
    g_rer = {k: _dev(np.zeros(PCOLS)) for k in ("te", "tw")}
    crm_glue.check_energy_chng(hip, NCOL, PCOLS, PVER, {k: _dev(v) for k, v in st.items()}, 1200.0,
                               {k: _dev(v) for k, v in flx.items()}, g_cur, g_tnd, g_rer)
    assert np.array_equal(_host(g_rer["te"])[:NCOL], te_rer)
    assert np.array_equal(_host(g_rer["tw"])[:NCOL], tw_rer)
    for k in ("te", "tw"):
        assert np.array_equal(_host(g_cur[k]), w_cur[k]), k
        assert np.array_equal(_host(g_tnd[k]), w_tnd[k]), k


def test_gcm_step_on_the_device(hip):
    """physics state -> crm_input -> crm() -> tendencies entirely on the device, against the
    host chain with the oracle CRM: the CRM parity bar (one GCM step of 3 substeps)."""
    st = glue_cases.physics_state(NCOL, PCOLS, PVER, seed=15)
    srf = glue_cases.surface(PCOLS)
    angle = np.zeros(PCOLS)
    case = _case()
    case.rad["qrad"][...] = 0.0
    # host chain
    want = case.copy()
    G.gcm_input(NCOL, CRM_NZ, st, srf, angle, want.inp, want.rad["qrad"])
    assert oracle_lib.oracle().run(want, 60.0) == 0
    # device chain
    dc = DeviceCase(case, "cuda:0")
    crm_glue.gcm_input(hip, NCOL, PCOLS, PVER, CRM_NZ, 8, {k: _dev(v) for k, v in st.items()},
                       {k: _dev(v) for k, v in srf.items()}, _dev(angle), dc)
    hip.run_device(dc, 60.0)
    got = dc.download(case.copy())
    errs = compare_case(got, want, dt_gl=60.0)
    assert max(errs.values()) <= 1e-10, sorted(errs.items(), key=lambda kv: -kv[1])[:4]


def test_rad_columns_and_qrad(hip):
    """The radiation's inputs in every CRM radiation column (radiation.F90:1584-1636) from the
    crm_rad averages of a real crm() call (radiation on 8 column groups), and crm_qrad from the
    per-column shortwave and longwave heating (:2393-2398), bitwise against the restatement;
    columns past ncol untouched."""
    import torch
    case = _case()
    hip.run(case, 200.0)  # crm_rad averages with clouds in some groups
    st = glue_cases.physics_state(NCOL, PCOLS, PVER, seed=21)
    rng = np.random.default_rng(4)
    cloud = {f: np.asfortranarray(rng.uniform(0, 1, (PCOLS, PVER))) for f in crm_glue.CLOUD}
    nrad = 8
    want = G.rad_columns(NCOL, CRM_NZ, st, cloud, case.rad)
    dc = DeviceCase(case, "cuda:0")
    out = {f: torch.full((PVER, nrad, PCOLS), -7.0, dtype=torch.float64, device="cuda:0")
           for f in crm_glue.RAD_COLUMN}
    crm_glue.rad_columns(hip, NCOL, PCOLS, PVER, CRM_NZ, nrad, {k: _dev(v) for k, v in st.items()},
                         {k: _dev(v) for k, v in cloud.items()}, dc, out)
    for f in crm_glue.RAD_COLUMN:
        got = out[f].cpu().numpy().T  # (pcols, nrad, pver)
        assert np.array_equal(got[:NCOL], want[f][:NCOL]), f
        assert np.all(got[NCOL:] == -7.0), f
    assert (want["cld"][:NCOL, :, PVER - CRM_NZ:] > 0).any()  # some cloud in the averages
    qrs = rng.normal(0, 1e-5, case.rad["qrad"].shape)
    qrl = rng.normal(0, 1e-5, case.rad["qrad"].shape)
    want_q = case.rad["qrad"].copy()
    G.rad_qrad(NCOL, qrs, qrl, want_q)
    crm_glue.rad_qrad(hip, NCOL, PCOLS, nrad, CRM_NZ, _dev4(qrs), _dev4(qrl), dc)
    got_q = dc.download(case.copy()).rad["qrad"]
    assert np.array_equal(got_q, want_q)


def _dev4(a):
    """A (pcols, nx_rad, ny_rad, nz) Fortran array as a device tensor in the same memory order."""
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(3, 2, 1, 0))).cuda()
