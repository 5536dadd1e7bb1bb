"""The CPU restatement of the GCM-side code (oracle/glue_oracle.py) pinned by what the physics
fixes: closed forms of the column integrals, the closure of check_energy_chng for a state
advanced by exactly the fluxes it is told about (a dropped latent or sensible term fails it),
rotation invariants of the input wind, and the Q dp round trip of qrad."""
import numpy as np
import pytest

import glue_cases
import glue_oracle as G

NCOL, PCOLS, PVER, CRM_NZ = 5, 7, 30, 28


def _state():
    return glue_cases.physics_state(NCOL, PCOLS, PVER, seed=3)


def _inp(pcols, pver):
    F = lambda *s: np.zeros(s, order="F")
    inp = {k: F(pcols, pver) for k in ("zmid", "tl", "ql", "qccl", "qiil", "pmid", "pdel", "ul", "vl")}
    inp.update({k: F(pcols, pver + 1) for k in ("zint", "pint")})
    inp.update({k: F(pcols) for k in ("ps", "phis", "ocnfrac", "tau00", "wndls", "bflxls", "fluxu00",
                                      "fluxv00", "fluxt00", "fluxq00")})
    return inp


def test_column_integrals_closed_forms():
    st = _state()
    dry = {k: v.copy() for k, v in st.items()}
    for k in ("u", "v", "qv", "qcl", "qci"):
        dry[k][...] = 0.0
    dry["phis"][...] = 0.0
    te, tw = G.energy_integrals(NCOL, dry)
    want = (dry["t"][:NCOL] * dry["pdel"][:NCOL]).sum(1) * G.CPAIR / G.GRAVIT
    assert np.allclose(te, want, rtol=1e-14, atol=0) and np.all(tw == 0.0)
    # vapour adds (latvap + latice) per unit of column water, condensate latice (liquid) or 0 (ice)
    wet = {k: v.copy() for k, v in dry.items()}
    wet["qv"][...] = st["qv"]
    wet["qcl"][...] = st["qcl"]
    wet["qci"][...] = st["qci"]
    te2, tw2 = G.energy_integrals(NCOL, wet)
    col = lambda q: (q[:NCOL] * st["pdel"][:NCOL]).sum(1) / G.GRAVIT
    assert np.allclose(te2 - te, (G.LATVAP + G.LATICE) * col(st["qv"]) + G.LATICE * col(st["qcl"]), rtol=1e-9)
    assert np.allclose(tw2, col(st["qv"]) + col(st["qcl"]) + col(st["qci"]), rtol=1e-14)
    # kinetic energy and surface geopotential
    wind = {k: v.copy() for k, v in dry.items()}
    wind["u"][...] = st["u"]
    wind["v"][...] = st["v"]
    wind["phis"][...] = st["phis"]
    te3, _ = G.energy_integrals(NCOL, wind)
    ke = (0.5 * (st["u"] ** 2 + st["v"] ** 2)[:NCOL] * st["pdel"][:NCOL]).sum(1) / G.GRAVIT
    assert np.allclose(te3 - te, ke + st["phis"][:NCOL] * st["ps"][:NCOL] / G.GRAVIT, rtol=1e-9)


def test_energy_check_closes_for_the_flux_bypass():
    """A state advanced by the bypass tendencies over dt matches what check_energy_chng expects
    from the same sensible and vapour fluxes: relative errors at round-off."""
    st = _state()
    srf = glue_cases.surface(PCOLS)
    te_cur, tw_cur = np.zeros(PCOLS), np.zeros(PCOLS)
    te_t, tw_t = np.zeros(PCOLS), np.zeros(PCOLS)
    z = np.zeros(PCOLS)
    G.check_energy_chng(NCOL, st, 1200.0, z, z, z, z, te_cur, tw_cur, te_t, tw_t)  # sets te/tw_cur
    ps_, pq = np.zeros((PCOLS, PVER), order="F"), np.zeros((PCOLS, PVER), order="F")
    G.flux_bypass(NCOL, st["pdel"], srf["shf"], srf["cflx"], ps_, pq)
    dt = 1200.0
    st["t"][:NCOL] += ps_[:NCOL] / G.CPAIR * dt  # physics_update: s -> t, q
    st["qv"][:NCOL] += pq[:NCOL] * dt
    te_rer, tw_rer = G.check_energy_chng(NCOL, st, dt, srf["cflx"], z, z, srf["shf"], te_cur, tw_cur, te_t, tw_t)
    print("energy / water relative errors", np.abs(te_rer).max(), np.abs(tw_rer).max())
    assert np.abs(te_rer).max() < 1e-14
    assert np.abs(tw_rer).max() < 1e-13
    # a scheme that forgot the latent heat of the vapour flux does not close
    te_cur2 = te_cur.copy()
    st["qv"][:NCOL] += pq[:NCOL] * dt
    bad, _ = G.check_energy_chng(NCOL, st, dt, z, z, z, srf["shf"] * 0.0, te_cur2, tw_cur.copy(), te_t, tw_t)
    assert np.abs(bad).max() > 1e-6


def test_input_wind_rotation():
    st = _state()
    srf = glue_cases.surface(PCOLS)
    qrad = np.ones((PCOLS, 4, 1, CRM_NZ), order="F")
    a0 = G.gcm_input(NCOL, CRM_NZ, st, srf, np.zeros(PCOLS), _inp(PCOLS, PVER), qrad.copy(order="F"))
    assert np.array_equal(a0["ul"][:NCOL], st["u"][:NCOL]) and np.array_equal(a0["vl"][:NCOL], st["v"][:NCOL])
    a1 = G.gcm_input(NCOL, CRM_NZ, st, srf, np.full(PCOLS, np.pi / 2), _inp(PCOLS, PVER), qrad.copy(order="F"))
    assert np.allclose(a1["ul"][:NCOL], st["v"][:NCOL], atol=1e-14)
    assert np.allclose(a1["vl"][:NCOL], -st["u"][:NCOL], atol=1e-14)
    ang = np.random.default_rng(2).uniform(0, 2 * np.pi, PCOLS)
    a2 = G.gcm_input(NCOL, CRM_NZ, st, srf, ang, _inp(PCOLS, PVER), qrad.copy(order="F"))
    sp = lambda u, v: np.hypot(u, v)
    assert np.allclose(sp(a2["ul"], a2["vl"])[:NCOL], sp(st["u"], st["v"])[:NCOL], rtol=1e-14)
    assert np.allclose(a2["tau00"][:NCOL], np.hypot(srf["wsx"], srf["wsy"])[:NCOL], rtol=4e-16, atol=0)
    assert np.all(a2["tl"][NCOL:] == 0.0)  # columns past ncol untouched


def test_qrad_round_trip_and_tendencies():
    st = _state()
    rng = np.random.default_rng(5)
    nrx = 4
    qrad0 = np.asfortranarray(rng.normal(0, 1e-5, (PCOLS, nrx, 1, CRM_NZ)))
    qrad = qrad0.copy(order="F")
    # carried as Q dp between steps: the input step divides, the tendency step multiplies
    for m in range(CRM_NZ):
        qrad[:NCOL, :, :, m] *= st["pdel"][:NCOL, PVER - 1 - m][:, None, None]
    G.gcm_input(NCOL, CRM_NZ, st, glue_cases.surface(PCOLS), np.zeros(PCOLS), _inp(PCOLS, PVER), qrad)
    assert np.allclose(qrad[:NCOL], qrad0[:NCOL], rtol=4e-16, atol=0)
    F = lambda: np.asfortranarray(rng.normal(0, 1e-4, (PCOLS, PVER)))
    out = {k: F() for k in ("qltend", "qcltend", "qiltend", "sltend")}
    out.update({k: rng.uniform(0, 1e-7, PCOLS) for k in ("precc", "precl", "precsc", "precsl")})
    tot, tots = out["precc"] + out["precl"], out["precsc"] + out["precsl"]
    qrs, qrl = F() * st["pdel"], F() * st["pdel"]
    qsum = (qrs + qrl)[:NCOL].sum(1) / G.GRAVIT
    ptend = {k: np.zeros((PCOLS, PVER), order="F") for k in ("s", "qv", "qcl", "qci")}
    f = G.gcm_tend(NCOL, CRM_NZ, st["pdel"], out, qrad, qrs, qrl, ptend)
    assert np.allclose(f, qsum, rtol=1e-13)
    assert np.array_equal(out["precc"][:NCOL], tot[:NCOL]) and np.all(out["precl"][:NCOL] == 0)
    assert np.array_equal(out["precsc"][:NCOL], tots[:NCOL])
    top = PVER - CRM_NZ + 2
    assert np.array_equal(ptend["s"][:NCOL, top:], out["sltend"][:NCOL, top:])
    assert np.array_equal(ptend["s"][:NCOL, :top], (qrs + qrl)[:NCOL, :top])
    assert np.array_equal(ptend["qcl"][:NCOL], out["qcltend"][:NCOL])


def _rad_column_case():
    """pver 5, crm_nz 3, radiation columns 2 x 1: GCM values 100 + level (state) and 0.2 + 0.01
    level (clouds); CRM level m of radiation column r holds distinct averages."""
    pcols, pver, nz, ncol = 3, 5, 3, 2
    lev = np.arange(pver, dtype=float)
    state = {f: np.asfortranarray(np.tile(base + lev, (pcols, 1)))
             for f, base in (("t", 100.0), ("qv", 1e-3), ("qcl", 2e-3), ("qci", 3e-3))}
    state["pdel"] = np.asfortranarray(np.tile(1000.0 + 10 * lev, (pcols, 1)))
    cloud = {f: np.asfortranarray(np.tile(0.2 + 0.01 * lev, (pcols, 1))) for f in ("cld", "fice", "cicewp", "cliqwp")}
    rad = {f: np.zeros((pcols, 2, 1, nz), order="F") for f in ("temperature", "qv", "qc", "qi", "cld")}
    rad["temperature"][:, 0, 0, :] = [250.0, 260.0, 270.0]
    rad["temperature"][:, 1, 0, :] = [251.0, 261.0, 271.0]
    rad["qv"][:, :, 0, :] = 5e-3
    rad["qv"][:, 1, 0, 2] = -1e-12                      # a negative average: the 1e-9 floor
    rad["cld"][:, 0, 0, :] = [0.5, 0.0, 0.001]          # cloudy, clear, a trace (the 1% floor)
    rad["qc"][:, 0, 0, :] = [1e-4, 0.0, 1e-6]
    rad["qi"][:, 0, 0, :] = [3e-4, 0.0, 0.0]
    return ncol, nz, state, cloud, rad


def test_rad_columns_closed_forms():
    """radiation.F90:1584-1636: per radiation column, the CRM levels (GCM level pver-1-m of CRM
    level m) take the crm_rad averages; the in-cloud water paths are q pdel / g over the cloud
    fraction floored at 1%; the ice fraction is qi / (qc + qi); vapour is floored at 1e-9; the
    levels above the CRM keep the GCM column's values."""
    ncol, nz, state, cloud, rad = _rad_column_case()
    out = G.rad_columns(ncol, nz, state, cloud, rad)
    # levels above the CRM (k = 0, 1) untouched, in both radiation columns
    for f, src in (("t", state["t"]), ("qc", state["qcl"]), ("cld", cloud["cld"])):
        assert np.array_equal(out[f][:ncol, :, :2], np.repeat(src[:ncol, None, :2], 2, axis=1))
    # level order: CRM level 0 is the bottom GCM level
    assert out["t"][0, 0, 4] == 250.0 and out["t"][0, 0, 2] == 270.0 and out["t"][1, 1, 3] == 261.0
    # the cloudy cell: 0.5 cover, 0.1 g/kg liquid, 0.3 g/kg ice, pdel 1040 Pa at k = 4
    assert out["fice"][0, 0, 4] == pytest.approx(0.75, rel=1e-15)
    assert out["cicewp"][0, 0, 4] == pytest.approx(3e-4 * 1040.0 / 9.80616 / 0.5, rel=1e-15)
    assert out["cliqwp"][0, 0, 4] == pytest.approx(1e-4 * 1040.0 / 9.80616 / 0.5, rel=1e-15)
    # clear: zero paths and ice fraction; a trace of cover counts as 1%
    assert out["cicewp"][0, 0, 3] == 0.0 and out["fice"][0, 0, 3] == 0.0 and out["cld"][0, 0, 3] == 0.0
    assert out["cliqwp"][0, 0, 2] == pytest.approx(1e-6 * 1020.0 / 9.80616 / 0.01, rel=1e-15)
    # vapour floor
    assert out["qv"][0, 1, 2] == 1e-9 and out["qv"][0, 0, 2] == 5e-3
    # columns past ncol stay zero
    assert not out["t"][ncol:].any()


def test_rad_columns_planted_mistakes(monkeypatch):
    """The checks notice a dropped cover floor and a flipped level order."""
    ncol, nz, state, cloud, rad = _rad_column_case()
    # the in-cloud path divided by the cover itself: the trace cell comes out 10x too large
    monkeypatch.setattr(G, "incloud_path", lambda q, dp, cld: q * dp / G.GRAVIT / cld)
    got = G.rad_columns(ncol, nz, state, cloud, rad)
    assert got["cliqwp"][0, 0, 2] != pytest.approx(1e-6 * 1020.0 / 9.80616 / 0.01, rel=1e-3)
    monkeypatch.undo()
    # levels taken top-first: the CRM's bottom temperature lands at the wrong GCM level
    flipped = {f: v[..., ::-1].copy() for f, v in rad.items()}
    got = G.rad_columns(ncol, nz, state, cloud, flipped)
    assert got["t"][0, 0, 4] != 250.0


def test_rad_qrad_is_shortwave_plus_longwave():
    q = np.zeros((3, 2, 1, 4))
    qrs, qrl = np.full_like(q, 1.5), np.full_like(q, -0.25)
    G.rad_qrad(2, qrs, qrl, q)
    assert np.all(q[:2] == 1.25) and not q[2].any()
