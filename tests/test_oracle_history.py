"""The history restatement (oracle/history_oracle.py, crm_physics.F90:1201-1461) pinned by what
the quantities are: a constructed cloud deck's in-cloud and column water paths, the vapour of a
state built from its parts, the CRM heating with the radiation taken back out, the precise
4-byte 1e-3 of the reference build, and time means of samples with a known mean.  Each group
also fails for a planted mistake of the kind it guards against."""
import numpy as np
import pytest

import history_oracle as H

NCOL, PVER = 3, 5


def _out(gice, gliq, cld):
    names = ["qcl", "qci", "qpl", "qpi", "prec_crm", "tk", "tkh", "mctot", "mcup", "mcdn", "mcuup",
             "mcudn", "qc_mean", "qi_mean", "qs_mean", "qg_mean", "qr_mean", "flux_qt", "flux_u",
             "flux_v", "tkez", "tkesgsz", "tkz", "fluxsgs_qt", "flux_qp", "precflux", "qt_ls",
             "qt_trans", "qp_trans", "qp_evp", "qp_fall", "qp_src", "t_ls", "cltot", "clhgh",
             "clmed", "cllow", "cldtop", "timing_factor", "mu_crm", "md_crm", "eu_crm", "du_crm",
             "ed_crm"]
    o = {n: np.zeros((NCOL, PVER)) for n in names}
    o.update(gicewp=gice, gliqwp=gliq, cld=cld)
    return o


def _deck():
    """Column 0: a half-covered liquid layer of 100 g/m2 in level 3; column 1: an overcast
    ice layer of 20 g/m2 over two levels; column 2: clear sky."""
    gi, gl, cld = (np.zeros((NCOL, PVER)) for _ in range(3))
    gl[0, 3], cld[0, 3] = 100.0, 0.5
    gi[1, 1:3], cld[1, 1:3] = 10.0, 1.0
    return gi, gl, cld


def test_water_paths_of_a_cloud_deck():
    gi, gl, cld = _deck()
    f = H.fields(out=_out(gi, gl, cld))
    # in-cloud: 100 g/m2 spread over half the column is 0.2 kg/m2 where the cloud is
    assert f["ICLDTWP"][0, 3] == pytest.approx(0.2, rel=1e-7)
    assert f["ICLDIWP"][1, 1] == pytest.approx(0.01, rel=1e-7)
    # column totals in kg/m2, liquid and ice apart and together
    assert f["TGCLDLWP"] == pytest.approx([0.1, 0.0, 0.0], rel=1e-7)
    assert f["TGCLDIWP"] == pytest.approx([0.0, 0.02, 0.0], rel=1e-7)
    assert f["TGCLDCWP"] == pytest.approx([0.1, 0.02, 0.0], rel=1e-7)
    # the grid-mean path stays in g/m2
    assert f["GCLDLWP"][0, 3] == 100.0 and f["GCLDLWP"][1, 2] == 10.0
    # clear sky is zero everywhere, and the 1% floor of the cover keeps it finite
    for k in ("ICLDTWP", "ICLDIWP", "GCLDLWP"):
        assert np.all(f[k][2] == 0.0) and np.all(np.isfinite(f[k]))


def test_in_cloud_path_floor_and_literal():
    """max(0.01, cld): a trace of water with no cover counts as 1% cover; 1.0e-3 is the 4-byte
    literal (1.0000000474974513e-3), not the exact decimal."""
    gl = np.zeros((NCOL, PVER))
    cld = np.zeros((NCOL, PVER))
    gl[0, 0] = 1.0
    f = H.fields(out=_out(np.zeros_like(gl), gl, cld))
    assert f["ICLDTWP"][0, 0] == float(np.float32(1e-3)) / 0.01
    assert f["ICLDTWP"][0, 0] != 1e-3 / 0.01
    assert H.KG_PER_G == 1.0000000474974513e-3


def test_planted_mistakes_in_water_paths_are_caught(monkeypatch):
    gi, gl, cld = _deck()
    # a path divided by the cover without the floor: clear-sky cells blow up
    monkeypatch.setattr(H, "incloud", lambda p, c: p * H.KG_PER_G / c)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = H.fields(out=_out(gi, gl, cld))
    assert not np.all(np.isfinite(f["ICLDTWP"]))
    monkeypatch.undo()
    # grid-mean column path taken in g/m2 (the conversion dropped)
    monkeypatch.setattr(H, "KG_PER_G", 1.0)
    f = H.fields(out=_out(gi, gl, cld))
    assert f["TGCLDLWP"][0] != pytest.approx(0.1, rel=1e-7)


def test_vapour_and_heating():
    """CRM_QV is what is left of the total water after the condensate; SPDT is the CRM's own
    heating: the radiative heating put into ptend_s is taken back out, in K/s."""
    rng = np.random.default_rng(1)
    qv, qcl, qci = rng.uniform(0, 1e-2, (3, NCOL, 4, 1, 6))
    out = _out(*[np.zeros((NCOL, PVER))] * 3)
    out.update(qcl=qcl, qci=qci)
    state = {"u_wind": qv, "v_wind": qv, "w_wind": qv, "temperature": qv, "qt": qv + qcl + qci}
    f = H.fields(state=state, out=out)
    assert np.allclose(f["CRM_QV"], qv, rtol=0, atol=1e-17)
    heat = rng.normal(0, 1e-4, (NCOL, PVER))           # K/s of the CRM
    qrs, qrl = rng.normal(0, 1e-2, (2, NCOL, PVER))    # W/kg after the division by pdel
    ptend = {"s": heat * H.CPAIR + qrs + qrl, "qv": heat, "qcl": heat, "qci": heat}
    f = H.fields(ptend=ptend, qrs=qrs, qrl=qrl)
    assert np.allclose(f["SPDT"], heat, rtol=1e-9, atol=1e-18)
    assert np.allclose(f["SPQRS"] * H.CPAIR, qrs, rtol=1e-15)
    # the radiation left in (a dropped term) is caught
    wrong = (ptend["s"] - qrs) / H.CPAIR
    assert not np.allclose(wrong, heat, rtol=1e-6, atol=1e-12)


def test_time_mean():
    """Three calls of a linear ramp average to the middle call; counts per field."""
    h = H.History()
    base = np.arange(NCOL * PVER, dtype=float).reshape(NCOL, PVER)
    for j in range(3):
        h.accumulate(NCOL, pmid=base + 10.0 * j, pdel=base)
    m, n = h.mean("PRES")
    assert n == 3 and np.allclose(m, base + 10.0, rtol=1e-15)
    m, n = h.mean("DPRES")
    assert n == 3 and np.allclose(m, base, rtol=1e-15)
    # only the chunk's columns are averaged
    h2 = H.History()
    h2.accumulate(2, pmid=base)
    assert h2.mean("PRES")[0].shape == (2, PVER)


def test_library_fields_are_the_restatement_fields():
    """The device history (host-side table of the library, no GPU call) holds exactly the
    fields the restatement defines, in the reference's outfld order starting with CRM_QRAD."""
    import crm_history
    names = crm_history.field_names()
    inputs = dict(state={k: np.zeros(1) for k in ("u_wind", "v_wind", "w_wind", "temperature", "qt")},
                  out=_out(*[np.zeros((NCOL, PVER))] * 3), qrad=0, pmid=0, pdel=0,
                  ptend={k: 0 for k in ("s", "qv", "qcl", "qci")}, qrs=0, qrl=0)
    inputs["out"].update(qcl=np.zeros(1), qci=np.zeros(1))
    assert sorted(names) == sorted(H.fields(**inputs))
    assert names[0] == "CRM_QRAD" and names[-1] == "ED_CRM" and len(names) == 65
