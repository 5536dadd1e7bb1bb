"""Closed forms of the statistics that crm() accumulates every substep
(crm_module.F90:896-1073) and converts at the end of the call (crm_module.F90:1078-1659).

These outputs are restated (the reference's crm_module.F90 needs CAM modules and is not compiled
here), so they are pinned by what a constructed state fixes: a cloud deck in known columns and
levels, uniform updrafts and downdrafts, a prescribed shear, known surface rain rates.  The
oracle's statistics run on such a state through oracle_stats_probe, and each group of outputs is
compared with the values the construction gives.

Each group also runs with a planted mistake in the oracle (oracle_mutate: a swapped branch, a
wrong level index, a dropped factor) and must fail, which shows the checks notice that kind of
slip.  The GPU tests hold the library's statistics kernels to the oracle at 1e-10 on full
calls, so these closed forms pin the library's statistics too."""
import numpy as np
import pytest

import oracle_lib
from crm_inputs import make_case

G = 9.80616
NX, NZ, DX, DT = 32, 28, 4000.0, 20.0
PLEV = NZ + 2
THR = float(np.float32(0.001))  # cwp_threshold (crm_module.F90:134), kg/m2


def gcm(k):
    """GCM level (0-based, top first) of CRM level k (0-based, bottom first)."""
    return PLEV - 1 - k


def cldtop_slot(k):
    """GCM level (0-based) where crm() records a cloud top at CRM level k.  The cloud-top scan
    runs over CRM levels from the top, nz-k for k = 1..nzm, but stores at l = plev-k+1
    (crm_module.F90:913-922), the bottom-up mapping of the loop counter: the record sits at the
    GCM level mirrored about the CRM's levels, plev - nzm + k (0-based), not at gcm(k).  The
    reference does this and the library keeps it."""
    return PLEV - NZ + k


def gcm_iface(k0):
    """GCM interface (0-based; 0 the top, PLEV the surface) of CRM interface k0."""
    return PLEV - k0


class Built:
    """A constructed state on the profiles of a quiet tropical sounding."""

    def __init__(self, nx_rad=None):
        self.case = make_case(1, NX, NZ, DX, DT, nx_rad=nx_rad)
        self.probe = oracle_lib.StatsProbe()
        self.probe.set_up(self.case)
        get = self.probe.get
        self.rho = get("rho", NZ)
        self.adz = get("adz", NZ)
        self.pres = get("pres", NZ)
        self.rhow = get("rhow", NZ + 1)
        self.dz = float(get("dz", 1)[0])
        self.mass = self.rho * self.adz * self.dz  # kg/m2 per CRM level
        self.pdel = self.case.inp["pdel"][0].copy()
        t = self.case.state["temperature"][0, :, 0, :].T.copy()  # (NZ, NX)
        self.f = {
            "u": np.zeros((NZ, NX)), "w": np.zeros((NZ + 1, NX)), "tabs": t,
            "qv": self.case.state["qt"][0, :, 0, :].T.copy(), "qcl": np.zeros((NZ, NX)),
            "qci": np.zeros((NZ, NX)), "qpl": np.zeros((NZ, NX)), "qpi": np.zeros((NZ, NX)),
            "tke": np.zeros((NZ, NX)), "tk": np.zeros((NZ, NX)), "precsfc": np.zeros(NX),
            "precssfc": np.zeros(NX), "u0": np.zeros(NZ)}
        for name in ("uwle", "uwsb", "mkwle1", "mkwle2", "mkwsb1", "mkwsb2", "mkadv1", "mkadv2",
                     "mkdiff1", "mkdiff2", "qpsrc", "qpevp", "qpfall", "precflux"):
            self.f[name] = np.zeros(NZ + 1)
        # pressure bands of the cloud-layer split (crm_module.F90:925-935)
        self.low = [k for k in range(NZ) if self.pres[k] >= 700.0]
        self.mid = [k for k in range(NZ) if 400.0 <= self.pres[k] < 700.0]
        self.high = [k for k in range(NZ) if self.pres[k] < 400.0 and self.pres[k] > 150.0]

    def run(self, nsteps=1, dt_gl=1200.0, dtn=None):
        for name, v in self.f.items():
            self.probe.set(name, v)
        c = self.case
        for f in ("cltot", "clhgh", "clmed", "cllow", "prectend", "precstend"):
            c.out[f][...] = 0.0
        dtn = dt_gl / nsteps if dtn is None else dtn
        assert self.probe.run(c, dt_gl, dtn, nsteps) == 0
        return c.out, c.rad


# ---- groups of checks: each builds its state, runs the statistics and asserts ----------------

def cloud_decks(b):
    """Three decks in disjoint columns: liquid in the 400-700 hPa band (columns 4-11), liquid
    below 700 hPa (20-23), ice above 400 hPa (0-1).  Every deck level holds more than the
    cloud threshold, so each deck level is cloudy and its top level is the column's cloud top.
    Returns {deck: (columns, levels)}."""
    decks = {"mid": (range(4, 12), b.mid[len(b.mid) // 2 - 1: len(b.mid) // 2 + 2]),
             "low": (range(20, 24), b.low[2:4]), "high": (range(0, 2), b.high[-2:])}
    for name, (cols, levs) in decks.items():
        for k in levs:
            q = max(1e-4, 3.0 * THR / b.mass[k])
            b.f["qci" if name == "high" else "qcl"][k, list(cols)] = q
    return decks


def check_cloud_fractions(b, nsteps=3):
    """cld: fraction of columns whose level holds more than the threshold; cldtop: fraction
    whose cloud top (first level from the top where the column's cumulative water path
    exceeds the threshold) is that level, stored at cldtop_slot; cltot / cllow / clmed /
    clhgh: fraction of columns with a cloudy column / low (>= 700 hPa) / middle / high
    (< 400 hPa) layer.  Time means over the
    substeps, so nsteps identical substeps give the single-substep values."""
    decks = cloud_decks(b)
    out, _ = b.run(nsteps)
    cld = np.zeros(PLEV)
    top = np.zeros(PLEV)
    for cols, levs in decks.values():
        for k in levs:
            cld[gcm(k)] += len(cols) / NX
        top[cldtop_slot(max(levs))] += len(cols) / NX
    np.testing.assert_allclose(out["cld"][0], cld, rtol=1e-14, atol=0)
    np.testing.assert_allclose(out["cldtop"][0], top, rtol=1e-14, atol=0)
    frac = {n: len(c) / NX for n, (c, _) in decks.items()}
    np.testing.assert_allclose(out["cltot"][0], sum(frac.values()), rtol=1e-14)
    np.testing.assert_allclose(out["cllow"][0], frac["low"], rtol=1e-14)
    np.testing.assert_allclose(out["clmed"][0], frac["mid"], rtol=1e-14)
    np.testing.assert_allclose(out["clhgh"][0], frac["high"], rtol=1e-14)


def check_thin_cloud_top(b):
    """A thin cloud: six levels each holding 0.35 of the threshold.  No level is cloudy, but
    the cumulative path from the top passes the threshold at the third level from the top,
    which is therefore the cloud top of those columns (and the column counts in cltot)."""
    levs = b.mid[:6] if len(b.mid) >= 6 else (b.low + b.mid)[-6:]
    cols = list(range(8, 16))
    for k in levs:
        b.f["qcl"][k, cols] = 0.35 * THR / b.mass[k]
    out, _ = b.run(2)
    top = np.zeros(PLEV)
    top[cldtop_slot(sorted(levs)[-3])] = len(cols) / NX
    np.testing.assert_allclose(out["cldtop"][0], top, rtol=1e-14, atol=0)
    assert np.all(out["cld"][0] == 0.0)
    np.testing.assert_allclose(out["cltot"][0], len(cols) / NX, rtol=1e-14)


def check_water_paths(b):
    """gliqwp / gicewp [g/m2]: the level-mean liquid / ice mixing ratio times the layer mass
    pdel/g, per GCM level."""
    rng = np.random.default_rng(11)
    b.f["qcl"][:] = np.where(rng.random((NZ, NX)) < 0.4, rng.uniform(0, 1e-3, (NZ, NX)), 0.0)
    b.f["qci"][:] = np.where(rng.random((NZ, NX)) < 0.3, rng.uniform(0, 2e-4, (NZ, NX)), 0.0)
    out, _ = b.run(2)
    for field, q in (("gliqwp", b.f["qcl"]), ("gicewp", b.f["qci"])):
        want = np.zeros(PLEV)
        for k in range(NZ):
            want[gcm(k)] = q[k].mean() * b.pdel[gcm(k)] / G * 1000.0
        np.testing.assert_allclose(out[field][0], want, rtol=1e-13, atol=0)


def check_cloud_mass_fluxes(b):
    """mcup / mcdn: rho * w (cell-centre w, the mean of the bounding interfaces) summed over
    cloudy cells whose centre w exceeds +2 m/s (below -2 m/s), as a domain mean; mcuup /
    mcudn: the same over clear cells; mctot: their sum.  Columns with a cloudy centre w of
    1.5 m/s contribute nothing."""
    decks = cloud_decks(b)
    mid_cols, mid = decks["mid"]
    low_cols, low = decks["low"]
    w = b.f["w"]
    w[:, list(mid_cols)] = 3.0
    weak = [12, 13]
    for k in mid:
        b.f["qcl"][k, weak] = 1e-4 + 3.0 * THR / b.mass[k]
    w[:, weak] = 1.5
    w[:, list(low_cols)] = -2.5
    w[:, 24:28] = -3.0
    w[:, 28:32] = 2.2
    out, _ = b.run(2)
    want = {f: np.zeros(PLEV) for f in ("mcup", "mcdn", "mcuup", "mcudn")}
    for k in range(NZ):
        l, r = gcm(k), b.rho[k]
        if k in mid:
            want["mcup"][l] += r * 3.0 * len(mid_cols) / NX
        else:
            want["mcuup"][l] += r * 3.0 * len(mid_cols) / NX
        if k in low:
            want["mcdn"][l] += r * -2.5 * len(low_cols) / NX
        else:
            want["mcudn"][l] += r * -2.5 * len(low_cols) / NX
        want["mcudn"][l] += r * -3.0 * 4 / NX
        want["mcuup"][l] += r * 2.2 * 4 / NX
    for f, v in want.items():
        np.testing.assert_allclose(out[f][0], v, rtol=1e-13, atol=0, err_msg=f)
    np.testing.assert_allclose(out["mctot"][0], sum(want.values()), rtol=1e-13, atol=1e-300)


def check_entrainment(b, nsteps=3):
    """Interface mass fluxes (Xu et al. 2002 as in crm_module.F90:1017-1046): an updraft
    counts at an interface when the level below it is cloudy, a downdraft when the level above
    is cloudy or holds more than 1e-4 kg/kg of precipitation; they are summed over columns and
    substeps.  Then per GCM level (crm_module.F90:1483-1511): mu / md = mean of the bounding
    interface fluxes * g/100 [hPa/s]; eu - du = g * (flux at the top - flux at the bottom) /
    pdel with eu, du >= 0 and not both nonzero; ed the increase of the downward flux below;
    jt / mx the first / last level (1-based, from the top) with a nonzero mu or md.  Every
    entrained mass is detrained again, so sum (eu - du) * pdel / g = 0."""
    cols_up = list(range(4, 12))
    deck = b.mid[len(b.mid) // 2 - 1: len(b.mid) // 2 + 2]
    for k in deck:
        b.f["qcl"][k, cols_up] = 1e-4
    b.f["w"][:, cols_up] = 3.0
    cols_dn = list(range(24, 28))
    rain = b.low[1:4] + b.mid[:1]
    for k in rain:
        b.f["qpl"][k, cols_dn] = 2e-4
    b.f["w"][:, cols_dn] = -2.0
    b.f["w"][:, 28:32] = 2.2  # clear updraft: no interface flux
    out, _ = b.run(nsteps)
    mui = np.zeros(PLEV + 1)
    mdi = np.zeros(PLEV + 1)
    for k0 in range(NZ + 1):  # CRM interfaces, 0 = surface
        below, above = max(k0 - 1, 0), min(k0 + 1, NZ - 1)
        if below in deck:
            mui[gcm_iface(k0)] += nsteps * len(cols_up) * b.rhow[k0] * 3.0
        if above in rain:
            mdi[gcm_iface(k0)] += nsteps * len(cols_dn) * b.rhow[k0] * -2.0
    assert mui.any() and mdi.any()
    mu = 0.5 * (mui[:-1] + mui[1:]) * G / 100.0
    md = 0.5 * (mdi[:-1] + mdi[1:]) * G / 100.0
    np.testing.assert_allclose(out["mu_crm"][0], mu, rtol=1e-13, atol=0)
    np.testing.assert_allclose(out["md_crm"][0], md, rtol=1e-13, atol=0)
    eu, du, ed = out["eu_crm"][0], out["du_crm"][0], out["ed_crm"][0]
    assert np.all(eu >= 0) and np.all(du >= 0) and np.all(ed >= 0)
    assert not np.any((eu > 0) & (du > 0))
    np.testing.assert_allclose(eu - du, G * (mui[:-1] - mui[1:]) / b.pdel, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(ed, np.maximum(mdi[:-1] - mdi[1:], 0.0) * G / b.pdel, rtol=1e-13, atol=0)
    assert abs(((eu - du) * b.pdel / G).sum()) <= 1e-12 * mui.max()
    # cloud base entrains, the level above the deck top detrains
    assert eu[gcm(min(deck))] > 0 and du[gcm(max(deck) + 1)] > 0
    levels = np.nonzero((np.abs(mu) > 1e-15) | (np.abs(md) > 1e-15))[0]
    assert out["jt_crm"][0] == levels.min() + 1
    assert out["mx_crm"][0] == levels.max() + 1


def check_tke_profiles(b):
    """A prescribed shear u = u0 + A cos(2 pi (i + 1/2) / nx) about the domain mean u0, a
    uniform w = B and a uniform SGS energy E: tkez = rho (1/2 (A^2/2 + B^2) + E), tkesgsz =
    rho E [kg/m/s2]; tkz the level-mean eddy viscosity."""
    rng = np.random.default_rng(3)
    A = rng.uniform(0.5, 4.0, NZ)
    u0 = rng.uniform(-5, 5, NZ)
    E = rng.uniform(0, 0.8, NZ)
    B = 1.3
    i = np.arange(NX)
    b.f["u"][:] = u0[:, None] + A[:, None] * np.cos(2 * np.pi * (i + 0.5) / NX)[None, :]
    b.f["u0"][:] = u0
    b.f["w"][:] = B
    b.f["tke"][:] = E[:, None]
    b.f["tk"][:] = rng.uniform(0, 50, (NZ, NX))
    out, _ = b.run(2)
    want_sgs = np.zeros(PLEV)
    want = np.zeros(PLEV)
    want_tkz = np.zeros(PLEV)
    for k in range(NZ):
        want_sgs[gcm(k)] = b.rho[k] * E[k]
        want[gcm(k)] = b.rho[k] * (0.5 * (A[k] ** 2 / 2 + B ** 2) + E[k])
        want_tkz[gcm(k)] = b.f["tk"][k].mean()
    np.testing.assert_allclose(out["tkesgsz"][0], want_sgs, rtol=1e-13, atol=0)
    np.testing.assert_allclose(out["tkez"][0], want, rtol=1e-12, atol=0)
    np.testing.assert_allclose(out["tkz"][0], want_tkz, rtol=1e-13, atol=0)


def check_flux_profiles(b):
    """Vertical flux profiles from what the phase routines accumulate in the last substep
    (each resets its sums per call: advect2_mom_z.F90:30, advect_scalar2D.F90:118,
    diffuse_scalar2D.F90:97), so each output is that substep's domain-mean flux / nstop:
    * uwle = sum_i rhow w u / dz (advect2_mom_z.F90:73-78): uniform w, u give flux_u = w u;
    * mkwle = sum_i rhow w q dtn / dz, MPDATA's donor-cell flux of a uniform q in Courant
      units (adams.F90:29-34, advect_scalar2D.F90:111-136) over the last subcycle dtn: gives
      flux_qt = rhow w q (the resolved part);
    * mkwsb = sum_i F / dz for a turbulent flux F (diffuse_scalar2D.F90:106-113): fluxsgs_qt.
    Tendency sums (kg/kg over each substep, summed over columns and substeps) become mean rates
    over dt_gl: qt_trans, qp_trans, qp_src, qp_evp, qp_fall; precflux (kg/m2 per dz, summed)
    becomes a mean surface-flux rate in m/s."""
    nsteps, dt_gl, dtn = 4, 80.0, 10.0  # substeps of 20 s, the last one in two subcycles
    rng = np.random.default_rng(5)
    W, U, Q, Qp = 1.7, -3.2, 4e-3, 6e-4
    Fs = rng.uniform(-1e-4, 1e-4, NZ + 1)
    Fsp = rng.uniform(-1e-5, 1e-5, NZ + 1)
    b.f["uwle"][:] = NX * b.rhow * W * U / b.dz
    b.f["mkwle1"][:] = NX * b.rhow * W * Q * dtn / b.dz
    b.f["mkwle2"][:] = NX * b.rhow * W * Qp * dtn / b.dz
    b.f["mkwsb1"][:] = NX * Fs / b.dz
    b.f["mkwsb2"][:] = NX * Fsp / b.dz
    rates = {n: rng.uniform(-1e-7, 1e-7, NZ + 1) for n in ("mkadv1", "mkdiff1", "mkadv2", "mkdiff2",
                                                           "qpsrc", "qpevp", "qpfall")}
    for n, r in rates.items():  # per-substep change r * dt_sub in every column
        b.f[n][:] = NX * nsteps * r * (dt_gl / nsteps)
    P = rng.uniform(0, 1e-3, NZ + 1)  # kg/m2/s through each level
    b.f["precflux"][:] = NX * nsteps * P * (dt_gl / nsteps) / b.dz
    out, _ = b.run(nsteps, dt_gl, dtn)

    def lev(a):  # (NZ + 1) CRM-level profile -> GCM levels
        w = np.zeros(PLEV)
        for k in range(NZ):
            w[gcm(k)] = a[k]
        return w

    tol = dict(rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(out["flux_u"][0], lev(np.full(NZ + 1, W * U)) / nsteps, **tol)
    np.testing.assert_allclose(out["fluxsgs_qt"][0], lev(Fs) / nsteps, **tol)
    np.testing.assert_allclose(out["flux_qt"][0], lev(b.rhow * W * Q + Fs) / nsteps, **tol)
    np.testing.assert_allclose(out["flux_qp"][0], lev(b.rhow * W * Qp + Fsp) / nsteps, **tol)
    np.testing.assert_allclose(out["qt_trans"][0], lev(rates["mkadv1"] + rates["mkdiff1"]), **tol)
    np.testing.assert_allclose(out["qp_trans"][0], lev(rates["mkadv2"] + rates["mkdiff2"]), **tol)
    for f, n in (("qp_src", "qpsrc"), ("qp_evp", "qpevp"), ("qp_fall", "qpfall")):
        np.testing.assert_allclose(out[f][0], lev(rates[n]), **tol)
    np.testing.assert_allclose(out["precflux"][0], lev(P) / 1000.0, **tol)


def check_surface_precipitation(b):
    """Surface rain per column P_i [kg/m2/s] accumulated over the call (precsfc = P dt_gl / dz,
    microphysics.F90:762-790); prec_crm = P / 1000 [m/s]; columns above 10 mm/day count as
    convective (precc), the others as stratiform (precl), each as a domain mean in m/s
    (crm_module.F90:1419-1470); the frozen parts likewise from precssfc."""
    dt_gl = 1200.0
    P = np.where(np.arange(NX) < 12, 20.0, 5.0) / 86400.0  # mm/day -> kg/m2/s
    b.f["precsfc"][:] = P * dt_gl / b.dz
    b.f["precssfc"][:] = 0.25 * P * dt_gl / b.dz
    out, _ = b.run(3, dt_gl)
    np.testing.assert_allclose(out["prec_crm"][0, :, 0], P / 1000.0, rtol=1e-13)
    conv = P > 10.0 / 86400.0
    np.testing.assert_allclose(out["precc"][0], P[conv].sum() / NX / 1000.0, rtol=1e-13)
    np.testing.assert_allclose(out["precl"][0], P[~conv].sum() / NX / 1000.0, rtol=1e-13)
    np.testing.assert_allclose(out["precsc"][0], 0.25 * P[conv].sum() / NX / 1000.0, rtol=1e-13)
    np.testing.assert_allclose(out["precsl"][0], 0.25 * P[~conv].sum() / NX / 1000.0, rtol=1e-13)


def check_radiation_groups(b):
    """crm_rad (crm_module.F90:979-1015, 1086-1107): per group of nx/nx_rad adjacent columns,
    the time mean of the group mean of tabs, of max(qv, 0), of qcl and qci, and the fraction of
    cells holding any condensate."""
    rng = np.random.default_rng(9)
    b.f["tabs"][:] += rng.uniform(-2, 2, (NZ, NX))
    b.f["qv"][:] = rng.uniform(-2e-4, 1e-2, (NZ, NX))
    b.f["qcl"][:] = np.where(rng.random((NZ, NX)) < 0.3, rng.uniform(0, 1e-3, (NZ, NX)), 0.0)
    b.f["qci"][:] = np.where(rng.random((NZ, NX)) < 0.2, rng.uniform(0, 1e-4, (NZ, NX)), 0.0)
    _, rad = b.run(3)
    nxr = b.case.nx_rad
    g = NX // nxr

    def grouped(a):  # (NZ, NX) -> (nxr, NZ) group means
        return a.reshape(NZ, nxr, g).mean(axis=2).T

    tol = dict(rtol=1e-13, atol=0)
    np.testing.assert_allclose(rad["temperature"][0, :, 0, :], grouped(b.f["tabs"]), **tol)
    np.testing.assert_allclose(rad["qv"][0, :, 0, :], grouped(np.maximum(b.f["qv"], 0.0)), **tol)
    np.testing.assert_allclose(rad["qc"][0, :, 0, :], grouped(b.f["qcl"]), **tol)
    np.testing.assert_allclose(rad["qi"][0, :, 0, :], grouped(b.f["qci"]), **tol)
    np.testing.assert_allclose(rad["cld"][0, :, 0, :], grouped((b.f["qcl"] + b.f["qci"] > 0) * 1.0), **tol)


GROUPS = {
    "cloud_fractions": (check_cloud_fractions, {}),
    "thin_cloud_top": (check_thin_cloud_top, {}),
    "water_paths": (check_water_paths, {}),
    "cloud_mass_fluxes": (check_cloud_mass_fluxes, {}),
    "entrainment": (check_entrainment, {}),
    "tke_profiles": (check_tke_profiles, {}),
    "flux_profiles": (check_flux_profiles, {}),
    "surface_precipitation": (check_surface_precipitation, {}),
    "radiation_groups": (check_radiation_groups, {"nx_rad": 8}),
}

# The group that must notice each planted mistake.
CAUGHT_BY = {
    "eu_du_swapped": "entrainment",
    "cloud_layers_700_400_swapped": "cloud_fractions",
    "cld_not_time_mean": "cloud_fractions",
    "mc_threshold_wmin": "cloud_mass_fluxes",
    "jt_mx_swapped": "entrainment",
    "water_path_pdel_level": "water_paths",
    "cldtop_every_level": "thin_cloud_top",
    "tkez_no_half": "tke_profiles",
    "scalar_flux_dt_not_dtn": "flux_profiles",
    "updraft_cloud_level_above": "entrainment",
    "rad_qv_unclipped": "radiation_groups",
    "prec_split_per_hour": "surface_precipitation",
}


def _run_group(name):
    fn, kw = GROUPS[name]
    fn(Built(**kw))


@pytest.mark.parametrize("name", GROUPS)
def test_statistics_closed_form(name):
    _run_group(name)


@pytest.mark.parametrize("mutation", CAUGHT_BY)
def test_planted_mistake_is_caught(mutation):
    probe = oracle_lib.StatsProbe()
    probe.mutate(oracle_lib.MUTATIONS[mutation])
    try:
        with pytest.raises(AssertionError):
            _run_group(CAUGHT_BY[mutation])
    finally:
        probe.mutate(0)


def test_every_mutation_has_a_catching_group():
    assert set(CAUGHT_BY) == set(oracle_lib.MUTATIONS)
    assert set(CAUGHT_BY.values()) <= set(GROUPS)
