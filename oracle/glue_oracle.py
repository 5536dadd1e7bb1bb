"""CPU restatement of the GCM-side code around crm(), the checker of the library's glue kernels
(include/crm_glue.h).  TEST INFRASTRUCTURE ONLY: imported by tests/ and nothing else.

Arrays are numpy arrays in the physics types' Fortran layout: (pcols, pver) and (pcols,
pver+1) with order="F", per-column arrays (pcols).  Only columns 0..ncol-1 are touched.  Level
sums run in the reference's level order, one array operation per level, so every value is the
reference's expression in the reference's evaluation order.
"""
import numpy as np

# physconst via shr_const_mod (cime/src/share/util/shr_const_mod.F90:22,47,53,54)
GRAVIT = 9.80616
CPAIR = 1.00464e3
LATVAP = 2.501e6
LATICE = 3.337e5
C061 = float(np.float32(0.61))  # default-real literal of crm_physics.F90:1092

INPUT_FROM_STATE = {"zmid": "zm", "zint": "zi", "tl": "t", "ql": "qv", "qccl": "qcl", "qiil": "qci",
                    "ps": "ps", "pmid": "pmid", "pint": "pint", "pdel": "pdel", "phis": "phis"}


def gcm_input(ncol, crm_nz, state, srf, angle, inp, qrad):
    """crm_physics.F90:1005-1010 (qrad from Q dp back to a heating rate, in place),
    1050-1098 (crm_input; writes inp, leading dimension pcols) and the wind rotation of
    :1120-1125."""
    n = slice(0, ncol)
    pver = state["pdel"].shape[1]
    for m in range(crm_nz):
        qrad[n, :, :, m] = qrad[n, :, :, m] / state["pdel"][n, pver - 1 - m][:, None, None]
    for dst, src in INPUT_FROM_STATE.items():
        inp[dst][n] = state[src][n]
    inp["ocnfrac"][n] = srf["ocnfrac"][n]
    wsx, wsy, shf, lhf = (srf[k][n] for k in ("wsx", "wsy", "shf", "lhf"))
    tbot = state["t"][n, -1]
    inp["tau00"][n] = np.sqrt(wsx * wsx + wsy * wsy)                       # :1091
    inp["bflxls"][n] = shf / CPAIR + C061 * tbot * lhf / LATVAP            # :1092
    inp["fluxu00"][n] = wsx                                                # :1093-1096
    inp["fluxv00"][n] = wsy
    inp["fluxt00"][n] = shf / CPAIR
    inp["fluxq00"][n] = lhf / LATVAP
    ubot, vbot = state["u"][n, -1], state["v"][n, -1]
    inp["wndls"][n] = np.sqrt(ubot * ubot + vbot * vbot)                   # :1098
    ca, sa = np.cos(angle[n])[:, None], np.sin(angle[n])[:, None]
    u, v = state["u"][n], state["v"][n]
    inp["ul"][n] = u * ca + v * sa                                         # :1122
    inp["vl"][n] = v * ca - u * sa                                         # :1123
    return inp


def flux_bypass(ncol, pdel, shf, cflx, ptend_s, ptend_q):
    """crm_surface_flux_bypass_tend, crm_physics.F90:1690-1703 (rpdel = 1/pdel)."""
    n = slice(0, ncol)
    g_dp = GRAVIT * (1.0 / pdel[n, -1])
    ptend_s[n] = 0.0
    ptend_q[n] = 0.0
    ptend_s[n, -1] = g_dp * shf[n]
    ptend_q[n, -1] = g_dp * cflx[n]


def gcm_tend(ncol, crm_nz, pdel, out, qrad, qrs, qrl, ptend):
    """crm_physics.F90:1158-1169 (tendencies, precipitation merge), 1203-1209 (qrad to Q dp),
    1310-1330 (radiative heating above the CRM, sp_rad_flux).  qrad: (pcols, nx_rad, ny_rad,
    crm_nz); qrs, qrl changed in place; ptend: dict s, qv, qcl, qci.  Returns sp_rad_flux."""
    n = slice(0, ncol)
    pver = pdel.shape[1]
    ptend["qv"][n] = out["qltend"][n]
    ptend["qcl"][n] = out["qcltend"][n]
    ptend["qci"][n] = out["qiltend"][n]
    ptend["s"][n] = out["sltend"][n]
    out["precc"][n] = out["precc"][n] + out["precl"][n]
    out["precsc"][n] = out["precsc"][n] + out["precsl"][n]
    out["precl"][n] = 0.0
    out["precsl"][n] = 0.0
    for m in range(crm_nz):
        qrad[n, :, :, m] = qrad[n, :, :, m] * pdel[n, pver - 1 - m][:, None, None]
    qrs[n] = qrs[n] / pdel[n]
    qrl[n] = qrl[n] / pdel[n]
    top = pver - crm_nz + 2
    ptend["s"][n, :top] = qrs[n, :top] + qrl[n, :top]
    f = np.zeros(ncol)
    for k in range(pver):
        f = f + (qrs[n, k] + qrl[n, k]) * pdel[n, k] / GRAVIT
    return f


def incloud_path(q, pdel, cld):
    """radiation.F90:1601-1604: in-cloud water path [kg/m2] of a level, the cover floored at 1%."""
    return q * pdel / GRAVIT / max(0.01, cld)


def rad_columns(ncol, crm_nz, state, cloud, rad):
    """The radiation's inputs in every CRM radiation column (rrtmg/radiation.F90:1584-1636,
    sam1mom): for radiation column (ii, jj), r = ii + nx_rad jj, the physics state and the GCM
    cloud fields with the CRM levels (GCM level k = pver-1-m of CRM level m) replaced from the
    crm_rad averages.  state: dict t, qv, qcl, qci, pdel (pcols, pver); cloud: dict cld, fice,
    cicewp, cliqwp (pcols, pver); rad: dict temperature, qv, qc, qi, cld (pcols, nx_rad,
    ny_rad, crm_nz).  Returns dict t, qv, qc, qi, cld, fice, cicewp, cliqwp of shape
    (pcols, nrad, pver) (columns past ncol zero)."""
    pcols, pver = state["t"].shape
    nxr, nyr = rad["cld"].shape[1:3]
    nrad = nxr * nyr
    src = {"t": state["t"], "qv": state["qv"], "qc": state["qcl"], "qi": state["qci"],
           "cld": cloud["cld"], "fice": cloud["fice"], "cicewp": cloud["cicewp"], "cliqwp": cloud["cliqwp"]}
    out = {f: np.zeros((pcols, nrad, pver), order="F") for f in src}
    for jj in range(nyr):
        for ii in range(nxr):
            r = ii + nxr * jj
            for f in src:
                out[f][:ncol, r, :] = src[f][:ncol]
            for m in range(crm_nz):
                k = pver - 1 - m
                for i in range(ncol):
                    cld = rad["cld"][i, ii, jj, m]
                    qc, qi = rad["qc"][i, ii, jj, m], rad["qi"][i, ii, jj, m]
                    out["cld"][i, r, k] = cld                                             # :1595
                    if cld > 0:                                                           # :1598-1604
                        qtot = qc + qi
                        out["fice"][i, r, k] = qi / qtot
                        out["cicewp"][i, r, k] = incloud_path(qi, state["pdel"][i, k], cld)
                        out["cliqwp"][i, r, k] = incloud_path(qc, state["pdel"][i, k], cld)
                    else:                                                                 # :1605-1609
                        out["fice"][i, r, k] = 0.0
                        out["cicewp"][i, r, k] = 0.0
                        out["cliqwp"][i, r, k] = 0.0
                    out["qi"][i, r, k] = qi                                               # :1627-1630
                    out["qc"][i, r, k] = qc
                    out["qv"][i, r, k] = max(1.e-9, rad["qv"][i, ii, jj, m])
                    out["t"][i, r, k] = rad["temperature"][i, ii, jj, m]
    return out


def rad_qrad(ncol, qrs_crm, qrl_crm, qrad):
    """radiation.F90:2393-2398: the CRM's radiative heating, shortwave plus longwave, per
    radiation column and CRM level (pcols, nx_rad, ny_rad, crm_nz), written into qrad."""
    qrad[:ncol] = qrs_crm[:ncol] + qrl_crm[:ncol]
    return qrad


def energy_integrals(ncol, state):
    """Column integrals of check_energy.F90:440-484 (no rain or snow constituents)."""
    n = slice(0, ncol)
    ke = np.zeros(ncol)
    se = np.zeros(ncol)
    wv = np.zeros(ncol)
    wl = np.zeros(ncol)
    wi = np.zeros(ncol)
    for k in range(state["t"].shape[1]):
        u, v, dp = state["u"][n, k], state["v"][n, k], state["pdel"][n, k]
        ke = ke + 0.5 * (u * u + v * v) * dp / GRAVIT
        se = se + state["t"][n, k] * CPAIR * dp / GRAVIT
        wv = wv + state["qv"][n, k] * dp / GRAVIT
    se = se + state["phis"][n] * state["ps"][n] / GRAVIT
    for k in range(state["t"].shape[1]):
        wl = wl + state["qcl"][n, k] * state["pdel"][n, k] / GRAVIT
        wi = wi + state["qci"][n, k] * state["pdel"][n, k] / GRAVIT
    te = se + ke + (LATVAP + LATICE) * wv + LATICE * (wl + 0.0)
    tw = wv + wl + wi + 0.0 + 0.0
    return te, tw


def check_energy_chng(ncol, state, ztodt, flx_vap, flx_cnd, flx_ice, flx_sen, te_cur, tw_cur,
                      te_tnd_c, tw_tnd_c):
    """check_energy.F90:486-558; updates te_cur, tw_cur, te_tnd_c, tw_tnd_c in place and
    returns (te_rer, tw_rer)."""
    n = slice(0, ncol)
    te, tw = energy_integrals(ncol, state)
    te_tnd = flx_vap[n] * (LATVAP + LATICE) - (flx_cnd[n] - flx_ice[n]) * 1000.0 * LATICE + flx_sen[n]
    tw_tnd = flx_vap[n] - flx_cnd[n] * 1000.0
    te_tnd_c[n] = te_tnd_c[n] + te_tnd
    tw_tnd_c[n] = tw_tnd_c[n] + tw_tnd
    te_xpd = te_cur[n] + te_tnd * ztodt
    tw_xpd = tw_cur[n] + tw_tnd * ztodt
    with np.errstate(divide="ignore", invalid="ignore"):
        te_rer = (te_xpd - te) / te_cur[n]
        tw_rer = np.where(tw_cur[n] > 0.0, (tw_xpd - tw) / tw_cur[n], 0.0)
    te_cur[n] = te
    tw_cur[n] = tw
    return te_rer, tw_rer
