"""Shared helpers of the parity tests: field comparison and the standard test cases."""
import numpy as np

from crm_inputs import add_convection, make_case

# Fields of crm_state / crm_rad / crm_output compared after a call.
STATE = ["u_wind", "v_wind", "w_wind", "temperature", "qt", "qp", "qn"]
RAD = ["temperature", "qv", "qc", "qi", "cld"]
OUT = ["qcl", "qci", "qpl", "qpi", "tk", "tkh", "prec_crm", "cltot", "clhgh", "clmed", "cllow",
       "precc", "precl", "precsc", "precsl", "jt_crm", "mx_crm", "prectend", "precstend", "z0m",
       "cldtop", "qc_mean", "qi_mean", "qs_mean", "qg_mean", "qr_mean", "sltend", "qltend",
       "qcltend", "qiltend", "cld", "gicewp", "gliqwp", "mctot", "mcup", "mcdn", "mcuup",
       "mcudn", "mu_crm", "md_crm", "du_crm", "eu_crm", "ed_crm", "flux_qt", "fluxsgs_qt",
       "tkez", "tkesgsz", "tkz", "flux_u", "flux_v", "flux_qp", "precflux", "qt_ls",
       "qt_trans", "qp_trans", "qp_fall", "qp_src", "qp_evp", "t_ls"]

# Internal fields readable from both the oracle and the HIP library's debug entry.
GRID = ["u", "w", "t", "tabs", "qv", "qcl", "qci", "qpl", "qpi", "qt", "qp", "qn", "tk", "tkh",
        "tke", "dudt_na", "dudt_nb", "dudt_nc", "dwdt_na", "dwdt_nb", "dwdt_nc", "precsfc",
        "precssfc"]
PROF = ["t0", "q0", "qv0", "tabs0", "u0", "qn0", "qp0", "rho", "rhow", "adz", "adzw", "pres",
        "presi", "z", "zi", "bet", "gamaz", "ttend", "qtend", "utend", "dz"]
# 3D CRMs only (v and its history, means, tendencies and fluxes)
GRID3 = ["v", "dvdt_na", "dvdt_nb", "dvdt_nc"]
PROF3 = ["v0", "vtend", "uwle", "uwsb", "vwle", "vwsb"]


def relerr(a, b):
    """Normwise relative difference max|a-b| / max|b| (absolute when b is all zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return np.inf
    d = np.abs(a - b).max() if a.size else 0.0
    s = np.abs(b).max() if b.size else 0.0
    return d / s if s > 0 else d


# Outputs that are differences of two states divided by dt_gl (crm_module.F90:1109-1251 and
# 572-579): their error is measured against the scale of the states differenced, since
# cancellation makes the difference itself arbitrarily small.  Values: (state field whose
# magnitude sets the scale, factor).
CP = 1.00464e3
G = 9.80616
DIFFERENCE_OUTPUTS = {
    "sltend": ("temperature", CP), "t_ls": ("temperature", 1.0),
    "qltend": ("qt", 1.0), "qcltend": ("qt", 1.0), "qiltend": ("qt", 1.0),
    "qt_ls": ("qt", 1.0),
    "prectend": ("qp", 1.0e5 / G), "precstend": ("qp", 1.0e5 / G),
}


def compare_case(got, want, groups=("state", "rad", "out"), dt_gl=None):
    """{group.field: relerr} between two CrmCase objects over instances 0..ncrms-1.  With
    dt_gl, the outputs listed in DIFFERENCE_OUTPUTS are normalised by the scale of the states
    they difference (e.g. cp * max|T| / dt_gl for sltend)."""
    n = want.ncrms
    errs = {}
    names = {"state": STATE, "rad": RAD, "out": OUT}
    for g in groups:
        gd, wd = getattr(got, g), getattr(want, g)
        for f in names[g]:
            a, b = gd[f][:n], wd[f][:n]
            if g == "out" and dt_gl and f in DIFFERENCE_OUTPUTS:
                field, factor = DIFFERENCE_OUTPUTS[f]
                scale = factor * np.abs(want.state[field][:n]).max() / dt_gl
                scale = max(scale, np.abs(b).max())
                errs[f"{g}.{f}"] = np.abs(a - b).max() / scale if scale > 0 else 0.0
            else:
                errs[f"{g}.{f}"] = relerr(a, b)
    return errs


def convective_case(ncrms=4, nx=32, nz=28, dx=4000.0, dt=20.0, seed=1, **kw):
    """A case whose first substep exercises every process: a raining updraft, ice aloft,
    velocity noise, per-instance randomized soundings."""
    return add_convection(make_case(ncrms, nx, nz, dx, dt, randomize=True, **kw), seed=seed)


def load_golden(path):
    """(case, dt_gl, expected) from a fixture written by tests/golden/make_golden.py; expected
    maps "state.x" / "rad.x" / "out.x" to the reference's arrays after the call."""
    from crm_views import CrmCase
    z = np.load(path)
    m = {k[5:]: z[k].item() for k in z.files if k.startswith("meta.")}
    case = CrmCase(m["ncrms"], m["plev"], m["nx"], m["ny"], m["nz"], m["dx"], m["dt"],
                   nx_rad=m["nx_rad"], ny_rad=m["ny_rad"], ld=m["ld"])
    case.igstep = m["igstep"]
    case.gcol = z["gcol"].astype(np.int32)
    for k in z.files:
        if k.startswith("in:"):
            group, name = k[3:].split(".", 1)
            getattr(case, group)[name][...] = z[k]
    expected = {k[4:]: z[k] for k in z.files if k.startswith("ref:")}
    return case, m["dt_gl"], expected


# ---- column budgets -------------------------------------------------------------------------
LC, LF = 2.501e6, 3.337e5
_F32 = lambda x: float(np.float32(x))  # noqa: E731
_TBG = (_F32(253.16), _F32(273.16))
_TPR = (_F32(268.16), _F32(283.16))


def _levels(field, n, nz):
    """(n, plev) GCM-ordered array -> (n, nz) CRM-ordered (CRM level k = GCM level plev-1-k)."""
    return field[:n, ::-1][:, :nz]


def static_energy(case, gamaz):
    """Liquid-ice static energy / cp of the state (crm_module.F90:508-509 with the
    temperature ramps of micro_diagnose, microphysics.F90:331-357): (n, nx, ny, nz)."""
    n = case.ncrms
    T = case.state["temperature"][:n]
    qn, qp = case.state["qn"][:n], case.state["qp"][:n]
    omn = np.clip((T - _TBG[0]) * (1.0 / (_TBG[1] - _TBG[0])), 0, 1)
    omp = np.clip((T - _TPR[0]) * (1.0 / (_TPR[1] - _TPR[0])), 0, 1)
    fc, fs = LC / CP, (LC + LF) / CP
    return (T + gamaz[:, None, None, :] - fc * (qn * omn + qp * omp)
            - fs * (qn * (1 - omn) + qp * (1 - omp)))


def budget_residuals(before, after, dt_gl, get):
    """Relative residuals of the total-water and static-energy budgets of one crm() call.

    Water:  d/dt sum_k m_k <qt + qp>  =  sum_k m_k qt_ls  -  P
    Energy: d/dt sum_k m_k <t>        =  sum_k m_k t_ls   +  (Lc/cp) P + (Lf/cp) P_snow
    with m_k = rho adz dz, P the surface precipitation (1000 kg/m3 * prec_crm) and P_snow its
    frozen part (precsc + precsl).  Energy uses the prognostic static energy t at the end of
    the call (get("t")): the returned temperature satisfies the saturation adjustment only to
    its 0.01 K stopping tolerance (cloud.F90).  Returns (water, energy) residuals relative to
    the column totals."""
    n, nz, nx = before.ncrms, before.nz, before.nx * before.ny
    rho = get("rho", n * nz).reshape(n, nz)
    adz = get("adz", n * nz).reshape(n, nz)
    dz = get("dz", n)
    gamaz = get("gamaz", n * nz).reshape(n, nz)
    mass = rho * adz * dz[:, None]

    def water(c):
        q = (c.state["qt"][:n] + c.state["qp"][:n]).mean(axis=(1, 2))
        return (mass * q).sum(axis=1)

    P = 1000.0 * after.out["prec_crm"][:n].mean(axis=(1, 2)) * dt_gl
    Ps = 1000.0 * (after.out["precsc"][:n] + after.out["precsl"][:n]) * dt_gl
    Fq = (mass * _levels(after.out["qt_ls"], n, nz)).sum(axis=1) * dt_gl
    Ft = (mass * _levels(after.out["t_ls"], n, nz)).sum(axis=1) * dt_gl
    Wb, Wa = water(before), water(after)
    Hb = (mass * static_energy(before, gamaz).mean(axis=(1, 2))).sum(axis=1)
    Ha = (mass * get("t", n * nz * nx).reshape(n, nz, nx).mean(axis=2)).sum(axis=1)
    rw = np.abs(Wa - Wb - Fq + P).max() / np.abs(Wb).max()
    rh = np.abs(Ha - Hb - Ft - (LC / CP) * P - (LF / CP) * Ps).max() / np.abs(Hb).max()
    return rw, rh
