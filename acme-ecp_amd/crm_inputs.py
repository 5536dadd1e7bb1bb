"""Synthetic GCM columns and CRM states for the benchmark workloads and the parity tests.

Workloads (all 2D unless stated):
  * tropical sounding: SST, T lapse 6.5 K/km to a 15 km tropopause then isothermal, RH
    falling linearly from its surface value to 20% at 12 km (20% above), hydrostatic
    pressure from 1000 hPa using the dry temperature, u = shear * z / 10 km, v = 0.
  * surface fluxes reach the CRM the way the caller's flux bypass delivers them
    (cam/crm_physics.F90:1697-1701): the lowest GCM level of tl/ql is incremented by
    g*SH*dt_gl/(cp*pdel) and g*(LH/Lv)*dt_gl/pdel; no other large-scale forcing, qrad = 0.
  * noise: uniform +/- amp K on the CRM temperature in the lowest 5 levels, seeded per
    instance (numpy PCG64, seed = instance id + seed_base).

The GCM column has plev = nz + 2 levels (CAM ordering, k=0 at the top); CRM level k sits at
GCM level plev-1-k (crm_module.F90:340).
"""
import numpy as np

from crm_views import CrmCase

# Physical constants as the CRM sees them under -DCRM (crm/params.F90:49-55 from
# cime/src/share/util/shr_const_mod.F90:22-55).
G = 9.80616
CP = 1.00464e3
LV = 2.501e6
RGAS = 6.02214e26 * 1.38065e-23 / 28.966


def _f(x):
    return float(np.float32(x))


_ESW = [_f(c) for c in (6.105851, 0.4440316, 0.1430341e-1, 0.2641412e-3, 0.2995057e-5,
                        0.2031998e-7, 0.6936113e-10, 0.2564861e-13, -0.3704404e-15)]


def qsatw(t, p_mb):
    """Saturation mixing ratio over water (Flatau et al. 1992 polynomial, crm/sat.F90:12-41)."""
    t = np.asarray(t, dtype=np.float64)
    dt = np.maximum(t - _f(273.16), -80.0)
    e = _ESW[8]
    for c in reversed(_ESW[:8]):
        e = c + dt * e
    return _f(0.622) * e / np.maximum(e, p_mb - e)


def stretched_interfaces(n, dz_bottom, dz_top_max, z_top):
    """n+1 interface heights from 0 to z_top; layer thickness grows geometrically from
    dz_bottom and is capped at dz_top_max.  The ratio is found by bisection."""
    def top(r):
        dz = np.minimum(dz_bottom * r ** np.arange(n), dz_top_max)
        return dz.sum(), dz
    lo, hi = 1.0, 3.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if top(mid)[0] < z_top:
            lo = mid
        else:
            hi = mid
    dz = top(hi)[1]
    dz *= z_top / dz.sum()
    return np.concatenate([[0.0], np.cumsum(dz)])


# Vertical grids of the size classes: (CRM levels) -> (dz_bottom, dz_max, CRM top [m]).
GRIDS = {28: (80.0, 1600.0, 24000.0), 58: (50.0, 1000.0, 40000.0)}


def sounding_profiles(zmid, zint, sst, rh_sfc, shear):
    """T, qv, u at mid levels and p at interfaces/mid levels for heights above the surface."""
    gam = 6.5e-3
    ztp = 15000.0
    t_tp = sst - gam * ztp
    expo = G / (RGAS * gam)

    def temp(z):
        return np.where(z <= ztp, sst - gam * z, t_tp)

    def pres(z):
        p_tp = 1.0e5 * (t_tp / sst) ** expo
        return np.where(z <= ztp, 1.0e5 * (temp(z) / sst) ** expo,
                        p_tp * np.exp(-G * (z - ztp) / (RGAS * t_tp)))

    tm = temp(zmid)
    pm = pres(zmid)
    pi = pres(zint)
    rh = np.where(zmid <= 12000.0, rh_sfc + (0.2 - rh_sfc) * zmid / 12000.0, 0.2)
    qv = rh * qsatw(tm, pm / 100.0)
    u = shear * zmid / 10000.0
    return tm, qv, u, pm, pi


def make_case(ncrms, nx, nz, dx, dt, ny=1, dt_gl=1200.0, sst=300.0, rh_sfc=0.8, shear=5.0,
              randomize=False, seed_base=0, noise_amp=0.1, sh=15.0, lh=100.0, ld=None,
              nx_rad=None, ny_rad=None):
    """Build a CrmCase.  randomize=True draws SST~U[290,304], RH_sfc~U[0.6,0.9] and
    shear~U[0,10] m/s per 10 km per instance (seed = instance id + seed_base)."""
    plev = nz + 2
    c = CrmCase(ncrms, plev, nx, ny, nz, dx, dt, nx_rad=nx_rad, ny_rad=ny_rad, ld=ld)
    dzb, dzm, ztop = GRIDS.get(nz, (80.0, 1600.0, 24000.0 * nz / 28.0))
    zi_crm = stretched_interfaces(nz, dzb, dzm, ztop)
    dz_last = zi_crm[-1] - zi_crm[-2]
    zi_col = np.concatenate([zi_crm, zi_crm[-1] + dz_last * np.array([1.0, 2.0])])
    zm_col = 0.5 * (zi_col[1:] + zi_col[:-1])
    # CAM ordering: index 0 is the top
    zint = zi_col[::-1].copy()
    zmid = zm_col[::-1].copy()
    for n in range(ncrms):
        rng = np.random.default_rng(n + seed_base)
        if randomize:
            s = rng.uniform(290.0, 304.0)
            r = rng.uniform(0.6, 0.9)
            sh_ = rng.uniform(0.0, 10.0)
        else:
            s, r, sh_ = sst, rh_sfc, shear
        tm, qv, u, pm, pi = sounding_profiles(zmid, zint, s, r, sh_)
        c.inp["zmid"][n, :] = zmid
        c.inp["zint"][n, :] = zint
        c.inp["pmid"][n, :] = pm
        c.inp["pint"][n, :] = pi
        c.inp["pdel"][n, :] = pi[1:] - pi[:-1]
        c.inp["tl"][n, :] = tm
        c.inp["ql"][n, :] = qv
        c.inp["ul"][n, :] = u
        if ny > 1:  # 3D: a cross wind of half the along-x shear
            c.inp["vl"][n, :] = 0.5 * u
        c.inp["ps"][n] = pi[-1]
        c.inp["ocnfrac"][n] = 1.0
        c.inp["wndls"][n] = abs(u[-1])
        c.inp["bflxls"][n] = sh / CP + 0.61 * tm[-1] * lh / LV
        pd = c.inp["pdel"][n, -1]
        c.inp["tl"][n, -1] += G * sh * dt_gl / (CP * pd)
        c.inp["ql"][n, -1] += G * (lh / LV) * dt_gl / pd
        # CRM state: CRM level k <-> GCM level plev-1-k
        col = slice(plev - 1, plev - 1 - nz, -1)
        tcrm, qcrm, ucrm = tm[col], qv[col], u[col]
        noise = rng.uniform(-noise_amp, noise_amp, size=(5, ny, nx))
        for k in range(nz):
            c.state["temperature"][n, :, :, k] = tcrm[k]
            c.state["qt"][n, :, :, k] = qcrm[k]
            c.state["u_wind"][n, :, :, k] = ucrm[k]
            if ny > 1:
                c.state["v_wind"][n, :, :, k] = 0.5 * ucrm[k]
        for k in range(min(5, nz)):
            c.state["temperature"][n, :, :, k] += noise[k].T
    return c


def add_convection(c, seed=1):
    """Make the state exercise every microphysics branch: a moist, cloudy and raining plume
    with an updraft, an ice cloud aloft, and random velocity noise."""
    rng = np.random.default_rng(seed)
    nz, nx = c.nz, c.nx
    for n in range(c.ncrms):
        p_mb = c.inp["pmid"][n, ::-1][:nz] / 100.0
        t = c.state["temperature"][n, :, 0, :]
        i0 = (3 * n) % nx
        cols = [(i0 + d) % nx for d in range(max(3, nx // 6))]
        klo, khi = max(1, nz // 12), max(3, nz // 2)
        # three draws per (column, level) in that order: condensate, precipitation, updraft
        r = rng.random((len(cols), khi - klo, 3))
        for j, i in enumerate(cols):
            qs = qsatw(t[i, klo:khi], p_mb[klo:khi])
            qn = 1.5e-3 * (1.0 + 0.3 * r[j, :, 0])
            c.state["qn"][n, i, 0, klo:khi] = qn
            c.state["qt"][n, i, 0, klo:khi] = qs * 1.01 + qn
            c.state["qp"][n, i, 0, klo:khi] = 2.0e-3 * r[j, :, 1]
            c.state["w_wind"][n, i, 0, klo:khi] = 3.0 + 2.0 * r[j, :, 2]
        # ice cloud aloft (T below 253 K) and falling snow
        kice = [k for k in range(nz) if t[0, k] < 245.0]
        for k in kice[: max(1, len(kice) // 3)]:
            for i in range(nx // 2, nx // 2 + max(2, nx // 8)):
                c.state["qn"][n, i, 0, k] = 2.0e-4
                c.state["qt"][n, i, 0, k] += 2.0e-4
                c.state["qp"][n, i, 0, k] = 5.0e-4
        c.state["u_wind"][n] += rng.uniform(-1.0, 1.0, size=c.state["u_wind"][n].shape)
        c.state["w_wind"][n, :, :, 1:] += rng.uniform(-0.5, 0.5, size=(nx, c.ny, nz - 1))
        if c.ny > 1:
            # 3D: the plume and the ice cloud span a few rows in y, and v gets noise too
            for j in range(1, max(2, c.ny // 3)):
                for f in ("qn", "qt", "qp", "w_wind"):
                    c.state[f][n, :, j, :] = c.state[f][n, :, 0, :]
            c.state["v_wind"][n] += rng.uniform(-1.0, 1.0, size=c.state["v_wind"][n].shape)
    return c
