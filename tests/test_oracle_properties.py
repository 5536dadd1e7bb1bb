"""Checks of the oracle that do not need the reference: closed forms, textbook algorithms
and invariants of the equations, each chosen so that a plausible restatement mistake (a
dropped term, a wrong sign, an index off by one, a transposed operand) fails it."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from crm_inputs import make_case
from parity import convective_case

dp = C.POINTER(C.c_double)


def _lib():
    return oracle_lib.oracle().lib


# ---- x-transform (crm/fft.F90:5-153) ---------------------------------------------------------
@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_fft991_is_the_scaled_real_dft(n):
    """Forward coefficients are a_m + i b_m = (1/n) sum_j x_j exp(-2 pi i j m / n) stored as
    (a0, b0, a1, b1, ..., a_{n/2}, b_{n/2}); the REAL(4) twiddles of the reference limit the
    agreement to single precision, which this test also pins (it is not double)."""
    lib = _lib()
    lib.oracle_fft991.argtypes = [dp, C.c_int, C.c_int]
    x = np.random.default_rng(n).standard_normal(n)
    a = np.zeros(n + 2)
    a[:n] = x
    lib.oracle_fft991(a.ctypes.data_as(dp), n, -1)
    X = np.fft.rfft(x) / n
    want = np.empty(n + 2)
    want[0::2] = X.real
    want[1::2] = X.imag
    err = np.abs(a - want).max() / np.abs(want).max()
    assert err < 1e-6
    assert err > 1e-13  # single-precision twiddles, as in the reference
    # inverse: x_j = sum over the stored coefficients (a0 + 2 sum Re(...) + a_{n/2} cos)
    b = a.copy()
    lib.oracle_fft991(b.ctypes.data_as(dp), n, 1)
    assert np.abs(b[:n] - x).max() < 1e-5


# ---- Mersenne twister (cam/RNG_MT.F90:60-139) -------------------------------------------------
def _mt_restated(seed, count, umask):
    """MT19937 of Matsumoto & Nishimura (1998) with the 69069 seeding of RNG_MT_set_seed
    and a configurable upper mask."""
    N, M = 624, 397
    mt = [0] * N
    mt[0] = seed & 0xFFFFFFFF
    for i in range(1, N):
        mt[i] = (69069 * mt[i - 1]) & 0xFFFFFFFF
    out, idx = [], N
    for _ in range(count):
        if idx >= N:
            for kk in range(N):
                y = (mt[kk] & umask) | (mt[(kk + 1) % N] & 0x7FFFFFFF)
                mt[kk] = mt[(kk + M) % N] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        out.append(y & 0xFFFFFFFF)
    return out, mt


def test_mt_restatement_is_mt19937():
    """With the standard upper mask the restatement is numpy's MT19937 given the same state."""
    seed = 1234
    mt0 = [seed]
    for i in range(1, 624):
        mt0.append((69069 * mt0[-1]) & 0xFFFFFFFF)
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937",
                "state": {"key": np.array(mt0, dtype=np.uint32), "pos": 624}}
    want = bg.random_raw(1500)
    got, _ = _mt_restated(seed, 1500, 0x80000000)
    assert np.array_equal(np.array(got, dtype=np.uint64), want.astype(np.uint64))


def test_oracle_generator_uses_reference_mask():
    """RNG_MT.F90:61 sets UMASK = -1147483648 (not -2147483648); outputs are mapped to (0,1)
    as y/2^32 for y < 0 (signed) and (y + 0.5)/2^32 otherwise (RNG_MT.F90:131-136)."""
    lib = _lib()
    lib.oracle_mt_seed.argtypes = [C.c_int]
    lib.oracle_mt_rand.restype = C.c_double
    seed = 17000
    lib.oracle_mt_seed(seed)
    got = np.array([lib.oracle_mt_rand() for _ in range(1300)])
    raw, _ = _mt_restated(seed, 1300, (-1147483648) & 0xFFFFFFFF)
    want = []
    for y in raw:
        ys = y - (1 << 32) if y >= (1 << 31) else y
        want.append((ys + 2.0 ** 32) / 2.0 ** 32 if ys < 0 else (ys + 0.5) / 2.0 ** 32)
    assert np.array_equal(got, np.array(want))
    std, _ = _mt_restated(seed, 1300, 0x80000000)
    assert raw != std  # the mask matters


# ---- saturation (crm/sat.F90:12-139) ----------------------------------------------------------
ESATW = (6.105851, 0.4440316, 0.1430341e-1, 0.2641412e-3, 0.2995057e-5, 0.2031998e-7,
         0.6936113e-10, 0.2564861e-13, -0.3704404e-15)
ESATI = (6.11147274, 0.503160820, 0.188439774e-1, 0.420895665e-3, 0.615021634e-5,
         0.602588177e-7, 0.385852041e-9, 0.146898966e-11, 0.252751365e-14)


def _poly(coefs, dt):
    c = [float(np.float32(v)) for v in coefs]
    e = c[8]
    for v in reversed(c[:8]):
        e = v + dt * e
    return e


@pytest.mark.parametrize("t", [320.0, 300.0, 273.16, 250.0, 210.0, 194.0])
def test_saturation_polynomials(t):
    """Flatau et al. (1992) 8th-order fits with REAL(4) coefficients, in mb; the Murphy-Koop
    form below -80 C; qsat = 0.622 e / max(e, p - e)."""
    lib = _lib()
    for f in ("oracle_esatw", "oracle_esati"):
        getattr(lib, f).restype = C.c_double
        getattr(lib, f).argtypes = [C.c_double]
    lib.oracle_qsatw.restype = C.c_double
    lib.oracle_qsatw.argtypes = [C.c_double, C.c_double]
    r4 = lambda v: float(np.float32(v))
    dt = t - r4(273.16)
    if dt > -80.0:
        ew, ei = _poly(ESATW, dt), _poly(ESATI, dt)
    else:
        low = math.exp(r4(9.550426) - r4(5723.265) / t + r4(3.53068) * math.log(t)
                       - r4(0.00728332) * t)
        ew, ei = r4(0.01) * 2.0 * low, r4(0.01) * low
    assert lib.oracle_esatw(t) == pytest.approx(ew, rel=1e-15)
    assert lib.oracle_esati(t) == pytest.approx(ei, rel=1e-15)
    # Clausius-Clapeyron sanity: within 2% of the textbook fit near the triple point
    if 250 < t < 310:
        bolton = 6.112 * math.exp(17.67 * (t - 273.15) / (t - 29.65))
        assert lib.oracle_esatw(t) == pytest.approx(bolton, rel=2e-2)
    p = 850.0
    assert lib.oracle_qsatw(t, p) == pytest.approx(r4(0.622) * ew / max(ew, p - ew), rel=1e-15)


def test_gamma_values():
    """precip_init.F90:24-45 uses gammafff (crm/gammafff.c:13) = REAL(4) exp(lgamma(x))."""
    for x in (3.8, 2.9, 4.8, 3.25, 2.625, 4.25, 3.5, 2.75, 4.5):
        g = float(np.float32(math.exp(math.lgamma(float(np.float32(x))))))
        assert g == pytest.approx(math.gamma(x), rel=1e-6)


# ---- invariants of the equations --------------------------------------------------------------
def _uniform_case(n=2):
    c = make_case(n, 32, 28, 4000.0, 20.0, noise_amp=0.0, shear=0.0, sh=0.0, lh=0.0)
    return c


def test_rest_state_stays_horizontally_uniform():
    """No noise, no wind, no surface fluxes: every column stays identical (x symmetry)."""
    c = _uniform_case()
    assert oracle_lib.oracle().run(c, 200.0) == 0
    for f in ("temperature", "qt", "u_wind", "w_wind"):
        a = c.state[f][: c.ncrms, :, 0, :]
        spread = np.abs(a - a[:, :1, :]).max()
        assert spread <= 1e-12 * max(1.0, np.abs(a).max()), (f, spread)


def test_buoyancy_of_uniform_state_and_warm_bubble():
    """buoyancy.F90:7-38 uses departures from the level means: round-off for a uniform
    state; a +1 K anomaly in one cell lifts the two w levels bounding it, by about
    g dT / T (1 - 1/nx) in total (each interface takes its adz-weighted share)."""
    c = _uniform_case()
    o = oracle_lib.oracle()
    n = c.ncrms * (c.nz + 1) * c.nx
    assert o.run(c.copy(), c.dt, 2) == 0
    assert np.abs(o.field("dwdt_na", n)).max() < 1e-12
    i0, k0, dT = 5, 10, 1.0
    c.state["temperature"][0, i0, 0, k0] += dT
    assert o.run(c, c.dt, 2) == 0
    b = o.field("dwdt_na", n).reshape(c.ncrms, c.nz + 1, c.nx)
    T = c.state["temperature"][0, 0, 0, k0]
    total = b[0, k0, i0] + b[0, k0 + 1, i0]
    assert b[0, k0, i0] > 0 and b[0, k0 + 1, i0] > 0
    assert total == pytest.approx(9.80616 * dT / T * (1 - 1 / c.nx), rel=0.05)
    # other columns feel the lowered level mean, with the opposite sign
    assert b[0, k0, (i0 + 3) % c.nx] < 0


@pytest.mark.parametrize("m,tol", [(16, 1e-12), (1, 1e-5), (5, 1e-5)])
def test_shift_equivariance(m, tol):
    """Every process is invariant under a periodic shift in x, so shifting the initial state
    shifts the result.  A half-domain shift commutes exactly with the reference's transform
    (its twiddles are then +-1); other shifts only to the single precision of its REAL(4)
    twiddles (fft.F90:247-295)."""
    c = make_case(2, 32, 28, 4000.0, 20.0, randomize=True)
    rng = np.random.default_rng(3)
    c.state["u_wind"][...] += rng.uniform(-2, 2, c.state["u_wind"].shape)
    c.state["w_wind"][:, :, :, 1:] += rng.uniform(-0.5, 0.5, c.state["w_wind"][:, :, :, 1:].shape)
    s = c.copy()
    for k, v in s.state.items():
        v[...] = np.roll(c.state[k], m, axis=1)
    o = oracle_lib.oracle()
    assert o.run(c, 100.0) == 0
    assert o.run(s, 100.0) == 0
    for k in ("temperature", "qt", "u_wind", "w_wind"):
        a = np.roll(c.state[k], m, axis=1)
        assert np.abs(a - s.state[k]).max() <= tol * np.abs(a).max(), k


def _divergence(o, c, u, w):
    """Discrete divergence of press_rhs.F90:91-99 on the C grid (periodic x)."""
    nx, nz, nzm = c.nx, c.nz + 1, c.nz
    rho = o.field("rho", c.ncrms * nzm).reshape(c.ncrms, nzm)
    rhow = o.field("rhow", c.ncrms * nz).reshape(c.ncrms, nz)
    adz = o.field("adz", c.ncrms * nzm).reshape(c.ncrms, nzm)
    dz = o.field("dz", c.ncrms)
    u = u.reshape(c.ncrms, nzm, nx)
    w = w.reshape(c.ncrms, nz, nx)
    dudx = (np.roll(u, -1, axis=2) - u) / c.dx
    dwdz = (rhow[:, 1:, None] * w[:, 1:, :] - rhow[:, :-1, None] * w[:, :-1, :]) / (
        rho[:, :, None] * adz[:, :, None] * dz[:, None, None])
    return dudx + dwdz, np.abs(dudx).max()


def test_pressure_projection_removes_divergence():
    """After pressure + adams (crm_module.F90:829-834) the new velocities u^{n+1}, w^{n+1}
    (held in dudt(nc), dwdt(nc)) are discretely divergence-free: the divergence drops by more
    than six orders of magnitude from that of the velocities entering the substep (the
    REAL(4) twiddles of the reference's transform bound it near 1e-7)."""
    c = convective_case(2, seed=6)
    o = oracle_lib.oracle()
    nx, nz, nzm = c.nx, c.nz + 1, c.nz
    assert o.run(c.copy(), c.dt, 11) == 0
    before, _ = _divergence(o, c, o.field("u", c.ncrms * nzm * nx), o.field("w", c.ncrms * nz * nx))
    assert o.run(c.copy(), c.dt, 13) == 0
    after, _ = _divergence(o, c, o.field("dudt_nc", c.ncrms * nzm * nx),
                           o.field("dwdt_nc", c.ncrms * nz * nx))
    assert np.abs(after).max() < 1e-6 * np.abs(before).max()


def test_scalar_advection_conserves_total_water():
    """MPDATA (advect_scalar2D.F90) is flux-form with no flux through the surface or lid, so
    the mass-weighted domain total of qt is unchanged by advect_all_scalars (qt does not
    sediment)."""
    c = convective_case(2, seed=8)
    o = oracle_lib.oracle()
    nx, nzm = c.nx, c.nz
    assert o.run(c.copy(), c.dt, 14) == 0
    rho = o.field("rho", c.ncrms * nzm).reshape(c.ncrms, nzm, 1)
    adz = o.field("adz", c.ncrms * nzm).reshape(c.ncrms, nzm, 1)
    before = o.field("qt", c.ncrms * nzm * nx).reshape(c.ncrms, nzm, nx)
    assert o.run(c.copy(), c.dt, 15) == 0
    after = o.field("qt", c.ncrms * nzm * nx).reshape(c.ncrms, nzm, nx)
    tb = (rho * adz * before).sum(axis=(1, 2))
    ta = (rho * adz * after).sum(axis=(1, 2))
    assert np.abs(ta - tb).max() < 1e-13 * np.abs(tb).max()
    assert after.min() >= 0.0
