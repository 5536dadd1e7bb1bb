// crm_kernels.hpp -- the CRM substep as HIP kernels for gfx950 (FP64 form).
//
// Each kernel is one phase group of the reference substep (crm/crm_module.F90:743-888) done
// as a single sweep over levels: lane = column, 64/NX instances per wave64, neighbours in x
// by lane rotation, neighbours in z by a register window.  Level index k is 0-based here
// (reference level k+1).
#pragma once
#include "../../include/crm_hip.h"
#include "crm_common.hpp"
#include "crm_mlp_dev.hpp"
#include <type_traits>

namespace CRM_NS {

// Launch bounds of the segment kernels: 256 threads, and at least CRM_WAVES_PER_EU waves per
// SIMD resident: the level sweeps are latency bound, and a second wave hides more than the
// few spilled registers cost where the FP64 sweeps exceed the 256-register file of two waves.
#ifndef CRM_WAVES_PER_EU
#define CRM_WAVES_PER_EU 2
#endif
#define CRM_LB __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CRM_WAVES_PER_EU)))

// A captured call runs every substep with one subcycle; k_decide sets gmax[GMAX_STOP] to the
// first substep that needs more, and every kernel of the call then returns at entry (the flag
// is the same for every thread, so no workgroup barrier is split); the host resumes that
// substep eagerly.  d.stop is null outside such a call.
#define CRM_STOP_CHECK(d) \
  if ((d).stop && *(const volatile unsigned long long*)(d).stop) return

// Clears KURANT_SLOTS slots of the batch-global maxima (a kernel rather than a memset, so that
// it can sit in a captured graph); launched with at least KURANT_SLOTS threads.
__global__ void k_clear_u64(unsigned long long* p, const unsigned long long* stop) {
  if (stop && *(const volatile unsigned long long*)stop) return;
  if (threadIdx.x < KURANT_SLOTS) p[threadIdx.x * KURANT_STRIDE] = 0ull;
}

// Zeroes n 8-byte words (a kernel rather than a memset, so that the call may sit in a graph a
// caller captures); any grid, grid-stride.
__global__ void k_zero_u64(unsigned long long* __restrict__ p, size_t n) {
  for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (size_t)gridDim.x * blockDim.x)
    p[j] = 0ull;
}

// The subcycle decision of kurant.F90:53-75 on the device, for a captured call: one subcycle
// when cfl / 0.7 <= 1; anything else (more subcycles, the reference's abort, a non-finite
// value) stops the captured call at this substep for the host to take over; so does
// nstep == stop_at (a test hook, 0 when off).
__global__ void k_decide(unsigned long long* gmax, int nstep, int stop_at) {
  if (threadIdx.x != 0 || gmax[GMAX_STOP]) return;
  unsigned long long cbits = 0;
  for (int j = 0; j < KURANT_SLOTS; j++) cbits = ::max(cbits, gmax[GMAX_FLAGS + j * KURANT_STRIDE]);
  const double q = __longlong_as_double((long long)cbits) / (double)0.7f;
  if (!(q <= 1.0) || nstep == stop_at) gmax[GMAX_STOP] = (unsigned long long)nstep;
}

// ------------------------------------------------------------------------------------------
// Reference profiles from crm_input: crm_module.F90:320-375; grid factors sgs.F90:219-228;
// sponge profile damping.F90:44-61; roughness crm_module.F90:583-590 (z0_est
// crmsurface.F90:104-131).  One thread per instance.
// ------------------------------------------------------------------------------------------
__device__ double z0_est(double z, double bflx, double wnd, double ustar, double c1) {
  const double vonk = R4(0.4), am = R4(4.8), bm = R4(19.3), eps = R4(1.e-10);
  double rlmo = -bflx * vonk / (ustar * ustar * ustar + eps);
  double zeta = fmin(1.0, z * rlmo), psi1;
  if (zeta >= 0.) {
    psi1 = -am * zeta;
  } else {
    double x = sqrt(sqrt(1.0 - bm * zeta));
    psi1 = 2. * log(1.0 + x) + log(1.0 + x * x) - 2. * atan(x) + c1;
  }
  double lnz = fmax(0.0, vonk * wnd / (ustar + eps) + psi1);
  return z * exp(-lnz);
}

template <typename R>
__global__ void k_setup(DevT<R> d, crm_input_view in, const double* lat) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ncrms) return;
  if (lat) {
    // crm_module.F90:287,321-322: latitude in degrees from radians (57.296_r8), then the
    // Coriolis parameters with the REAL(4) pi and literals of the reference build
    const double pi = (double)3.141592653589793f;
    const double lat0 = lat[c] * 57.296;
    const double fcor = 4 * pi / 86400. * sin(lat0 * pi / 180.);
    const double om = 2 * pi / (3600. * 24.);
    inst(d, I_FCOR, c) = fcor;
    inst(d, I_FCORZ, c) = sqrt(4. * (om * om) - fcor * fcor);
  }
  const int ld = d.ld, plev = d.plev, nzm = d.nzm;
  auto IN = [&](const double* p, int l) { return p[(size_t)c + (size_t)ld * l]; };
  double* z = prof(d, P_Z, c);
  double* zi = prof(d, P_ZI, c);
  double* pres = prof(d, P_PRES, c);
  double* presi = prof(d, P_PRESI, c);
  double* adz = prof(d, P_ADZ, c);
  double* adzw = prof(d, P_ADZW, c);
  double* rho = prof(d, P_RHO, c);
  double* rhow = prof(d, P_RHOW, c);
  double* bet = prof(d, P_BET, c);
  double* gamaz = prof(d, P_GAMAZ, c);
  double* grdf = prof(d, P_GRDFX, c);
  double* tau = prof(d, P_TAU, c);
  double zs = IN(in.zint, plev);
  for (int k = 0; k < nzm; k++) {
    z[k] = IN(in.zmid, plev - 1 - k) - zs;
    zi[k] = IN(in.zint, plev - k) - zs;
    pres[k] = IN(in.pmid, plev - 1 - k) / 100.;
    presi[k] = IN(in.pint, plev - k) / 100.;
    bet[k] = GGR_D / IN(in.tl, plev - 1 - k);
    gamaz[k] = GGR_D / CP_D * z[k];
  }
  zi[nzm] = IN(in.zint, plev - nzm) - zs;
  presi[nzm] = IN(in.pint, plev - nzm) / 100.;
  double dz = 0.5 * (z[0] + z[1]);
  inst(d, I_DZ, c) = dz;
  for (int k = 1; k < nzm; k++) adzw[k] = (z[k] - z[k - 1]) / dz;
  adzw[0] = 1.;
  adzw[nzm] = adzw[nzm - 1];
  for (int k = 0; k < nzm; k++) adz[k] = (zi[k + 1] - zi[k]) / dz;
  for (int k = 0; k < nzm; k++) rho[k] = IN(in.pdel, plev - 1 - k) / GGR_D / (adz[k] * dz);
  for (int k = 1; k < nzm; k++) rhow[k] = (IN(in.pmid, plev - k) - IN(in.pmid, plev - 1 - k)) / GGR_D / (adzw[k] * dz);
  rhow[0] = 2. * rhow[1] - rhow[2];
  rhow[nzm] = 2. * rhow[nzm - 1] - rhow[nzm - 2];
  for (int k = 0; k < nzm; k++) {
    double a = adz[k] * dz;
    grdf[k] = fmin(16.0, d.dx * d.dx / (a * a));
  }
  // sponge: n_damp is the deepest level within 0.4 of the domain height from the top
  int n_damp = 0;
  for (int k = nzm - 1; k >= 0; k--)
    if (z[nzm - 1] - z[k] < R4(0.4) * z[nzm - 1]) n_damp = nzm - k;
  inst(d, I_NDAMP, c) = (double)n_damp;
  int kbot = nzm - 1 - n_damp;  // 0-based level of the reference's z(nzm-n_damp)
  if (kbot < 0) kbot = 0;
  for (int k = 0; k < nzm; k++) {
    tau[k] = 0.0;
    if (k + 1 >= nzm - n_damp) {
      double tt = 60. * pow(450. / 60., (z[nzm - 1] - z[k]) / (z[nzm - 1] - z[kbot]));
      tau[k] = 1. / tt;
    }
  }
  // per-level factors the sweeps would otherwise divide out per column and substep; the same
  // IEEE operations, done once per call
  double* irho = prof(d, P_IRHO, c);
  double* iadz = prof(d, P_IADZ, c);
  double* irhoadz = prof(d, P_IRHOADZ, c);
  double* irhowadz = prof(d, P_IRHOWADZ, c);
  double* ddadz = prof(d, P_DDADZ, c);
  double* pratio = prof(d, P_PRATIO, c);
  for (int k = 0; k < nzm; k++) {
    irho[k] = 1. / rho[k];
    iadz[k] = 1. / adz[k];
    irhoadz[k] = 1. / (rho[k] * adz[k]);
    irhowadz[k] = 1. / (rhow[k] * adz[k]);
    const int kc = (k + 1 < nzm) ? k + 1 : k, kb = (k >= 1) ? k - 1 : 0;
    ddadz[k] = (double)(2.0f / (float)(kc - kb)) / adz[k];  // advect_scalar2D.F90 vertical weights
    pratio[k] = sqrt(R4(1.29) / rho[k]);
  }
  // level factors of the Smagorinsky closure (shear_prod2D.F90:22-27,35-40, shear_prod3D.F90,
  // tke_full.F90:131,248-257), the sweep's own expressions
  {
    double* srdz = prof(d, P_SRDZ, c);
    double* srdx = prof(d, P_SRDX, c);
    double* srdzw = prof(d, P_SRDZW, c);
    double* srdxw = prof(d, P_SRDXW, c);
    double* sbetdz = prof(d, P_SBETDZ, c);
    double* stkmax = prof(d, P_STKMAX, c);
    const double dx = d.dx, rdx0 = 1. / dx;
    for (int k = 0; k < nzm; k++) {
      srdz[k] = 1. / (dz * adz[k]);
      srdx[k] = rdx0 * sqrt(dx * srdz[k]);
    }
    for (int k = 0; k <= nzm; k++) {
      srdzw[k] = 1. / (dz * adzw[k]);
      srdxw[k] = rdx0 * sqrt(dx * srdzw[k]);
    }
    sbetdz[nzm - 1] = 0.0;
    for (int k = 0; k + 1 < nzm; k++) sbetdz[k] = 0.5 * (bet[k + 1] + bet[k]) / dz / adzw[k + 1];
    for (int k = 0; k < nzm; k++) {
      const double cx = dx * dx / d.dt / grdf[k];
      const double dzm = dz * fmin(adzw[k], adzw[k + 1]);
      const double cz = dzm * dzm / d.dt / 1.0;
      stkmax[k] = R4(0.09) / (1. / cx + 1. / cx + 1. / cz);
    }
    // vertical density weight of the scalar diffusion flux (diffuse_scalar2D.F90:93)
    double* drhoi = prof(d, P_DRHOI, c);
    for (int k = 0; k < nzm; k++) drhoi[k] = rhow[k + 1] / adzw[k + 1];
  }
  double ustar = sqrt(in.tau00[c] / rho[0]);
  double z0 = z0_est(z[0], in.bflxls[c], in.wndls[c], ustar, d.z0c1);
  inst(d, I_Z0, c) = fmax(R4(0.00001), fmin(1.0, z0));
}

// FP32 copies of the level profiles (mixed-precision form only).
// crm_rad%qrad from the caller's (pcols, nx_rad, ny_rad, nz) order (crm_rad_module.F90:17-34)
// into [instance][level][radiation column]; consecutive threads read consecutive instances
__global__ void k_qrad_order(const double* __restrict__ qrad, double* __restrict__ out, int ncrms, int ld, int nzm,
                             int nxr) {
  const size_t n = (size_t)ncrms * nzm * nxr;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int c = (int)(t % ncrms);
  const size_t rk = t / ncrms;  // ir + nxr * k
  const int ir = (int)(rk % nxr), k = (int)(rk / nxr);
  out[((size_t)c * nzm + k) * nxr + ir] = qrad[(size_t)c + (size_t)ld * rk];
}

template <typename R>
__global__ void k_prof_mirror(const double* src, R* dst, size_t n) {
  size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) dst[idx] = (R)src[idx];
}

// micro_diagnose (MICRO_SAM1MOM/microphysics.F90:331-357) for one point.
template <typename T>
__device__ __forceinline__ void micro_diag(T qt, T qn, T qp, T tabs, T a_bg, T a_pr, T& qv, T& qcl, T& qci,
                                           T& qpl, T& qpi) {
  qv = qt - qn;
  T omn = fmax((T)0.0, fmin((T)1.0, (tabs - (T)TBGMIN) * a_bg));
  qcl = qn * omn;
  qci = qn * ((T)1. - omn);
  T omp = fmax((T)0.0, fmin((T)1.0, (tabs - (T)TPRMIN) * a_pr));
  qpl = qp * omp;
  qpi = qp * ((T)1. - omp);
}

__device__ __forceinline__ creal A_BG() { return 1. / (TBGMAX - TBGMIN); }
__device__ __forceinline__ creal A_PR() { return 1. / (TPRMAX - TPRMIN); }
__device__ __forceinline__ creal A_GR() { return 1. / (TGRMAX - TGRMIN); }

// ------------------------------------------------------------------------------------------
// Initial fields, means and large-scale tendencies: crm_module.F90:379-594, micro_init
// (microphysics.F90:160-245), sgs_init (sgs.F90:187-213).  Segment per instance.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_init(DevT<R> d, crm_input_view in, crm_state_view st, crm_output_view out, double dt_gl) {
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, ld = d.ld, plev = d.plev, i = s.i;
  const double umax = 0.5 * d.dx / d.dt, factor_xy = 1. / (double)NX, idt_gl = 1. / dt_gl;
  const bool y3 = d.ny > 1;  // 3D: NX == 64 columns = 8 x 8, lane 8 j + i is the state's (i, j)
  auto G4 = [&](const double* p, int cc, int ii, int k) {
    return p[(size_t)cc + (size_t)ld * ((size_t)ii + (size_t)NX * k)];
  };
  // first-substep wind clamp decided by instance 0 (crm_module.F90:399-417)
  bool clamp = G4(st.u_wind, 0, 0, 0) == G4(st.u_wind, 0, 1, 0) && G4(st.u_wind, 0, 2, 1) == G4(st.u_wind, 0, 3, 1);
  const double a_bg = A_BG(), a_pr = A_PR();
  double* gamaz = prof(d, P_GAMAZ, c);
  double colprec = 0.0, colprecs = 0.0;
  int kmin = nzm, kmax = -1;
  for (int k = 0; k < nzm; k++) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    size_t ow = ((size_t)c * nz + k) * NX + i;
    double u = G4(st.u_wind, c, i, k);
    if (clamp) u = fmin(umax, fmax(-umax, u));
    double v = 0.0;
    if (y3) {
      v = G4(st.v_wind, c, i, k);
      if (clamp) v = fmin(umax, fmax(-umax, v));
    }
    double w = G4(st.w_wind, c, i, k);
    double tabs = G4(st.temperature, c, i, k);
    double qt = G4(st.qt, c, i, k), qp = G4(st.qp, c, i, k), qn = G4(st.qn, c, i, k);
    double qv, qcl, qci, qpl, qpi;
    micro_diag(qt, qn, qp, tabs, a_bg, a_pr, qv, qcl, qci, qpl, qpi);
    double t = tabs + gamaz[k] - FAC_COND_D * qcl - FAC_SUB_D * qci - FAC_COND_D * qpl - FAC_SUB_D * qpi;
    int l = plev - 1 - k;
    double pd = in.pdel[(size_t)c + (size_t)ld * l];
    colprec += s.sum((qpl + qpi) * pd);
    colprecs += s.sum(qpi * pd);
    double u0 = s.sum(u) * factor_xy;
    double v0 = y3 ? s.sum(v) * factor_xy : 0.0;
    double t0 = s.sum(t) * factor_xy;
    double t00 = s.sum(t + FAC_COND_D * qpl + FAC_SUB_D * qpi) * factor_xy;
    double tabs0 = s.sum(tabs) * factor_xy;
    double q0 = s.sum(qv + qcl + qci) * factor_xy;
    double qv0 = s.sum(qv) * factor_xy;
    double qn0 = s.sum(qcl + qci) * factor_xy;
    double qp0 = s.sum(qpl + qpi) * factor_xy;
    if (qcl + qci > 0. && tabs < R4(273.15)) {
      kmin = ::min(kmin, k);
      kmax = ::max(kmax, k);
    }
    if (live) {
      d.u[o] = u;
      d.w[ow] = w;
      d.t[o] = t;
      d.qt[o] = qt;
      d.qp[o] = qp;
      d.qn[o] = qn;
      d.tabs[o] = tabs;
      d.qv[o] = qv;
      d.qcl[o] = qcl;
      d.qci[o] = qci;
      d.qpl[o] = qpl;
      d.qpi[o] = qpi;
      d.tk[o] = 0.0;
      d.tkh[o] = 0.0;
      d.tke[o] = 0.0;
      for (int sl = 0; sl < 3; sl++) {
        d.dudt[sl][o] = 0.0;
        d.dwdt[sl][ow] = 0.0;
      }
      if (y3) {
        d.v[o] = v;
        for (int sl = 0; sl < 3; sl++) d.dvdt[sl][o] = 0.0;
      }
      if (i == 0) {
        auto IN = [&](const double* p) { return p[(size_t)c + (size_t)ld * l]; };
        double uln = fmin(umax, fmax(-umax, IN(in.ul)));
        if (y3) {
          double vln = fmin(umax, fmax(-umax, IN(in.vl)));
          prof(d, P_V0, c)[k] = v0;
          prof(d, P_VTEND, c)[k] = (vln - v0) * idt_gl;
        }
        prof(d, P_U0, c)[k] = u0;
        prof(d, P_T0, c)[k] = t0;
        prof(d, P_TABS0, c)[k] = tabs0;
        prof(d, P_Q0, c)[k] = q0;
        prof(d, P_QV0, c)[k] = qv0;
        prof(d, P_QN0, c)[k] = qn0;
        prof(d, P_QP0, c)[k] = qp0;
        prof(d, P_TTEND, c)[k] = (IN(in.tl) + gamaz[k] - FAC_COND_D * (IN(in.qccl) + IN(in.qiil)) -
                                  FAC_FUS_D * IN(in.qiil) - t00) * idt_gl;
        prof(d, P_QTEND, c)[k] = (IN(in.ql) + IN(in.qccl) + IN(in.qiil) - q0) * idt_gl;
        prof(d, P_UTEND, c)[k] = (uln - u0) * idt_gl;
        prof(d, P_UG0, c)[k] = uln;  // crm_module.F90:576-577
        if (y3) prof(d, P_VG0, c)[k] = fmin(umax, fmax(-umax, IN(in.vl)));
      }
    }
  }
  kmin = s.imin(kmin);
  kmax = s.imax(kmax);
  if (!live) return;
  size_t ow = ((size_t)c * nz + nzm) * NX + i;
  d.w[ow] = 0.0;
  for (int sl = 0; sl < 3; sl++) d.dwdt[sl][ow] = 0.0;
  d.precsfc[(size_t)c * NX + i] = 0.0;
  d.precssfc[(size_t)c * NX + i] = 0.0;
  if (i == 0) {
    inst(d, I_KMIN, c) = kmin;
    inst(d, I_KMAX, c) = kmax;
    inst(d, I_COLPREC, c) = colprec;
    inst(d, I_COLPRECS, c) = colprecs;
    inst(d, I_CLTOT, c) = out.cltot[c];
    inst(d, I_CLHGH, c) = out.clhgh[c];
    inst(d, I_CLMED, c) = out.clmed[c];
    inst(d, I_CLLOW, c) = out.cllow[c];
  }
  for (int p = P_UWLE; p < NPROF; p++)
    for (int k = i; k < nz; k += NX) prof(d, p, c)[k] = 0.0;
  for (int q = 0; q < 5; q++)
    for (int k = i; k < nzm * d.nx_rad; k += NX) d.rad[((size_t)q * d.ncrms + c) * nzm * d.nx_rad + k] = R(0.0);
}

// ------------------------------------------------------------------------------------------
// Initial noise on the first GCM step: setperturb.F90:7-86 with the Mersenne twister of
// cam/RNG_MT.F90:60-139 (whose upper mask is -1147483648).  One thread per instance; the
// 624-word state lives in thread-private memory.  Only runs when igstep <= 1.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void k_setperturb(DevT<R> d, const int* gcol) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.ncrms) return;
  uint32_t mt[624];
  mt[0] = (uint32_t)(gcol[c] * 1000);
  for (int m = 1; m < 624; m++) mt[m] = 69069u * mt[m - 1];
  int mti = 624;
  const uint32_t UMASK = (uint32_t)-1147483648LL, LMASK = 2147483647u, MATA = (uint32_t)-1727483681LL;
  auto gen = [&]() -> double {
    uint32_t y;
    if (mti >= 624) {
      int kk;
      for (kk = 0; kk < 624 - 397; kk++) {
        y = (mt[kk] & UMASK) | (mt[kk + 1] & LMASK);
        mt[kk] = mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
      }
      for (; kk < 623; kk++) {
        y = (mt[kk] & UMASK) | (mt[kk + 1] & LMASK);
        mt[kk] = mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
      }
      y = (mt[623] & UMASK) | (mt[0] & LMASK);
      mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? MATA : 0u);
      mti = 0;
    }
    y = mt[mti++];
    y ^= y >> 11;
    y ^= (y << 7) & (uint32_t)-1658038656LL;
    y ^= (y << 15) & (uint32_t)-272236544LL;
    y ^= y >> 18;
    int32_t ys = (int32_t)y;
    if (ys < 0) return ((double)ys + 4294967296.0) / 4294967296.0;
    return ((double)ys + 0.5) / 4294967296.0;
  };
  const double factor_xy = 1. / (double)NX;
  for (int k = 0; k < 5 && k < d.nzm; k++) {
    double scaling = (double)(5 - k) / 5.0;
    double t02 = 0.0;
    R* t = d.t + ((size_t)c * d.nzm + k) * NX;
    for (int i = 0; i < NX; i++) {
      double r = 1. - 2. * gen();
      t[i] = t[i] + 1 * r * scaling;
      t02 = t02 + t[i] * factor_xy;
    }
    double t0 = prof(d, P_T0, c)[k];
    for (int i = 0; i < NX; i++) t[i] = t[i] * t0 / t02;
  }
}

// ------------------------------------------------------------------------------------------
// CFL check: kurant.F90:9-83 and kurant_sgs (sgs.F90:329-372).  Batch-global maximum via an
// atomic max on the (non-negative) double's bits.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_kurant(DevT<R> d) {
  CRM_STOP_CHECK(d);
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const double dz = inst(d, I_DZ, c);
  const cprof<double> adzw = profcd(d, P_ADZW, c);
  const cprof<double> grdf = profcd(d, P_GRDFX, c);
  const bool y3 = d.ny > 1;  // 3D: dy == dx
  const double dxf = sqrt((1. / d.dx) * (1. / d.dx) + (y3 ? (1. / d.dx) * (1. / d.dx) : 0.0));
  // Every term is a monotone function of a level maximum over the columns, and rounding is
  // monotone, so the maximum over columns is taken once, of the per-column maximum over
  // levels: the same value as per-level column maxima (kurant.F90:30-60).  A maximum is exact
  // in any order, so the levels are independent iterations (the w of level k+1 is read again
  // rather than carried) and several levels' loads are in flight together.
  double cfl = 0.0;
  auto W = [&](int k) { return (double)d.w[((size_t)c * nz + k) * NX + i]; };
  auto U = [&](int k) { return (double)d.u[((size_t)c * nzm + k) * NX + i]; };
  auto V = [&](int k) { return y3 ? (double)d.v[((size_t)c * nzm + k) * NX + i] : 0.0; };
  auto T = [&](int k) { return (double)d.tkh[((size_t)c * nzm + k) * NX + i]; };
#pragma unroll 4
  for (int k = 0; k < nzm; k++) {
    const double uu = U(k), vv = V(k), tkh = T(k);
    // the lid's w counts as zero (the reference's loop stops below it)
    const double wm = fabs(W(k)), wm_up = k + 1 < nzm ? fabs(W(k + 1)) : 0.0;
    const double uhm = sqrt(uu * uu + vv * vv);
    const double dzz = dz * adzw[k];
    const double sgs = fmax(fmax(0.5 * tkh * 1.0 * d.dt / (dzz * dzz), 0.5 * tkh * grdf[k] * d.dt / (d.dx * d.dx)), 0.0);
    cfl = fmax(cfl, sgs);
    cfl = fmax(cfl, fmax(uhm * d.dt * dxf, fmax(wm, wm_up) * d.dt / (dz * adzw[k])));
  }
  cfl = s.max(cfl);
  if (live && i == 0) atomic_max_nonneg(&d.gmax[GMAX_FLAGS + (c % KURANT_SLOTS) * KURANT_STRIDE], cfl);
}

// ------------------------------------------------------------------------------------------
// zero + buoyancy + forcing + radiative heating + sponge + cloud-ice fall, one sweep:
// zero.F90:8-25, buoyancy.F90:7-38, forcing.F90:7-79, crm_module.F90:766-777,
// damping.F90:9-117, ice_fall.F90:6-173.  Writes dudt(na), dwdt(na) from scratch.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_pre(DevT<R> d, int na, const double* qrad) {
  CRM_STOP_CHECK(d);
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const R dtn = d.dtn, dz = inst(d, I_DZ, c);
  const cprof<R> adz = profc(d, P_ADZ, c);
  const cprof<R> rho = profc(d, P_RHO, c);
  const cprof<R> bet = profc(d, P_BET, c);
  const cprof<R> tabs0 = profc(d, P_TABS0, c);
  const cprof<R> qv0 = profc(d, P_QV0, c);
  const cprof<R> qn0 = profc(d, P_QN0, c);
  const cprof<R> qp0 = profc(d, P_QP0, c);
  const cprof<R> ttend = profc(d, P_TTEND, c);
  const cprof<R> qtend = profc(d, P_QTEND, c);
  const cprof<R> utend = profc(d, P_UTEND, c);
  const cprof<R> vtend = profc(d, P_VTEND, c);
  const cprof<R> tau = profc(d, P_TAU, c);
  const bool y3 = d.ny > 1;
  const int n_damp = (int)inst(d, I_NDAMP, c);
  const int kmin = (int)inst(d, I_KMIN, c), kmax = (int)inst(d, I_KMAX, c);
  const int nxr = d.nx_rad, ir = rad_col(d, i);
  R* dudt = d.dudt[na];
  R* dwdt = d.dwdt[na];
  auto F = [&](const R* p, int k) { return p[((size_t)c * nzm + k) * NX + i]; };
  // cloud-ice flux fz(k) (ice_fall.F90:94-128) from qci at levels max(0, k-1), k and
  // min(nzm-1, k+1); zero outside [kmin-1, kmax]
  auto icefz = [&](int k, R qci_b, R qci_k, R qci_c) -> R {
    if (k < ::max(0, kmin - 1) || k > kmax || k > nzm - 1) return R(0.0);
    int kc = ::min(nzm - 1, k + 1), kb = ::max(0, k - 1);
    R coef = dtn / (R(0.5) * (adz[kb] + adz[k]) * dz);
    R qiu = rho[kc] * qci_c, qic = rho[k] * qci_k, qid = rho[kb] * qci_b;
    R vt = fmin(R4(0.4), R4(8.66) * fpow<R>(fmax(R(0.0), qic) + R4(1.e-10), R4(0.24)));
    R phi;
    if (fabs(qic - qid) < R4(1.0e-25)) {
      phi = R(0.);
    } else {
      R th = (qiu - qic) / (qic - qid);
      phi = fmax(R(0.0), fmin(fmin(R(0.5) * (R(1.) + th), R(2.0)), R(2.) * th));
    }
    return -vt * (qic - R(0.5) * (R(1.) - coef * vt) * phi * (qic - qid));
  };
  // qci of levels k, k+1, k+2 at level k (clamped at the top, where the clamped level is the
  // one the flux uses): the flux icefz(k+1) needs all three, so they are read at least one
  // level ahead, and level k+3 is read at the top of level k.  Vector loads complete in issue
  // order, so a read of this level issued after the prefetch of the next would wait for it.
  R qi0 = F(d.qci, 0), qi1 = F(d.qci, ::min(1, nzm - 1)), qi2 = F(d.qci, ::min(2, nzm - 1));
  // the surface precipitation the ice flux adds to at level 0, read before the sweep's loads
  const double psfc0 = d.precsfc[(size_t)c * NX + i], pssfc0 = d.precssfc[(size_t)c * NX + i];
  R fz_k = icefz(0, qi0, qi0, qi1);
  // level k-1 thermodynamics for the buoyancy average
  R pb_tabs = 0, pb_qv = 0, pb_qcl = 0, pb_qci = 0, pb_qpl = 0, pb_qpi = 0;
  // the fields of one level, read one level ahead of their use
  struct Lvl {
    R tabs, qv, qcl, qpl, qpi, u, w, t, qt, v;
    double qr;
  };
  auto load = [&](int k, Lvl& L) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    L.tabs = d.tabs[o];
    L.qv = d.qv[o];
    L.qcl = d.qcl[o];
    L.qpl = d.qpl[o];
    L.qpi = d.qpi[o];
    L.u = d.u[o];
    L.w = d.w[((size_t)c * nz + k) * NX + i];
    L.t = d.t[o];
    L.qt = d.qt[o];
    L.v = y3 ? d.v[o] : R(0.0);
    L.qr = qrad[((size_t)c * nzm + k) * nxr + ir];  // [instance][level][radiation column]
  };
  Lvl nxt;
  load(0, nxt);
  // the results of a level are stored at the top of the next one, before its reads: vector
  // memory operations complete in issue order, so the wait for the reads at the end of a
  // level then never includes stores issued after them
  R st_t = 0, st_qt = 0, st_du = 0, st_dv = 0, st_dw = 0;
  auto store_level = [&](int k) {
    const size_t o = ((size_t)c * nzm + k) * NX + i;
    d.t[o] = st_t;
    d.qt[o] = st_qt;
    dudt[o] = st_du;
    if (y3) d.dvdt[na][o] = st_dv;
    dwdt[((size_t)c * nz + k) * NX + i] = st_dw;
  };
  for (int k = 0; k < nzm; k++) {
    const Lvl L = nxt;
    if (k > 0 && live) store_level(k - 1);
    // clamped at the top, where the record is not used again: an unconditional load needs no
    // copy that waits for earlier loads
    load(::min(k + 1, nzm - 1), nxt);
    const R qi3 = F(d.qci, ::min(k + 3, nzm - 1));
    R tabs = L.tabs, qv = L.qv, qcl = L.qcl, qci = qi0, qpl = L.qpl, qpi = L.qpi;
    R u = L.u, w = L.w;
    R t = L.t, qt = L.qt;
    // buoyancy on w level k from scalar levels k-1 and k
    R dw = R(0.0);
    if (k >= 1) {
      int kb = k - 1;
      R betu = adz[kb] / (adz[k] + adz[kb]);
      R betd = adz[k] / (adz[k] + adz[kb]);
      R up = tabs0[k] * (EPSV * (qv - qv0[k]) - (qcl + qci - qn0[k] + qpl + qpi - qp0[k])) +
                  (tabs - tabs0[k]) * (R(1.) + EPSV * qv0[k] - qn0[k] - qp0[k]);
      R dn = tabs0[kb] * (EPSV * (pb_qv - qv0[kb]) - (pb_qcl + pb_qci - qn0[kb] + pb_qpl + pb_qpi - qp0[kb])) +
                  (pb_tabs - tabs0[kb]) * (R(1.) + EPSV * qv0[kb] - qn0[kb] - qp0[kb]);
      dw = R(0.0) + bet[k] * betu * up + bet[kb] * betd * dn;
    }
    pb_tabs = tabs;
    pb_qv = qv;
    pb_qcl = qcl;
    pb_qci = qci;
    pb_qpl = qpl;
    pb_qpi = qpi;
    // large-scale forcing with level-wise refill of negative vapour
    t = t + ttend[k] * dtn;
    qt = qt + qtend[k] * dtn;
    // the level sums only matter when some column went negative; the vote covers every lane
    // of the instance (and of the wave), so all of them take the same branch
    const bool neg = qt < R(0.);
    if (s.any(neg)) {
      R nneg = s.sum(neg ? R(1.0) : R(0.0));
      R qneg = s.sum(neg ? qt : R(0.0));
      R qpoz = s.sum(neg ? R(0.0) : qt);
      if (nneg > R(0.5) && qpoz + qneg > R(0.)) qt = fmax(R(0.0), qt * (R(1.) + qneg / qpoz));
    }
    R du = R(0.0) + utend[k];
    R v = L.v, dv = R(0.0);
    if (y3) dv = dv + vtend[k];
    // radiative heating on column groups
    t = t + L.qr * dtn;
    // sponge layer
    if (k + 1 >= nzm - n_damp) {
      R tk = tau[k];
      R u0l = s.sum(u / (R)NX), t0l = s.sum(t / (R)NX);
      du = du - (u - u0l) * tk;
      if (y3) {
        R v0l = s.sum(v / (R)NX);
        dv = dv - (v - v0l) * tk;
      }
      dw = dw - w * tk;
      t = t - dtn * (t - t0l) * tk;
      qt = qt - dtn * (qv - qv0[k]) * tk;
    }
    // cloud-ice sedimentation
    R fz_kp = icefz(k + 1, qi0, qi1, qi2);
    if (k >= ::max(0, kmin - 2) && k <= kmax) {
      R coef = dtn / (dz * adz[k] * rho[k]);
      R dqi = coef * (fz_k - fz_kp);
      qt = qt + dqi;
      t = t - (FAC_COND + FAC_FUS) * dqi;
    }
    if (k == 0 && live) {
      R dqi = -(dtn / dz) * fz_k;
      d.precsfc[(size_t)c * NX + i] = psfc0 + dqi;
      d.precssfc[(size_t)c * NX + i] = pssfc0 + dqi;
    }
    fz_k = fz_kp;
    qi0 = qi1;
    qi1 = qi2;
    qi2 = qi3;
    st_t = t;
    st_qt = qt;
    st_du = du;
    st_dv = dv;
    st_dw = dw;
  }
  if (live) {
    store_level(nzm - 1);
    dwdt[((size_t)c * nz + nzm) * NX + i] = R(0.0);
  }
}

// ------------------------------------------------------------------------------------------
// Deformation squared of a 3D CRM at level k < nzm-1 (shear_prod3D.F90:26-80 for k >= 1,
// :81-106 at the surface level), dy == dx.  Window: u, v at k-1, k, k+1 and w at k, k+1.
// The four horizontal shear terms are one staggered quantity H at the (i+1/2, j+1/2) corners
// read at the four corners of the cell.
// ------------------------------------------------------------------------------------------
template <typename R, typename G>
__device__ __forceinline__ R shear3(const G& s, int k, cprof<R> srdz, cprof<R> srdx, cprof<R> srdzw, cprof<R> srdxw,
                                    cprof<R> u0, cprof<R> v0, R um, R u, R up, R vm, R v, R vp, R w0, R w1) {
  const int kc = k + 1, kb = k - 1;
  const R rdz = srdz[k], rdx = srdx[k];
  const R rdzw_up = srdzw[kc], rdx_up = srdxw[kc];
  // rdy* == rdx* for dy == dx
  R u_ip = s.rot(u, 1), v_jp = s.roty(v, 1);
  R a = (u_ip - u) * rdx, b = (v_jp - v) * rdx, cz = (w1 - w0) * rdz;
  R def2 = R(2.) * (a * a + b * b + cz * cz);
  R H = (s.roty(u_ip, 1) - u_ip) * rdx + (s.rot(v_jp, 1) - v_jp) * rdx;  // corner (i+1/2, j+1/2)
  R H_im = s.rot(H, -1), H_jm = s.roty(H, -1), H_ijm = s.roty(H_im, -1);
  def2 = def2 + R(0.25) * (H * H + H_im * H_im + H_jm * H_jm + H_ijm * H_ijm);
  // u-w and v-w shear at the upper interface
  R up_ip = s.rot(up, 1), w1p = s.rot(w1, 1), w1m = s.rot(w1, -1);
  R a1 = (up_ip - u0[kc] - u_ip + u0[k]) * rdzw_up + (w1p - w1) * rdx_up;
  R a2 = (up - u0[kc] - u + u0[k]) * rdzw_up + (w1 - w1m) * rdx_up;
  R vp_jp = s.roty(vp, 1), w1jp = s.roty(w1, 1), w1jm = s.roty(w1, -1);
  R b1 = (vp_jp - v0[kc] - v_jp + v0[k]) * rdzw_up + (w1jp - w1) * rdx_up;
  R b2 = (vp - v0[kc] - v + v0[k]) * rdzw_up + (w1 - w1jm) * rdx_up;
  if (k >= 1) {
    const R rdzw_dn = srdzw[k], rdx_dn = srdxw[k];
    R um_ip = s.rot(um, 1), w0p = s.rot(w0, 1), w0m = s.rot(w0, -1);
    R a3 = (u_ip - u0[k] - um_ip + u0[kb]) * rdzw_dn + (w0p - w0) * rdx_dn;
    R a4 = (u - u0[k] - um + u0[kb]) * rdzw_dn + (w0 - w0m) * rdx_dn;
    R vm_jp = s.roty(vm, 1), w0jp = s.roty(w0, 1), w0jm = s.roty(w0, -1);
    R b3 = (v_jp - v0[k] - vm_jp + v0[kb]) * rdzw_dn + (w0jp - w0) * rdx_dn;
    R b4 = (v - v0[k] - vm + v0[kb]) * rdzw_dn + (w0 - w0jm) * rdx_dn;
    def2 = def2 + R(0.25) * (a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4);
    def2 = def2 + R(0.25) * (b1 * b1 + b2 * b2 + b3 * b3 + b4 * b4);
  } else {
    def2 = def2 + R(0.5) * (b1 * b1 + b2 * b2);
    def2 = def2 + R(0.5) * (a1 * a1 + a2 * a2);
  }
  return def2;
}

// ------------------------------------------------------------------------------------------
// Smagorinsky eddy diffusivity: shear_prod2D.F90:7-96 (3D: shear_prod3D.F90:7-140) and
// tke_full.F90:11-317 (dosmagor).
// ------------------------------------------------------------------------------------------
template <int NX, typename R, typename G = SegFor<NX>>
__global__ void CRM_LB k_sgs(DevT<R> d) {
  CRM_STOP_CHECK(d);
  G s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const R dz = inst(d, I_DZ, c), dx = d.dx;
  const cprof<R> adz = profc(d, P_ADZ, c);
  const cprof<R> adzw = profc(d, P_ADZW, c);
  const cprof<R> bet = profc(d, P_BET, c);
  const cprof<R> presi = profc(d, P_PRESI, c);
  const cprof<R> u0 = profc(d, P_U0, c);
  const cprof<R> srdz = profc(d, P_SRDZ, c);
  const cprof<R> srdx = profc(d, P_SRDX, c);
  const cprof<R> srdzw = profc(d, P_SRDZW, c);
  const cprof<R> srdxw = profc(d, P_SRDXW, c);
  const cprof<R> sbetdz = profc(d, P_SBETDZ, c);
  const cprof<R> stkmax = profc(d, P_STKMAX, c);
  const R Cs = R4(0.15), Ck = R4(0.1), Pr = R(1.);
  const R Ce = Ck * Ck * Ck / (Cs * Cs * Cs * Cs);
  auto F = [&](const R* p, int k) { return p[((size_t)c * nzm + k) * NX + i]; };
  auto W = [&](int k) { return d.w[((size_t)c * nz + k) * NX + i]; };
  // window: u (and v in 3D) at k-1,k,k+1 and w at k,k+1
  R um = 0, u = F(d.u, 0), up = F(d.u, 1);
  R vm = 0, v = 0, vp = 0;
  const cprof<R> v0 = profc(d, P_V0, c);
  if constexpr (G::Y3) {
    v = F(d.v, 0);
    vp = F(d.v, 1);
  }
  R w0 = W(0), w1 = W(1);
  R bsv_prev = R(0.0);
  // thermodynamics at k (carried) and k+1; level k+2 and tk at k+1 are read one iteration
  // ahead of their use
  R th_tabs = F(d.tabs, 0), th_qcl = F(d.qcl, 0), th_qci = F(d.qci, 0), th_qv = F(d.qv, 0);
  R th_qpl = F(d.qpl, 0), th_qpi = F(d.qpi, 0), th_t = F(d.t, 0);
  R nx_tabs = F(d.tabs, 1), nx_qcl = F(d.qcl, 1), nx_qci = F(d.qci, 1), nx_qv = F(d.qv, 1);
  R nx_qpl = F(d.qpl, 1), nx_qpi = F(d.qpi, 1), nx_t = F(d.t, 1);
  R tk_cur = F(d.tk, 0);
  for (int k = 0; k < nzm - 1; k++) {
    int kc = k + 1, kb = k - 1;
    const R n_tabs = nx_tabs, n_qcl = nx_qcl, n_qci = nx_qci, n_qv = nx_qv, n_qpl = nx_qpl, n_qpi = nx_qpi;
    const R n_t = nx_t, tk_k = tk_cur;
    // level kc+1, clamped at the top: past it the values are not used (the loop ends), and a
    // load that always issues needs no copy that waits for earlier loads
    const int kn = ::min(kc + 1, nzm - 1);
    nx_tabs = F(d.tabs, kn);
    nx_qcl = F(d.qcl, kn);
    nx_qci = F(d.qci, kn);
    nx_qv = F(d.qv, kn);
    nx_qpl = F(d.qpl, kn);
    nx_qpi = F(d.qpi, kn);
    nx_t = F(d.t, kn);
    tk_cur = F(d.tk, kc);
    R up_next = F(d.u, kn);
    R vp_next = R(0.0);
    if constexpr (G::Y3) vp_next = F(d.v, kn);
    R w2 = W(kc + 1);
    // shear production def2 at level k
    R def2;
    if constexpr (G::Y3) {
      def2 = shear3(s, k, srdz, srdx, srdzw, srdxw, u0, v0, um, u, up, vm, v, vp, w0, w1);
    } else {
      const R rdz = srdz[k], rdx = srdx[k];
      R u_ip = s.rot(u, 1);
      R du = (u_ip - u) * rdx;
      R dwz = (w1 - w0) * rdz;
      def2 = R(2.) * (du * du + dwz * dwz);
      R w0m = s.rot(w0, -1), w0p = s.rot(w0, 1), w1m = s.rot(w1, -1), w1p = s.rot(w1, 1);
      R up_ip = s.rot(up, 1);
      {
        const R rdzw_up = srdzw[kc], rdx_up = srdxw[kc];
        R a1 = (up_ip - u0[kc] - u_ip + u0[k]) * rdzw_up + (w1p - w1) * rdx_up;
        R a2 = (up - u0[kc] - u + u0[k]) * rdzw_up + (w1 - w1m) * rdx_up;
        if (k >= 1) {
          const R rdzw_dn = srdzw[k], rdx_dn = srdxw[k];
          R um_ip = s.rot(um, 1);
          R a3 = (u_ip - u0[k] - um_ip + u0[kb]) * rdzw_dn + (w0p - w0) * rdx_dn;
          R a4 = (u - u0[k] - um + u0[kb]) * rdzw_dn + (w0 - w0m) * rdx_dn;
          def2 = def2 + R(0.5) * (a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4);
        } else {
          def2 = def2 + R(0.5) * (a1 * a1 + a2 * a2);
        }
      }
    }
    // SGS buoyancy flux at the interface between k and k+1 (tke_full.F90:120-237)
    R bsv;
    {
      const R betdz = sbetdz[k];
      R tabs_if = R(0.5) * (n_tabs + FAC_COND * n_qcl + FAC_SUB * n_qci + th_tabs + FAC_COND * th_qcl + FAC_SUB * th_qci);
      R qtot_if = R(0.5) * (n_qv + n_qcl + n_qci + th_qv + th_qcl + th_qci);
      R qp_if = R(0.5) * (n_qpl + n_qpi + th_qpl + th_qpi);
      R bbb = R(1.) + EPSV * qtot_if - qp_if;
      R dtt = n_t - th_t;
      R dqt = n_qv + n_qcl + n_qci - th_qv - th_qcl - th_qci;
      bsv = betdz * (bbb * dtt + EPSV * tabs_if * dqt + (bbb * FAC_COND - tabs_if) * (n_qpl - th_qpl) +
                     (bbb * FAC_SUB - tabs_if) * (n_qpi - th_qpi));
      R qctot = n_qcl + n_qci + th_qcl + th_qci;
      if (qctot > R(0.)) {
        R omn = (n_qcl + th_qcl) / (qctot + R4(1.e-20));
        R pp = presi[k + 1];
        R qsat_check = omn * qsatw(tabs_if, pp) + (R(1.) - omn) * qsati(tabs_if, pp);
        if (qtot_if > qsat_check) {
          R lstarn = FAC_COND + (R(1.) - omn) * FAC_FUS;
          R tif = R(0.5) * (n_tabs + th_tabs);
          R dqsat = omn * dtqsatw(tif, pp) + (R(1.) - omn) * dtqsati(tif, pp);
          R qsatt = omn * qsatw(tif, pp) + (R(1.) - omn) * qsati(tif, pp);
          R b2 = R(1.) + EPSV * qsatt + qsatt - qtot_if - qp_if + R4(1.61) * tif * dqsat;
          b2 = b2 / (R(1.) + lstarn * dqsat);
          R tk_ = th_tabs;
          bsv = betdz * (b2 * dtt + (b2 * lstarn - (R(1.) + lstarn * dqsat) * tif) * dqt +
                         (b2 * FAC_COND - (R(1.) + FAC_COND * dqsat) * tk_) * (n_qpl - th_qpl) +
                         (b2 * FAC_SUB - (R(1.) + FAC_SUB * dqsat) * tk_) * (n_qpi - th_qpi));
        }
      }
    }
    // eddy viscosity at level k (tke_full.F90:246-308)
    {
      R grd = dz * adz[k];
      R Ce1 = Ce / R4(0.7) * R4(0.19);
      R Ce2 = Ce / R4(0.7) * R4(0.51);
      const R tkmax = stkmax[k];
      R buoy = R(0.5) * (bsv_prev + bsv);
      size_t o = ((size_t)c * nzm + k) * NX + i;
      R smix;
      if (buoy <= R(0.))
        smix = grd;
      else
        smix = fmin(grd, fmax(R4(0.1) * grd, sqrt(R4(0.76) * tk_k / Ck / sqrt(buoy + R4(1.e-10)))));
      R ratio = smix / grd;
      R Cee = Ce1 + Ce2 * ratio;
      R tk = sqrt(Ck * Ck * Ck / Cee * fmax(R(0.0), def2 - Pr * buoy)) * (smix * smix);
      R tt = tk / (Ck * smix);
      tk = fmin(tk, tkmax);
      if (live) {
        d.tke[o] = tt * tt;
        d.tk[o] = tk;
        d.tkh[o] = Pr * tk;
      }
    }
    bsv_prev = bsv;
    th_tabs = n_tabs;
    th_qcl = n_qcl;
    th_qci = n_qci;
    th_qv = n_qv;
    th_qpl = n_qpl;
    th_qpi = n_qpi;
    th_t = n_t;
    um = u;
    u = up;
    up = up_next;
    vm = v;
    v = vp;
    vp = vp_next;
    w0 = w1;
    w1 = w2;
  }
}

// ------------------------------------------------------------------------------------------
// Momentum advection and SGS momentum diffusion: advect2_mom_xy.F90:60-84,
// advect2_mom_z.F90:7-118, diffuse_mom2D.F90:7-161.  Each level's four contributions are
// added in the reference's order.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__device__ __forceinline__ void mom2(const DevT<R>& d, int na) {
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const R dz = inst(d, I_DZ, c), dx = d.dx;
  const cprof<R> adz = profc(d, P_ADZ, c);
  const cprof<R> adzw = profc(d, P_ADZW, c);
  const cprof<R> rho = profc(d, P_RHO, c);
  const cprof<R> rhow = profc(d, P_RHOW, c);
  const cprof<R> grdf = profc(d, P_GRDFX, c);
  R* dudt = d.dudt[na];
  R* dwdt = d.dwdt[na];
  const R dx25 = R(0.25) / dx, dz25 = R(1.) / (R(4.) * dz);
  const R rdx2 = R(1.) / dx / dx, rdx25 = R(0.25) * rdx2;
  const R rdz = R(1.) / dz, dzx = dz / dx;
  auto F = [&](const R* p, int k) { return p[((size_t)c * nzm + k) * NX + i]; };
  auto W = [&](int k) { return d.w[((size_t)c * nz + k) * NX + i]; };
  R u_k = F(d.u, 0), u_n = F(d.u, 1);
  R w_k = W(0), w_n = W(1);
  R tk_k = F(d.tk, 0), tk_n = F(d.tk, 1);
  R fuz_k = R(0.0);    // fuz at level k (0 at the surface)
  R fwz_prev = R(0.0); // fwz(k-1)
  R fuv_k = R(0.0);    // vertical SGS u flux at the bottom of level k
  R fwv_k = R(0.0);    // vertical SGS w flux at w level k
  R pend_adv = R(0.0), pend_dif = R(0.0);  // horizontal contributions for w level k
  R uwsb_sum = R(0.0);
  // tendencies of the next level, read one iteration ahead
  R du_pre = dudt[((size_t)c * nzm) * NX + i];
  R dw_pre = dwdt[((size_t)c * nz + 1) * NX + i];
  for (int k = 0; k < nzm; k++) {
    int kn = k + 1;
    bool has_n = kn < nzm;
    int kcu = has_n ? kn : k;
    const R du_k = du_pre, dw_k = dw_pre;
    // the next levels, clamped at the top: values read past it enter only terms of levels
    // that have a level above (has_n), and a load that always issues needs no copy that waits
    // for earlier loads
    const int kc1 = ::min(kn, nzm - 1), kc2 = ::min(kn + 1, nzm - 1);
    du_pre = dudt[((size_t)c * nzm + kc1) * NX + i];
    dw_pre = dwdt[((size_t)c * nz + kn) * NX + i];
    R u_nn = F(d.u, kc2);
    R w_nn = W(::min(kn + 1, nz - 1));
    R tk_nn = F(d.tk, kc2);
    R u_m = s.rot(u_k, -1), u_p = s.rot(u_k, 1);
    R ucu = has_n ? u_n : u_k, ucu_p = s.rot(ucu, 1);
    R wn_m = s.rot(w_n, -1), wn_p = s.rot(w_n, 1);
    R tkcu = has_n ? tk_n : tk_k;
    R tk_p = s.rot(tk_k, 1), tkcu_p = s.rot(tkcu, 1);
    R tk_m = s.rot(tk_k, -1), tkn_m = s.rot(tk_n, -1);
    // ---- dudt(k)
    size_t o = ((size_t)c * nzm + k) * NX + i;
    R du = du_k;
    R fu1 = dx25 * (u_k + u_m) * (u_m + u_k);
    R fu2 = dx25 * (u_p + u_k) * (u_k + u_p);
    du = du - (fu2 - fu1);
    R fuz_n = R(0.0);
    if (has_n) {
      R rhoi = dz25 * rhow[kn];
      R www = rhoi * (w_n + wn_m);
      fuz_n = www * (u_n + u_k);
    }
    R rhoi_k = R(1.) / (rho[k] * adz[k]);
    du = du - (fuz_n - fuz_k) * rhoi_k;
    R fuh = -R(2.) * (rdx2 * grdf[k] * tk_k) * (u_p - u_k);
    R fuh_m = s.rot(fuh, -1);
    du = du - (fuh - fuh_m);
    // vertical SGS fluxes at the top of level k (w level k+1)
    R fuv_n = R(0.0), fwv_n;
    if (has_n) {
      R rdz2 = rdz * rdz * R(1.0);
      R rdz25 = R(0.25) * rdz2;
      R iadz = R(1.) / adz[k], iadzw = R(1.) / adzw[kn];
      R tkz = rdz2 * tk_k;
      fwv_n = -R(2.) * tkz * (w_n - w_k) * rho[k] * iadz;
      tkz = rdz25 * (tk_k + tk_m + tk_n + tkn_m);
      fuv_n = -tkz * ((u_n - u_k) * iadzw + (w_n - wn_m) * dzx) * rhow[kn];
      uwsb_sum = s.sum(fuv_n);
      if (live && i == 0) prof(d, P_UWSB, c)[kn] = uwsb_sum;
    } else {
      R rdz2 = rdz * rdz * R(1.0);
      R tkz = rdz2 * R(1.0) * tk_k;
      fwv_n = -R(2.) * tkz * (w_n - w_k) / adz[k] * rho[k];
    }
    du = du - (fuv_n - fuv_k) * rhoi_k;
    if (live) dudt[o] = du;
    R fuz_sum = s.sum(fuz_n);
    if (has_n && live && i == 0) prof(d, P_UWLE, c)[kn] = fuz_sum;
    // ---- dwdt(k) for w levels 1..nzm-1
    R fwz_k = dz25 * (w_n * rhow[kn] + w_k * rhow[k]) * (w_n + w_k);
    if (k >= 1) {
      size_t ow = ((size_t)c * nz + k) * NX + i;
      R dw = dw_k;
      R rhoiw = R(1.) / (rhow[k] * adzw[k]);
      dw = dw - pend_adv;
      dw = dw - (fwz_k - fwz_prev) * rhoiw;
      dw = dw - pend_dif;
      dw = dw - (fwv_n - fwv_k) * rhoiw;
      if (live) dwdt[ow] = dw;
    }
    // ---- horizontal contributions to w level k+1 (applied next iteration)
    {
      R irho = R(1.) / (rhow[kn] * adzw[kn]);
      R fw1 = dx25 * (u_k * rho[k] * adz[k] + ucu * rho[kcu] * adz[kcu]) * (wn_m + w_n);
      R fw2 = dx25 * (u_p * rho[k] * adz[k] + ucu_p * rho[kcu] * adz[kcu]) * (w_n + wn_p);
      pend_adv = irho * (fw2 - fw1);
      R dxz = dx / (dz * adzw[kn]);
      R tkx = rdx25 * grdf[k] * (tk_k + tk_p + tkcu + tkcu_p);
      R fwh = -tkx * (wn_p - w_n + (ucu_p - u_p) * dxz);
      R fwh_m = s.rot(fwh, -1);
      pend_dif = fwh - fwh_m;
    }
    fuz_k = fuz_n;
    fwz_prev = fwz_k;
    fuv_k = fuv_n;
    fwv_k = fwv_n;
    u_k = u_n;
    u_n = u_nn;
    w_k = w_n;
    w_n = w_nn;
    tk_k = tk_n;
    tk_n = tk_nn;
  }
  // lid (w level nzm): horizontal contributions only
  if (live) {
    size_t ow = ((size_t)c * nz + nzm) * NX + i;
    R dw = dwdt[ow];
    dw = dw - pend_adv;
    dw = dw - pend_dif;
    dwdt[ow] = dw;
    if (i == 0) {
      prof(d, P_UWLE, c)[0] = R(0.0);
      prof(d, P_UWSB, c)[0] = R(0.0);
    }
  }
}

// 3D momentum (NX = nx*ny columns of G's x-y geometry per instance): advect2_mom_xy.F90:22-58,
// advect2_mom_z.F90:7-118 (3D branch :47-63), diffuse_mom3D.F90:7-207, dy == dx.  Each
// tendency receives its contributions in the reference's order: x then y advection, vertical
// advection, x then y diffusion, vertical diffusion.
template <int NX, typename R, typename G>
__device__ __forceinline__ void mom3(const DevT<R>& d, int na) {
  G s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const R dz = inst(d, I_DZ, c), dx = d.dx;
  const cprof<R> adz = profc(d, P_ADZ, c);
  const cprof<R> adzw = profc(d, P_ADZW, c);
  const cprof<R> rho = profc(d, P_RHO, c);
  const cprof<R> rhow = profc(d, P_RHOW, c);
  const cprof<R> grdf = profc(d, P_GRDFX, c);  // grdf_y == grdf_x
  R* dudt = d.dudt[na];
  R* dvdt = d.dvdt[na];
  R* dwdt = d.dwdt[na];
  const R dx25 = R(0.25) / dx, dz25 = R(1.) / (R(4.) * dz);
  const R rdx2 = R(1.) / (dx * dx), rdx25 = R(0.25) * rdx2;
  const R rdz = R(1.) / dz, dzx = dz / dx, dxy = dx / dx;
  auto F = [&](const R* p, int k) { return p[((size_t)c * nzm + k) * NX + i]; };
  auto W = [&](int k) { return d.w[((size_t)c * nz + k) * NX + i]; };
  R u_k = F(d.u, 0), u_n = F(d.u, 1);
  R v_k = F(d.v, 0), v_n = F(d.v, 1);
  R w_k = W(0), w_n = W(1);
  R tk_k = F(d.tk, 0), tk_n = F(d.tk, 1);
  R fuz_k = R(0.0), fvz_k = R(0.0);  // vertical advective fluxes at the bottom of level k
  R fwz_prev = R(0.0);
  R fuv_k = R(0.0), fvv_k = R(0.0);  // vertical SGS fluxes at the bottom of level k
  R fwv_k = R(0.0);
  R pax = R(0.0), pay = R(0.0), pdx = R(0.0), pdy = R(0.0);  // horizontal terms for w level k
  // tendencies of the next level, read one iteration ahead
  R du_pre = dudt[((size_t)c * nzm) * NX + i], dv_pre = dvdt[((size_t)c * nzm) * NX + i];
  R dw_pre = dwdt[((size_t)c * nz + 1) * NX + i];
  for (int k = 0; k < nzm; k++) {
    const int kn = k + 1;
    const bool has_n = kn < nzm;
    const int kcu = has_n ? kn : k;
    const R du_k = du_pre, dv_k = dv_pre, dw_k = dw_pre;
    // the next levels, clamped at the top (as in mom2)
    const int kc1 = ::min(kn, nzm - 1), kc2 = ::min(kn + 1, nzm - 1);
    du_pre = dudt[((size_t)c * nzm + kc1) * NX + i];
    dv_pre = dvdt[((size_t)c * nzm + kc1) * NX + i];
    dw_pre = dwdt[((size_t)c * nz + kn) * NX + i];
    R u_nn = F(d.u, kc2);
    R v_nn = F(d.v, kc2);
    R w_nn = W(::min(kn + 1, nz - 1));
    R tk_nn = F(d.tk, kc2);
    // neighbours at level k
    R u_m = s.rot(u_k, -1), u_p = s.rot(u_k, 1), u_jm = s.roty(u_k, -1), u_jp = s.roty(u_k, 1);
    R u_p_jm = s.roty(u_p, -1);
    R v_m = s.rot(v_k, -1), v_p = s.rot(v_k, 1), v_jm = s.roty(v_k, -1), v_jp = s.roty(v_k, 1);
    R v_m_jp = s.roty(v_m, 1);
    R tk_m = s.rot(tk_k, -1), tk_p = s.rot(tk_k, 1), tk_jm = s.roty(tk_k, -1), tk_jp = s.roty(tk_k, 1);
    R tk_p_jm = s.roty(tk_p, -1), tk_m_jp = s.roty(tk_m, 1);
    R ucu = has_n ? u_n : u_k, vcu = has_n ? v_n : v_k, tkcu = has_n ? tk_n : tk_k;
    R ucu_p = s.rot(ucu, 1), vcu_jp = s.roty(vcu, 1);
    R tkcu_p = s.rot(tkcu, 1), tkcu_jp = s.roty(tkcu, 1);
    R tkn_m = s.rot(tk_n, -1), tkn_jm = s.roty(tk_n, -1);
    R wn_m = s.rot(w_n, -1), wn_p = s.rot(w_n, 1), wn_jm = s.roty(w_n, -1), wn_jp = s.roty(w_n, 1);
    const R rhoi_k = R(1.) / (rho[k] * adz[k]);
    // ---- vertical fluxes at w level k+1
    R fuz_n = R(0.0), fvz_n = R(0.0), fuv_n = R(0.0), fvv_n = R(0.0), fwv_n;
    if (has_n) {
      R rhoi = dz25 * rhow[kn];
      fuz_n = rhoi * (w_n + wn_m) * (u_n + u_k);
      fvz_n = rhoi * (w_n + wn_jm) * (v_n + v_k);
      R rdz2 = rdz * rdz * R(1.0);
      R rdz25 = R(0.25) * rdz2;
      R iadz = R(1.) / adz[k], iadzw = R(1.) / adzw[kn];
      R tkz = rdz2 * tk_k;
      fwv_n = -R(2.) * tkz * (w_n - w_k) * rho[k] * iadz;
      tkz = rdz25 * (tk_k + tk_m + tk_n + tkn_m);
      fuv_n = -tkz * ((u_n - u_k) * iadzw + (w_n - wn_m) * dzx) * rhow[kn];
      tkz = rdz25 * (tk_k + tk_jm + tk_n + tkn_jm);
      fvv_n = -tkz * ((v_n - v_k) * iadzw + (w_n - wn_jm) * dzx) * rhow[kn];
      R s_uwle = s.sum(fuz_n), s_vwle = s.sum(fvz_n), s_uwsb = s.sum(fuv_n), s_vwsb = s.sum(fvv_n);
      if (live && i == 0) {
        prof(d, P_UWLE, c)[kn] = s_uwle;
        prof(d, P_VWLE, c)[kn] = s_vwle;
        prof(d, P_UWSB, c)[kn] = s_uwsb;
        prof(d, P_VWSB, c)[kn] = s_vwsb;
      }
    } else {
      R rdz2 = rdz * rdz * R(1.0);
      R tkz = rdz2 * R(1.0) * tk_k;
      fwv_n = -R(2.) * tkz * (w_n - w_k) / adz[k] * rho[k];
    }
    // ---- dudt(k)
    size_t o = ((size_t)c * nzm + k) * NX + i;
    R du = du_k;
    {
      R fu1 = dx25 * (u_k + u_m) * (u_m + u_k);
      R fu2 = dx25 * (u_p + u_k) * (u_k + u_p);
      du = du - (fu2 - fu1);
      fu1 = dx25 * (v_k + v_m) * (u_jm + u_k);
      fu2 = dx25 * (v_jp + v_m_jp) * (u_k + u_jp);
      du = du - (fu2 - fu1);
      du = du - (fuz_n - fuz_k) * rhoi_k;
      R fx = -R(2.) * (rdx2 * grdf[k] * tk_k) * (u_p - u_k);
      du = du - (fx - s.rot(fx, -1));
      R tky = rdx25 * grdf[k] * (tk_k + tk_m + tk_jp + tk_m_jp);
      R fy = -tky * (u_jp - u_k + (v_jp - v_m_jp) * dxy);
      du = du - (fy - s.roty(fy, -1));
      du = du - (fuv_n - fuv_k) * rhoi_k;
    }
    // ---- dvdt(k)
    R dv = dv_k;
    {
      R fv1 = dx25 * (u_k + u_jm) * (v_m + v_k);
      R fv2 = dx25 * (u_p + u_p_jm) * (v_k + v_p);
      dv = dv - (fv2 - fv1);
      fv1 = dx25 * (v_k + v_jm) * (v_jm + v_k);
      fv2 = dx25 * (v_jp + v_k) * (v_k + v_jp);
      dv = dv - (fv2 - fv1);
      dv = dv - (fvz_n - fvz_k) * rhoi_k;
      R tkx = rdx25 * grdf[k] * (tk_k + tk_jm + tk_p + tk_p_jm);
      R fx = -tkx * (v_p - v_k + (u_p - u_p_jm) * dxy);
      dv = dv - (fx - s.rot(fx, -1));
      R fy = -R(2.) * (rdx2 * grdf[k] * tk_k) * (v_jp - v_k);
      dv = dv - (fy - s.roty(fy, -1));
      dv = dv - (fvv_n - fvv_k) * rhoi_k;
    }
    if (d.docoriolis) {  // coriolis.F90:15-31, after advect_mom and sgs_mom (crm_module.F90:825)
      const R fcor = inst(d, I_FCOR, c), fcorz = inst(d, I_FCORZ, c);
      const R w_m = s.rot(w_k, -1);
      const R v_av = R(0.25) * (v_k + v_jp + v_m + v_m_jp);
      const R w_av = R(0.25) * (w_n + wn_m + w_k + w_m);
      du = du + fcor * (v_av - (R)prof(d, P_VG0, c)[k]) - fcorz * w_av;
      const R u_av = R(0.25) * (u_k + u_p + u_jm + u_p_jm);
      dv = dv - R(0.5) * (fcor + fcor) * (u_av - (R)prof(d, P_UG0, c)[k]);
    }
    if (live) {
      dudt[o] = du;
      dvdt[o] = dv;
    }
    // ---- dwdt(k) for w levels 1..nzm-1
    R fwz_k = dz25 * (w_n * rhow[kn] + w_k * rhow[k]) * (w_n + w_k);
    if (k >= 1) {
      size_t ow = ((size_t)c * nz + k) * NX + i;
      R dw = dw_k;
      R rhoiw = R(1.) / (rhow[k] * adzw[k]);
      dw = dw - pax;
      dw = dw - pay;
      dw = dw - (fwz_k - fwz_prev) * rhoiw;
      dw = dw - pdx;
      dw = dw - pdy;
      dw = dw - (fwv_n - fwv_k) * rhoiw;
      if (live) dwdt[ow] = dw;
    }
    // ---- horizontal terms of w level k+1 (applied next iteration)
    {
      R irho = R(1.) / (rhow[kn] * adzw[kn]);
      R fw1 = dx25 * (u_k * rho[k] * adz[k] + ucu * rho[kcu] * adz[kcu]) * (wn_m + w_n);
      R fw2 = dx25 * (u_p * rho[k] * adz[k] + ucu_p * rho[kcu] * adz[kcu]) * (w_n + wn_p);
      pax = irho * (fw2 - fw1);
      fw1 = dx25 * (v_k * rho[k] * adz[k] + vcu * rho[kcu] * adz[kcu]) * (wn_jm + w_n);
      fw2 = dx25 * (v_jp * rho[k] * adz[k] + vcu_jp * rho[kcu] * adz[kcu]) * (w_n + wn_jp);
      pay = irho * (fw2 - fw1);
      R dxz = dx / (dz * adzw[kn]);
      R tkx = rdx25 * grdf[k] * (tk_k + tk_p + tkcu + tkcu_p);
      R fx = -tkx * (wn_p - w_n + (ucu_p - u_p) * dxz);
      pdx = fx - s.rot(fx, -1);
      R tky = rdx25 * grdf[k] * (tk_k + tk_jp + tkcu + tkcu_jp);
      R fy = -tky * (wn_jp - w_n + (vcu_jp - v_jp) * dxz);
      pdy = fy - s.roty(fy, -1);
    }
    fuz_k = fuz_n;
    fvz_k = fvz_n;
    fwz_prev = fwz_k;
    fuv_k = fuv_n;
    fvv_k = fvv_n;
    fwv_k = fwv_n;
    u_k = u_n;
    u_n = u_nn;
    v_k = v_n;
    v_n = v_nn;
    w_k = w_n;
    w_n = w_nn;
    tk_k = tk_n;
    tk_n = tk_nn;
  }
  // lid (w level nzm): horizontal terms only
  if (live) {
    size_t ow = ((size_t)c * nz + nzm) * NX + i;
    R dw = dwdt[ow];
    dw = dw - pax;
    dw = dw - pay;
    dw = dw - pdx;
    dw = dw - pdy;
    dwdt[ow] = dw;
    if (i == 0) {
      prof(d, P_UWLE, c)[0] = R(0.0);
      prof(d, P_UWSB, c)[0] = R(0.0);
      prof(d, P_VWLE, c)[0] = R(0.0);
      prof(d, P_VWSB, c)[0] = R(0.0);
    }
  }
}

template <int NX, typename R, typename G = SegFor<NX>>
__global__ void CRM_LB k_mom(DevT<R> d, int na) {
  CRM_STOP_CHECK(d);
  if constexpr (G::Y3)
    mom3<NX, R, G>(d, na);
  else
    mom2<NX, R>(d, na);
}

// ------------------------------------------------------------------------------------------
// Pressure: press_rhs.F90:8-106, the fft991 x-transform as dense operators (see
// fft991_operator.hpp; in 3D the x-y Kronecker operator, dy == dx), the vertical Thomas sweep
// of pressure.F90:160-237, the inverse transform, press_grad.F90:8-76, then adams.F90:7-40.
// 4 waves share the 64/NX instances of one wave; levels are dealt round-robin to the waves;
// the slab lives in LDS; the Thomas sweep runs on wave 0 with its factor in registers (NZM
// compile-time) or, for NZM == 0 (any level count, read from d.nzm), in LDS after the slab.
// ------------------------------------------------------------------------------------------
template <int NX, int NZM, typename R, typename G = SegFor<NX>>
__global__ void CRM_LB k_press(DevT<R> d, int na, int nb, int nc) {
  CRM_STOP_CHECK(d);
  extern __shared__ double slab[];  // [64/NX instances][nzm][NX] == [nzm][64] per lane column
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane & (NX - 1);
  const int cg = blockIdx.x * (64 / NX) + lane / NX;
  const bool live = cg < d.ncrms;
  const int c = live ? cg : d.ncrms - 1;
  G s;
  const int nzm = NZM > 0 ? NZM : d.nzm, nz = nzm + 1;
  const double dz = inst(d, I_DZ, c);
  const double* adz = prof(d, P_ADZ, c);
  const double* adzw = prof(d, P_ADZW, c);
  const double* rho = prof(d, P_RHO, c);
  const double* rhow = prof(d, P_RHOW, c);
  auto P = [&](int k, int col) -> double& { return slab[(size_t)k * 64 + (lane - i) + col]; };
  const double rdx = 1. / d.dx, btat = d.bt / d.at, ctat = d.ct / d.at;
  const double dta = 1. / d.dt3na / d.at;
  const R *ua = d.dudt[na], *ub = d.dudt[nb], *uc = d.dudt[nc];
  const R *wa = d.dwdt[na], *wb = d.dwdt[nb], *wc = d.dwdt[nc];
  // 1. right-hand side
  for (int k = wave; k < nzm; k += 4) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    size_t ow = ((size_t)c * nz + k) * NX + i;
    size_t own = ow + NX;
    double rdz = 1. / (adz[k] * dz);
    double rup = rhow[k + 1] / rho[k] * rdz;
    double rdn = rhow[k] / rho[k] * rdz;
    double uu = d.u[o], uap = ua[o], ubp = ub[o], ucp = uc[o];
    double pv;
    if constexpr (G::Y3) {  // press_rhs.F90:45-69
      double vv = d.v[o], vap = d.dvdt[na][o], vbp = d.dvdt[nb][o], vcp = d.dvdt[nc][o];
      pv = (rdx * (s.rot(uu, 1) - uu) + rdx * (s.roty(vv, 1) - vv) + (d.w[own] * rup - d.w[ow] * rdn)) * dta +
           (rdx * (s.rot(uap, 1) - uap) + rdx * (s.roty(vap, 1) - vap) + (wa[own] * rup - wa[ow] * rdn)) +
           btat * (rdx * (s.rot(ubp, 1) - ubp) + rdx * (s.roty(vbp, 1) - vbp) + (wb[own] * rup - wb[ow] * rdn)) +
           ctat * (rdx * (s.rot(ucp, 1) - ucp) + rdx * (s.roty(vcp, 1) - vcp) + (wc[own] * rup - wc[ow] * rdn));
    } else {
      pv = (rdx * (s.rot(uu, 1) - uu) + (d.w[own] * rup - d.w[ow] * rdn)) * dta +
           (rdx * (s.rot(uap, 1) - uap) + (wa[own] * rup - wa[ow] * rdn)) +
           btat * (rdx * (s.rot(ubp, 1) - ubp) + (wb[own] * rup - wb[ow] * rdn)) +
           ctat * (rdx * (s.rot(ucp, 1) - ucp) + (wc[own] * rup - wc[ow] * rdn));
    }
    P(k, i) = pv * rho[k];
  }
  __syncthreads();
  // 2. forward transform: slot m of level k (every lane reads the whole row before any
  // lane overwrites its entry, hence the barrier between the sum and the store)
  for (int k0 = 0; k0 < nzm; k0 += 4) {
    const int k = k0 + wave;
    double acc = 0.0;
    if (k < nzm) {
#pragma unroll 8
      for (int q = 0; q < NX; q++) acc += d.Ft[q * NX + i] * P(k, q);
    }
    __syncthreads();
    if (k < nzm) P(k, i) = acc;
  }
  __syncthreads();
  // 3. Thomas sweep per slot (wave 0)
  if (wave == 0) {
    const double eig = d.eign[i];
    constexpr int NA = NZM > 0 ? NZM : 1;
    double alfa_r[NA];
    double* alfa_l = slab + (size_t)nzm * 64;
    auto alfa = [&](int k) -> double& {
      if constexpr (NZM > 0)
        return alfa_r[k];
      else
        return alfa_l[(size_t)k * 64 + lane];
    };
    auto acoef = [&](int k) { return rhow[k] / (adz[k] * adzw[k] * dz * dz); };
    auto ccoef = [&](int k) { return rhow[k + 1] / (adz[k] * adzw[k + 1] * dz * dz); };
    double b;
    if (i == 0)
      b = 1.0 / (eig * rho[0] - acoef(0) - ccoef(0));
    else
      b = 1.0 / (eig * rho[0] - ccoef(0));
    alfa(0) = -ccoef(0) * b;
    double beta = P(0, i) * b;
    P(0, i) = beta;
    auto down = [&](int k) {
      double ak = acoef(k), ck = ccoef(k);
      double e = 1.0 / (eig * rho[k] - ak - ck + ak * alfa(k - 1));
      alfa(k) = -ck * e;
      beta = (P(k, i) - ak * beta) * e;
      P(k, i) = beta;
    };
    // unrolled with a compile-time level count, so the factors stay in registers
    if constexpr (NZM > 0) {
#pragma unroll
      for (int k = 1; k < NZM - 1; k++) down(k);
    } else {
      for (int k = 1; k < nzm - 1; k++) down(k);
    }
    {
      int k = nzm - 1;
      double ak = acoef(k);
      double top = (P(k, i) - ak * beta) / (eig * rho[k] - ak + ak * alfa(k - 1));
      P(k, i) = top;
      auto up = [&](int kk) {
        top = alfa(kk) * top + P(kk, i);
        P(kk, i) = top;
      };
      if constexpr (NZM > 0) {
#pragma unroll
        for (int kk = NZM - 2; kk >= 0; kk--) up(kk);
      } else {
        for (int kk = nzm - 2; kk >= 0; kk--) up(kk);
      }
    }
  }
  __syncthreads();
  // 4. inverse transform (same read-all-then-store pattern)
  for (int k0 = 0; k0 < nzm; k0 += 4) {
    const int k = k0 + wave;
    double acc = 0.0;
    if (k < nzm) {
#pragma unroll 8
      for (int q = 0; q < NX; q++) acc += d.Gt[q * NX + i] * P(k, q);
    }
    __syncthreads();
    if (k < nzm) P(k, i) = acc;
  }
  __syncthreads();
  // 5. pressure gradient, then Adams-Bashforth and the half-step Courant velocities
  const double dtdx = d.dtn / d.dx, dtdz = d.dtn / dz, dt3 = d.dt3na;
  R* uan = d.dudt[na];
  R* wan = d.dwdt[na];
  R* unc = d.dudt[nc];
  R* wnc = d.dwdt[nc];
  for (int k = wave; k < nzm; k += 4) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    size_t ow = ((size_t)c * nz + k) * NX + i;
    double pk = P(k, i);
    double pim = s.rot(pk, -1);
    double pkb = P(k > 0 ? k - 1 : 0, i);
    double rdz = 1. / (dz * adzw[k]);
    double du_a = ua[o] - (pk - pim) * rdx;
    double dw_a = wa[ow] - (pk - pkb) * rdz;
    double uu = d.u[o], ww = d.w[ow];
    double un = uu + dt3 * (d.at * du_a + d.bt * ub[o] + d.ct * uc[o]);
    double wn = ww + dt3 * (d.at * dw_a + d.bt * wb[ow] + d.ct * wc[ow]);
    double dv_a = 0.0, vn = 0.0, vv = 0.0;
    if constexpr (G::Y3) {  // press_grad.F90:29, adams.F90:29,32
      double pjm = s.roty(pk, -1);
      dv_a = d.dvdt[na][o] - (pk - pjm) * rdx;
      vv = d.v[o];
      vn = vv + dt3 * (d.at * dv_a + d.bt * d.dvdt[nb][o] + d.ct * d.dvdt[nc][o]);
    }
    if (live) {
      uan[o] = du_a;
      wan[ow] = dw_a;
      unc[o] = un;
      wnc[ow] = wn;
      d.u[o] = 0.5 * (uu + un) * (rho[k] * dtdx);
      d.w[ow] = 0.5 * (ww + wn) * (rhow[k] * dtdz);
      if constexpr (G::Y3) {
        d.dvdt[na][o] = dv_a;
        d.dvdt[nc][o] = vn;
        d.v[o] = 0.5 * (vv + vn) * (rho[k] * dtdx);
      }
    }
  }
  if (wave == 0 && live) wnc[((size_t)c * nz + nzm) * NX + i] = 0.0;
}

// ------------------------------------------------------------------------------------------
// Pressure for 128-column CRMs: k_press's algorithm and operation order with one instance per
// block of 4 waves, as two groups of 128 columns to which the levels are dealt alternately.
// The slab [NZM][128] lives in LDS.  Column neighbours are read from memory (right-hand side)
// or from the slab (gradient), so no cross-wave exchange is needed; the Thomas sweep runs on
// group 0.
// ------------------------------------------------------------------------------------------
template <int NZM, typename R>
__global__ void CRM_LB k_press128(DevT<R> d, int na, int nb, int nc) {
  CRM_STOP_CHECK(d);
  constexpr int NX = 128;
  extern __shared__ double slab[];  // [nzm][128]
  const int i = threadIdx.x & (NX - 1), grp = threadIdx.x >> 7;
  const int c = blockIdx.x;  // one block per instance
  const int ip = (i + 1) & (NX - 1), im = (i - 1) & (NX - 1);
  const int nzm = NZM > 0 ? NZM : d.nzm, nz = nzm + 1;
  const double dz = inst(d, I_DZ, c);
  const double* adz = prof(d, P_ADZ, c);
  const double* adzw = prof(d, P_ADZW, c);
  const double* rho = prof(d, P_RHO, c);
  const double* rhow = prof(d, P_RHOW, c);
  auto P = [&](int k, int col) -> double& { return slab[(size_t)k * NX + col]; };
  const double rdx = 1. / d.dx, btat = d.bt / d.at, ctat = d.ct / d.at;
  const double dta = 1. / d.dt3na / d.at;
  const R *ua = d.dudt[na], *ub = d.dudt[nb], *uc = d.dudt[nc];
  const R *wa = d.dwdt[na], *wb = d.dwdt[nb], *wc = d.dwdt[nc];
  // 1. right-hand side
  for (int k = grp; k < nzm; k += 2) {
    const size_t row = ((size_t)c * nzm + k) * NX, o = row + i, op = row + ip;
    const size_t ow = ((size_t)c * nz + k) * NX + i, own = ow + NX;
    double rdz = 1. / (adz[k] * dz);
    double rup = rhow[k + 1] / rho[k] * rdz;
    double rdn = rhow[k] / rho[k] * rdz;
    double uu = d.u[o], uap = ua[o], ubp = ub[o], ucp = uc[o];
    double pv = (rdx * ((double)d.u[op] - uu) + (d.w[own] * rup - d.w[ow] * rdn)) * dta +
                (rdx * ((double)ua[op] - uap) + (wa[own] * rup - wa[ow] * rdn)) +
                btat * (rdx * ((double)ub[op] - ubp) + (wb[own] * rup - wb[ow] * rdn)) +
                ctat * (rdx * ((double)uc[op] - ucp) + (wc[own] * rup - wc[ow] * rdn));
    P(k, i) = pv * rho[k];
  }
  // one 128-point transform of every level of the slab in place, P(k, i) <- sum_q M[q][i]
  // P(k, q): each operator entry is read once and applied to all of the group's levels (an
  // accumulator per level in registers, each summed in q order as a per-level loop would);
  // every level is read before any is overwritten
  constexpr int NJ = (NZM > 0 ? NZM : 64) / 2 + 1;  // levels per group, an upper bound
  auto transform = [&](const double* __restrict__ M) {
    double acc[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) acc[j] = 0.0;
    __syncthreads();
    for (int q = 0; q < NX; q++) {
      const double m = M[q * NX + i];
#pragma unroll
      for (int j = 0; j < NJ; j++) {
        const int k = grp + 2 * j;
        if (k < nzm) acc[j] += m * P(k, q);
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NJ; j++) {
      const int k = grp + 2 * j;
      if (k < nzm) P(k, i) = acc[j];
    }
    __syncthreads();
  };
  // 2. forward transform
  transform(d.Ft);
  // 3. Thomas sweep per slot (group 0)
  if (grp == 0) {
    const double eig = d.eign[i];
    constexpr int NA = NZM > 0 ? NZM : 1;
    double alfa_r[NA];
    double* alfa_l = slab + (size_t)nzm * NX;  // NZM == 0: the factors in LDS after the slab
    auto alfa = [&](int k) -> double& {
      if constexpr (NZM > 0)
        return alfa_r[k];
      else
        return alfa_l[(size_t)k * NX + i];
    };
    auto acoef = [&](int k) { return rhow[k] / (adz[k] * adzw[k] * dz * dz); };
    auto ccoef = [&](int k) { return rhow[k + 1] / (adz[k] * adzw[k + 1] * dz * dz); };
    double b;
    if (i == 0)
      b = 1.0 / (eig * rho[0] - acoef(0) - ccoef(0));
    else
      b = 1.0 / (eig * rho[0] - ccoef(0));
    alfa(0) = -ccoef(0) * b;
    double beta = P(0, i) * b;
    P(0, i) = beta;
    auto down = [&](int k) {
      double ak = acoef(k), ck = ccoef(k);
      double e = 1.0 / (eig * rho[k] - ak - ck + ak * alfa(k - 1));
      alfa(k) = -ck * e;
      beta = (P(k, i) - ak * beta) * e;
      P(k, i) = beta;
    };
    // unrolled with a compile-time level count, so the factors stay in registers
    if constexpr (NZM > 0) {
#pragma unroll
      for (int k = 1; k < NZM - 1; k++) down(k);
    } else {
      for (int k = 1; k < nzm - 1; k++) down(k);
    }
    {
      int k = nzm - 1;
      double ak = acoef(k);
      double top = (P(k, i) - ak * beta) / (eig * rho[k] - ak + ak * alfa(k - 1));
      P(k, i) = top;
      auto up = [&](int kk) {
        top = alfa(kk) * top + P(kk, i);
        P(kk, i) = top;
      };
      if constexpr (NZM > 0) {
#pragma unroll
        for (int kk = NZM - 2; kk >= 0; kk--) up(kk);
      } else {
        for (int kk = nzm - 2; kk >= 0; kk--) up(kk);
      }
    }
  }
  // 4. inverse transform (its leading barrier also ends the Thomas sweep)
  transform(d.Gt);
  // 5. pressure gradient, then Adams-Bashforth and the half-step Courant velocities
  const double dtdx = d.dtn / d.dx, dtdz = d.dtn / dz, dt3 = d.dt3na;
  R* uan = d.dudt[na];
  R* wan = d.dwdt[na];
  R* unc = d.dudt[nc];
  R* wnc = d.dwdt[nc];
  for (int k = grp; k < nzm; k += 2) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    size_t ow = ((size_t)c * nz + k) * NX + i;
    double pk = P(k, i);
    double pim = P(k, im);
    double pkb = P(k > 0 ? k - 1 : 0, i);
    double rdz = 1. / (dz * adzw[k]);
    double du_a = ua[o] - (pk - pim) * rdx;
    double dw_a = wa[ow] - (pk - pkb) * rdz;
    double uu = d.u[o], ww = d.w[ow];
    double un = uu + dt3 * (d.at * du_a + d.bt * ub[o] + d.ct * uc[o]);
    double wn = ww + dt3 * (d.at * dw_a + d.bt * wb[ow] + d.ct * wc[ow]);
    uan[o] = du_a;
    wan[ow] = dw_a;
    unc[o] = un;
    wnc[ow] = wn;
    d.u[o] = 0.5 * (uu + un) * (rho[k] * dtdx);
    d.w[ow] = 0.5 * (ww + wn) * (rhow[k] * dtdz);
  }
  if (grp == 0) wnc[((size_t)c * nz + nzm) * NX + i] = 0.0;
}

// ------------------------------------------------------------------------------------------
// Pressure for 64-column CRMs with the x-transforms on the matrix cores.  Same algorithm and
// operation order per level as k_press; the two transforms are the GEMMs
//   C[k][s] = sum_i P[k][i] Ft[i][s]   and   P'[k][i] = sum_s C[k][s] Gt[s][i]
// (levels padded to 64 rows) done with v_mfma_f64_16x16x4f64: wave w owns output columns
// 16w..16w+15 and keeps the matching 16 k-step fragments of Ft (then Gt) in registers; the
// A fragments stream from the LDS slab.  One instance per block of 4 waves.
//
// MFMA f64 16x16x4 operand layout (lane l of the wave):
//   A[16x4]: row l%16, col l/16      B[4x16]: row l/16, col l%16
//   D[16x16]: element j of the result vector is row 4*j + l/16, col l%16
// ------------------------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int PLD = 66;  // slab row stride in doubles: conflict-free A-fragment reads

// UNR: levels of the right-hand side and gradient passes unrolled together, so that their
// loads are in flight at once (a launch variant the autotuner may choose).  RALF: the Thomas
// factors in registers of the sweeping wave instead of LDS (NZM > 0 only), so that the block
// needs only the slab and three blocks fit a CU instead of two (another launch variant).
template <int NZM, typename R, typename G = Seg<64>, int UNR = 1, bool RALF = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RALF ? 3 : CRM_WAVES_PER_EU)))
k_press64(DevT<R> d, int na, int nb, int nc) {
  CRM_STOP_CHECK(d);
  static_assert(!RALF || NZM > 0, "factors in registers: compile-time level count");
  constexpr int NX = 64, NK = 64;
  extern __shared__ double lds[];
  const int c = blockIdx.x;  // grid is exactly ncrms blocks
  double* P = lds;              // [NK][PLD] slab: rhs, coefficients, pressure
  double* ALF = lds + NK * PLD; // [nzm][64] Thomas factors (not allocated with RALF)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane;
  G s;  // 3D (8 x 8 columns): the transforms are the x-y Kronecker operators, dy == dx
  const int nzm = NZM > 0 ? NZM : d.nzm, nz = nzm + 1;
  const double dz = inst(d, I_DZ, c);
  const double* adz = prof(d, P_ADZ, c);
  const double* adzw = prof(d, P_ADZW, c);
  const double* rho = prof(d, P_RHO, c);
  const double* rhow = prof(d, P_RHOW, c);
  const double rdx = 1. / d.dx, btat = d.bt / d.at, ctat = d.ct / d.at;
  const double dta = 1. / d.dt3na / d.at;
  const R *ua = d.dudt[na], *ub = d.dudt[nb], *uc = d.dudt[nc];
  const R *wa = d.dwdt[na], *wb = d.dwdt[nb], *wc = d.dwdt[nc];
  const int r16 = lane & 15, q4 = lane >> 4;
  // 1. right-hand side (rows >= nzm of the slab are zero).  The loads of UNR levels are
  // issued before any of them is used (2D), so that they are in flight together.
  static_assert(UNR == 1 || !G::Y3, "levels per trip: 2D only");
  for (int k0 = wave; k0 < NK; k0 += 4 * UNR) {
    double ld_[UNR][12];
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      const int k = k0 + 4 * q;
      if (!G::Y3 && k < nzm) {
        const size_t o = ((size_t)c * nzm + k) * NX + i, ow = ((size_t)c * nz + k) * NX + i, own = ow + NX;
        ld_[q][0] = d.u[o];
        ld_[q][1] = ua[o];
        ld_[q][2] = ub[o];
        ld_[q][3] = uc[o];
        ld_[q][4] = d.w[ow];
        ld_[q][5] = wa[ow];
        ld_[q][6] = wb[ow];
        ld_[q][7] = wc[ow];
        ld_[q][8] = d.w[own];
        ld_[q][9] = wa[own];
        ld_[q][10] = wb[own];
        ld_[q][11] = wc[own];
      }
    }
#pragma unroll
    for (int q = 0; q < UNR; q++) {
    const int k = k0 + 4 * q;
    double pv = 0.0;
    if (k < nzm) {
      size_t o = ((size_t)c * nzm + k) * NX + i;
      size_t ow = ((size_t)c * nz + k) * NX + i;
      size_t own = ow + NX;
      double rdz = 1. / (adz[k] * dz);
      double rup = rhow[k + 1] / rho[k] * rdz;
      double rdn = rhow[k] / rho[k] * rdz;
      if constexpr (G::Y3) {  // press_rhs.F90:45-69
        double uu = d.u[o], uap = ua[o], ubp = ub[o], ucp = uc[o];
        double vv = d.v[o], vap = d.dvdt[na][o], vbp = d.dvdt[nb][o], vcp = d.dvdt[nc][o];
        pv = (rdx * (s.rot(uu, 1) - uu) + rdx * (s.roty(vv, 1) - vv) + (d.w[own] * rup - d.w[ow] * rdn)) * dta +
             (rdx * (s.rot(uap, 1) - uap) + rdx * (s.roty(vap, 1) - vap) + (wa[own] * rup - wa[ow] * rdn)) +
             btat * (rdx * (s.rot(ubp, 1) - ubp) + rdx * (s.roty(vbp, 1) - vbp) + (wb[own] * rup - wb[ow] * rdn)) +
             ctat * (rdx * (s.rot(ucp, 1) - ucp) + rdx * (s.roty(vcp, 1) - vcp) + (wc[own] * rup - wc[ow] * rdn));
      } else {
        const double uu = ld_[q][0], uap = ld_[q][1], ubp = ld_[q][2], ucp = ld_[q][3];
        pv = (rdx * (s.rot(uu, 1) - uu) + (ld_[q][8] * rup - ld_[q][4] * rdn)) * dta +
             (rdx * (s.rot(uap, 1) - uap) + (ld_[q][9] * rup - ld_[q][5] * rdn)) +
             btat * (rdx * (s.rot(ubp, 1) - ubp) + (ld_[q][10] * rup - ld_[q][6] * rdn)) +
             ctat * (rdx * (s.rot(ucp, 1) - ucp) + (ld_[q][11] * rup - ld_[q][7] * rdn));
      }
      pv = pv * rho[k];
    }
    if (k < NK) P[k * PLD + i] = pv;
    }
  }
  // one 64x64x64 transform of the slab in place: P <- P * M (M row-major 64x64)
  auto transform = [&](const double* M) {
    double bf[16];
#pragma unroll
    for (int j = 0; j < 16; j++) bf[j] = M[(4 * j + q4) * NX + 16 * wave + r16];
    v4d acc[4];
#pragma unroll
    for (int r = 0; r < 4; r++) acc[r] = v4d{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double a = P[(16 * r + r16) * PLD + 4 * j + q4];
        acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bf[j], acc[r], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int jj = 0; jj < 4; jj++) P[(16 * r + 4 * jj + q4) * PLD + 16 * wave + r16] = acc[r][jj];
    __syncthreads();
  };
  // 2. forward transform
  transform(d.Ft);
  // 3. Thomas sweep per slot (wave 0, lane = slot), factors in LDS
  if (wave == 0) {
    static_assert(NZM <= 64, "one level per lane");  // NZM == 0: the host admits nzm <= 64
    const double eig = d.eign[i];
    // the level coefficients do not depend on the recurrence: lane k computes level k's,
    // and the sweep broadcasts them (the sweep then waits on one division per level, not
    // three plus the profile loads)
    const int kl = i < nzm ? i : nzm - 1;
    const double ac_l = rhow[kl] / (adz[kl] * adzw[kl] * dz * dz);
    const double cc_l = rhow[kl + 1] / (adz[kl] * adzw[kl + 1] * dz * dz);
    const double rho_l = rho[kl];
    auto lane_of = [](double v, int k) {
      const long long b = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_readlane((int)b, k), hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
      return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    };
    auto acoef = [&](int k) { return lane_of(ac_l, k); };
    auto ccoef = [&](int k) { return lane_of(cc_l, k); };
    double b;
    if (i == 0)
      b = 1.0 / (eig * rho[0] - acoef(0) - ccoef(0));
    else
      b = 1.0 / (eig * rho[0] - ccoef(0));
    // the factors: LDS, or (RALF) an array the unrolled sweeps index with constants, which
    // the compiler keeps in registers
    double alf_r[RALF ? (NZM > 0 ? NZM : 1) : 1];
    auto put_alf = [&](int k, double v) {
      if constexpr (RALF)
        alf_r[k] = v;
      else
        ALF[k * 64 + i] = v;
    };
    auto get_alf = [&](int k) {
      if constexpr (RALF)
        return alf_r[k];
      else
        return ALF[k * 64 + i];
    };
    double alfa = -ccoef(0) * b;
    put_alf(0, alfa);
    double beta = P[i] * b;
    P[i] = beta;
    auto fwd = [&](int k) {
      double ak = acoef(k), ck = ccoef(k);
      double e = 1.0 / (eig * lane_of(rho_l, k) - ak - ck + ak * alfa);
      alfa = -ck * e;
      put_alf(k, alfa);
      beta = (P[k * PLD + i] - ak * beta) * e;
      P[k * PLD + i] = beta;
    };
    if constexpr (RALF) {
      // the same operations as fwd, with the terms that do not depend on the recurrence (the
      // coefficient broadcasts and the slab read) fetched one level ahead, so that they
      // overlap the previous level's division instead of waiting behind it
      double ak_n = acoef(1), ck_n = ccoef(1), rk_n = lane_of(rho_l, 1), pk_n = P[PLD + i];
#pragma unroll
      for (int k = 1; k < NZM - 1; k++) {
        const double ak = ak_n, ck = ck_n, rk = rk_n, pk = pk_n;
        if (k + 1 < NZM - 1) {
          ak_n = acoef(k + 1);
          ck_n = ccoef(k + 1);
          rk_n = lane_of(rho_l, k + 1);
          pk_n = P[(k + 1) * PLD + i];
        }
        double e = 1.0 / (eig * rk - ak - ck + ak * alfa);
        alfa = -ck * e;
        put_alf(k, alfa);
        beta = (pk - ak * beta) * e;
        P[k * PLD + i] = beta;
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      for (int k = 1; k < nzm - 1; k++) fwd(k);
    }
    {
      int k = nzm - 1;
      double ak = acoef(k);
      double top = (P[k * PLD + i] - ak * beta) / (eig * rho[k] - ak + ak * alfa);
      P[k * PLD + i] = top;
      auto bwd = [&](int kk) {
        top = get_alf(kk) * top + P[kk * PLD + i];
        P[kk * PLD + i] = top;
      };
      if constexpr (RALF) {
        // bwd with the slab read one level ahead
        double pn = P[(NZM - 2) * PLD + i];
#pragma unroll
        for (int kk = NZM - 2; kk >= 0; kk--) {
          const double pk = pn;
          if (kk > 0) pn = P[(kk - 1) * PLD + i];
          top = get_alf(kk) * top + pk;
          P[kk * PLD + i] = top;
          __builtin_amdgcn_sched_barrier(0);
        }
      } else {
        for (int kk = nzm - 2; kk >= 0; kk--) bwd(kk);
      }
    }
  }
  // 4. inverse transform (its leading barrier also ends the Thomas sweep)
  transform(d.Gt);
  // 5. pressure gradient, then Adams-Bashforth and the half-step Courant velocities
  const double dtdx = d.dtn / d.dx, dtdz = d.dtn / dz, dt3 = d.dt3na;
  R* uan = d.dudt[na];
  R* wan = d.dwdt[na];
  R* unc = d.dudt[nc];
  R* wnc = d.dwdt[nc];
  for (int k0 = wave; k0 < nzm; k0 += 4 * UNR) {
    double lg_[UNR][8];
#pragma unroll
    for (int q = 0; q < UNR; q++) {
      const int k = k0 + 4 * q;
      if (k < nzm) {
        const size_t o = ((size_t)c * nzm + k) * NX + i, ow = ((size_t)c * nz + k) * NX + i;
        lg_[q][0] = ua[o];
        lg_[q][1] = wa[ow];
        lg_[q][2] = d.u[o];
        lg_[q][3] = d.w[ow];
        lg_[q][4] = ub[o];
        lg_[q][5] = uc[o];
        lg_[q][6] = wb[ow];
        lg_[q][7] = wc[ow];
      }
    }
#pragma unroll
    for (int q = 0; q < UNR; q++) {
    const int k = k0 + 4 * q;
    if (k >= nzm) break;
    size_t o = ((size_t)c * nzm + k) * NX + i;
    size_t ow = ((size_t)c * nz + k) * NX + i;
    double pk = P[k * PLD + i];
    double pim = s.rot(pk, -1);
    double pkb = P[(k > 0 ? k - 1 : 0) * PLD + i];
    double rdz = 1. / (dz * adzw[k]);
    double du_a = lg_[q][0] - (pk - pim) * rdx;
    double dw_a = lg_[q][1] - (pk - pkb) * rdz;
    double uu = lg_[q][2], ww = lg_[q][3];
    // the b and c terms of the Adams-Bashforth step are summed first, as k_press64r keeps
    // that sum in registers: both kernels give the same bits
    double un = uu + dt3 * (d.at * du_a + (d.bt * lg_[q][4] + d.ct * lg_[q][5]));
    double wn = ww + dt3 * (d.at * dw_a + (d.bt * lg_[q][6] + d.ct * lg_[q][7]));
    if constexpr (G::Y3) {  // press_grad.F90:29, adams.F90:29,32
      double pjm = s.roty(pk, -1);
      double dv_a = d.dvdt[na][o] - (pk - pjm) * rdx;
      double vv = d.v[o];
      double vn = vv + dt3 * (d.at * dv_a + d.bt * d.dvdt[nb][o] + d.ct * d.dvdt[nc][o]);
      d.dvdt[na][o] = dv_a;
      d.dvdt[nc][o] = vn;
      d.v[o] = 0.5 * (vv + vn) * (rho[k] * dtdx);
    }
    uan[o] = du_a;
    wan[ow] = dw_a;
    unc[o] = un;
    wnc[ow] = wn;
    d.u[o] = 0.5 * (uu + un) * (rho[k] * dtdx);
    d.w[ow] = 0.5 * (ww + wn) * (rhow[k] * dtdz);
    }
  }
  if (wave == 0) wnc[((size_t)c * nz + nzm) * NX + i] = 0.0;
}

// k_press64 for 2D CRMs of 64 columns on 8 waves (one instance per 512-thread workgroup):
// wave w owns the slab rows k = w + 8 j.  The right-hand side pass keeps, per row, what the
// gradient pass needs (u, w, their tendencies dudt(na), dwdt(na), and the b/c terms of the
// Adams-Bashforth step) in registers, so the velocities and the Adams-Bashforth history are
// read from memory once per launch instead of twice.  Eight waves of one workgroup fill a CU
// as two four-wave workgroups of k_press64 do.
template <int NZM, typename R>
__global__ void __launch_bounds__(512) k_press64r(DevT<R> d, int na, int nb, int nc) {
  CRM_STOP_CHECK(d);
  constexpr int NX = 64, NK = 64, NW = 8;
  constexpr int NJ = (NZM + NW - 1) / NW;  // rows of the level range per wave
  extern __shared__ double lds[];
  double* P = lds;              // [NK][PLD] slab: rhs, coefficients, pressure
  double* ALF = lds + NK * PLD; // [NZM][64] Thomas factors
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = blockIdx.x;  // grid is exactly ncrms blocks
  const int i = lane;
  Seg<64> s;
  const int nz = NZM + 1;
  const double dz = inst(d, I_DZ, c);
  const double* adz = prof(d, P_ADZ, c);
  const double* adzw = prof(d, P_ADZW, c);
  const double* rho = prof(d, P_RHO, c);
  const double* rhow = prof(d, P_RHOW, c);
  const double rdx = 1. / d.dx, btat = d.bt / d.at, ctat = d.ct / d.at;
  const double dta = 1. / d.dt3na / d.at;
  const R *ua = d.dudt[na], *ub = d.dudt[nb], *uc = d.dudt[nc];
  const R *wa = d.dwdt[na], *wb = d.dwdt[nb], *wc = d.dwdt[nc];
  const int r16 = lane & 15, q4 = lane >> 4;
  // the gradient pass's inputs of each owned level
  double cu[NJ], cua[NJ], cub[NJ], cw[NJ], cwa[NJ], cwb[NJ];
  // 1. right-hand side (rows >= NZM of the slab are zero)
#pragma unroll
  for (int j = 0; j < NK / NW; j++) {
    const int k = wave + NW * j;
    double pv = 0.0;
    if (k < NZM) {
      size_t o = ((size_t)c * NZM + k) * NX + i;
      size_t ow = ((size_t)c * nz + k) * NX + i;
      size_t own = ow + NX;
      double rdz = 1. / (adz[k] * dz);
      double rup = rhow[k + 1] / rho[k] * rdz;
      double rdn = rhow[k] / rho[k] * rdz;
      double uu = d.u[o], uap = ua[o], ubp = ub[o], ucp = uc[o];
      double ww = d.w[ow], wap = wa[ow], wbp = wb[ow], wcp = wc[ow];
      pv = (rdx * (s.rot(uu, 1) - uu) + (d.w[own] * rup - ww * rdn)) * dta +
           (rdx * (s.rot(uap, 1) - uap) + (wa[own] * rup - wap * rdn)) +
           btat * (rdx * (s.rot(ubp, 1) - ubp) + (wb[own] * rup - wbp * rdn)) +
           ctat * (rdx * (s.rot(ucp, 1) - ucp) + (wc[own] * rup - wcp * rdn));
      pv = pv * rho[k];
      if (j < NJ) {
        cu[j] = uu;
        cua[j] = uap;
        cub[j] = d.bt * ubp + d.ct * ucp;
        cw[j] = ww;
        cwa[j] = wap;
        cwb[j] = d.bt * wbp + d.ct * wcp;
      }
    }
    P[k * PLD + i] = pv;
  }
  // one 64x64x64 transform of the slab in place: P <- P * M (M row-major 64x64); wave w makes
  // output columns 16 (w & 3) .. +15 of row tiles 2 (w >> 2) and 2 (w >> 2) + 1
  auto transform = [&](const double* M) {
    const int cb = wave & 3, rb = (wave >> 2) * 2;
    double bf[16];
#pragma unroll
    for (int j = 0; j < 16; j++) bf[j] = M[(4 * j + q4) * NX + 16 * cb + r16];
    v4d acc[2];
#pragma unroll
    for (int r = 0; r < 2; r++) acc[r] = v4d{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) {
#pragma unroll
      for (int r = 0; r < 2; r++) {
        double a = P[(16 * (rb + r) + r16) * PLD + 4 * j + q4];
        acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bf[j], acc[r], 0, 0, 0);
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
      for (int jj = 0; jj < 4; jj++) P[(16 * (rb + r) + 4 * jj + q4) * PLD + 16 * cb + r16] = acc[r][jj];
    __syncthreads();
  };
  // 2. forward transform
  transform(d.Ft);
  // 3. Thomas sweep per slot (wave 0, lane = slot), factors in LDS; as in k_press64
  if (wave == 0) {
    static_assert(NZM <= 64, "one level per lane");
    const double eig = d.eign[i];
    const int kl = i < NZM ? i : NZM - 1;
    const double ac_l = rhow[kl] / (adz[kl] * adzw[kl] * dz * dz);
    const double cc_l = rhow[kl + 1] / (adz[kl] * adzw[kl + 1] * dz * dz);
    const double rho_l = rho[kl];
    auto lane_of = [](double v, int k) {
      const long long b = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_readlane((int)b, k), hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
      return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    };
    auto acoef = [&](int k) { return lane_of(ac_l, k); };
    auto ccoef = [&](int k) { return lane_of(cc_l, k); };
    double b;
    if (i == 0)
      b = 1.0 / (eig * rho[0] - acoef(0) - ccoef(0));
    else
      b = 1.0 / (eig * rho[0] - ccoef(0));
    double alfa = -ccoef(0) * b;
    ALF[i] = alfa;
    double beta = P[i] * b;
    P[i] = beta;
    for (int k = 1; k < NZM - 1; k++) {
      double ak = acoef(k), ck = ccoef(k);
      double e = 1.0 / (eig * lane_of(rho_l, k) - ak - ck + ak * alfa);
      alfa = -ck * e;
      ALF[k * 64 + i] = alfa;
      beta = (P[k * PLD + i] - ak * beta) * e;
      P[k * PLD + i] = beta;
    }
    {
      int k = NZM - 1;
      double ak = acoef(k);
      double top = (P[k * PLD + i] - ak * beta) / (eig * rho[k] - ak + ak * alfa);
      P[k * PLD + i] = top;
      for (int kk = NZM - 2; kk >= 0; kk--) {
        top = ALF[kk * 64 + i] * top + P[kk * PLD + i];
        P[kk * PLD + i] = top;
      }
    }
  }
  // 4. inverse transform (its leading barrier also ends the Thomas sweep)
  transform(d.Gt);
  // 5. pressure gradient, then Adams-Bashforth and the half-step Courant velocities, from the
  // registers of pass 1
  const double dtdx = d.dtn / d.dx, dtdz = d.dtn / dz, dt3 = d.dt3na;
  R* uan = d.dudt[na];
  R* wan = d.dwdt[na];
  R* unc = d.dudt[nc];
  R* wnc = d.dwdt[nc];
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int k = wave + NW * j;
    if (k < NZM) {
      size_t o = ((size_t)c * NZM + k) * NX + i;
      size_t ow = ((size_t)c * nz + k) * NX + i;
      double pk = P[k * PLD + i];
      double pim = s.rot(pk, -1);
      double pkb = P[(k > 0 ? k - 1 : 0) * PLD + i];
      double rdz = 1. / (dz * adzw[k]);
      double du_a = cua[j] - (pk - pim) * rdx;
      double dw_a = cwa[j] - (pk - pkb) * rdz;
      double uu = cu[j], ww = cw[j];
      double un = uu + dt3 * (d.at * du_a + cub[j]);
      double wn = ww + dt3 * (d.at * dw_a + cwb[j]);
      uan[o] = du_a;
      wan[ow] = dw_a;
      unc[o] = un;
      wnc[ow] = wn;
      d.u[o] = 0.5 * (uu + un) * (rho[k] * dtdx);
      d.w[ow] = 0.5 * (ww + wn) * (rhow[k] * dtdz);
    }
  }
  if (wave == 0) wnc[((size_t)c * nz + NZM) * NX + i] = 0.0;
}

template <int NZM>
constexpr size_t press64_lds_bytes() {  // NZM == 0: room for the most levels, 64
  return (size_t)(64 * PLD + (NZM > 0 ? NZM : 64) * 64) * sizeof(double);
}
constexpr size_t press64_slab_bytes() {  // the slab alone (Thomas factors in registers)
  return (size_t)64 * PLD * sizeof(double);
}

// ------------------------------------------------------------------------------------------
// MPDATA scalar advection: ADV_MPDATA/advect_scalar2D.F90:8-244 with the statistics of
// advect_scalar.F90:9-76.  Software-pipelined over levels: stage s processes level K-s.
// ------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T andiff(T x1, T x2, T a, T b) {
  return (fabs(a) - a * a * b) * (T)0.5 * (x2 - x1);
}
template <typename T>
__device__ __forceinline__ T across(T x1, T a1, T a2) { return (T)0.03125 * a1 * a2 * x1; }
template <typename T>
__device__ __forceinline__ T ppos(T y) { return fmax((T)0.0, y); }
template <typename T>
__device__ __forceinline__ T pneg(T y) { return -fmin((T)0.0, y); }

// terminal velocity of precipitation, microphysics.F90:453-486.  The three regimes (all rain,
// no rain, mixed) share their powers: each of the rain, graupel and snow powers is taken once,
// by the lanes whose regime uses it, so a wave whose lanes straddle the regimes evaluates at
// most three powers instead of six.  The arguments are the reference's in every regime (with
// omp == 0, q - omp*q is q exactly), and a power a regime does not use multiplies a zero
// weight there, so each lane's velocity is the one its own regime's formula gives.
template <typename T>
__device__ __forceinline__ T term_vel(const TermVel& tv, T q, T rho, T tabs) {
  T v = 0;
  if (q > QP_THRESHOLD) {
    const T omp = fmax(T(0.0), fmin(T(1.0), (tabs - TPRMIN) * (T)A_PR()));
    const T omg = fmax(T(0.0), fmin(T(1.0), (tabs - TGRMIN) * (T)A_GR()));
    const T qrr = omp * q, qsg = q - qrr, qgg = omg * qsg, qss = qsg - qgg;
    T pr = 0, pg = 0, ps = 0;
    if (omp > T(0.)) pr = fpow_fall<T>(rho * qrr, (T)tv.crain);
    if (omp < T(1.) && omg > T(0.)) pg = fpow_fall<T>(rho * qgg, (T)tv.cgrau);
    if (omp < T(1.) && omg < T(1.)) ps = fpow_fall<T>(rho * qss, (T)tv.csnow);
    const T vsg = omg * (T)tv.vgrau * pg + (T(1.) - omg) * (T)tv.vsnow * ps;
    if (omp == T(1.))
      v = (T)tv.vrain * pr;
    else if (omp == T(0.))
      v = vsg;
    else
      v = omp * (T)tv.vrain * pr + (T(1.) - omp) * vsg;
  }
  return v;
}

// An upper bound, to within a relative 1e-5, of term_vel's FP64 value, with the three powers in
// FP32: the regime weights are term_vel's own (FP64), and each power's argument is raised to at
// least 1e-30 first.  x**c is increasing in x for c > 0, so that can only raise the velocity, and
// it keeps FP32's underflow of a tiny argument from turning a small velocity into zero.  The
// FP32 powers (hardware base-2 logarithm and exponential) are within ~1e-6 relative here.
__device__ __forceinline__ float term_vel_bound(const TermVel& tv, double q, double rho, double tabs) {
  float v = 0.f;
  if (q > QP_THRESHOLD) {
    const double omp = fmax(0.0, fmin(1.0, (tabs - TPRMIN) * A_PR()));
    const double omg = fmax(0.0, fmin(1.0, (tabs - TGRMIN) * A_GR()));
    const double qrr = omp * q, qsg = q - qrr, qgg = omg * qsg, qss = qsg - qgg;
    auto pw = [](double x, double c) { return fpow<float>((float)fmax(x, 1e-30), (float)c); };
    float pr = 0.f, pg = 0.f, ps = 0.f;
    if (omp > 0.) pr = pw(rho * qrr, tv.crain);
    if (omp < 1. && omg > 0.) pg = pw(rho * qgg, tv.cgrau);
    if (omp < 1. && omg < 1.) ps = pw(rho * qss, tv.csnow);
    const float fomp = (float)omp, fomg = (float)omg;
    const float vsg = fomg * (float)tv.vgrau * pg + (1.f - fomg) * (float)tv.vsnow * ps;
    v = fomp * (float)tv.vrain * pr + (1.f - fomp) * vsg;
  }
  return v;
}

// One scalar of MPDATA, software-pipelined over levels.  At iteration K stage s works on
// level K-s; ring variables are named by how far behind the front they are at the top of the
// iteration.  Statistics: fadv (advect_scalar.F90:62-72) and the vertical flux of
// advect_scalar2D.F90:113-118,204-206.  With want_cfl, also the precipitation Courant number
// of precip_fall (microphysics.F90:596-611) on the advected field.
// PAIR: the steady state advances two levels per trip (the launch variant the autotuner may
// choose; the default for the 2D 64-column segment)
template <int NX, typename R, typename G, bool PAIR>
__device__ void mpdata_one(const DevT<R>& d, const G& s, int c, bool live, R* f, double* fadv,
                           double* flux, bool want_cfl) {
  constexpr bool Y3 = G::Y3;  // 3D: advect_scalar3D.F90:8-349 (dy == dx)
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const cprof<R> adz = profc(d, P_ADZ, c);
  const cprof<R> rho = profc(d, P_RHO, c);
  const cprof<R> irho = profc(d, P_IRHO, c);
  const cprof<R> iadz = profc(d, P_IADZ, c);
  const cprof<R> irhowadz = profc(d, P_IRHOWADZ, c);
  const cprof<R> ddadz = profc(d, P_DDADZ, c);
  const cprof<R> pratio = profc(d, P_PRATIO, c);
  const R eps = R4(1.e-10);
  auto F = [&](const R* p, int k) { return p[((size_t)c * nzm + k) * NX + i]; };
  auto Wv = [&](int k) { return d.w[((size_t)c * nz + k) * NX + i]; };
  const R dz = inst(d, I_DZ, c);
  R pcfl = R(0.0);
  R f_0 = 0, f_1 = 0, f_2 = 0, f_3 = 0, f_4 = 0, f_5 = 0;
  // inputs of the next two levels, loaded two iterations before use
  R nf1 = F(f, 0), nu1 = F(d.u, 0), nw1 = Wv(0);
  R nf2 = F(f, 1), nu2 = F(d.u, 1), nw2 = Wv(1);
  R nv1 = 0, nv2 = 0;
  if constexpr (Y3) {
    nv1 = F(d.v, 0);
    nv2 = F(d.v, 1);
  }
  R V_0 = 0, V_1 = 0, V_2 = 0, V_3 = 0;
  R vv1_0 = 0, vv1_1 = 0, vv2_a = 0, vv2_b = 0, vv3_a = 0;
  // tabs for the precipitation CFL of stage 5, loaded two iterations before use
  R tq_a = 0, tq_b = 0;
  R U_0 = 0, U_1 = 0, U_2 = 0, U_3 = 0;
  R W_0 = 0, W_1 = 0, W_2 = 0;
  R uu1_0 = 0, uu1_1 = 0, ww1_0 = 0, ww1_1 = 0;
  R f1_a = 0, f1_b = 0, f1_c = 0, f1_d = 0;
  R mx0_a = 0, mn0_a = 0;
  R uu2_a = 0, uu2_b = 0, ww2_a = 0, ww2_b = 0;
  R mx1_a = 0, mn1_a = 0;
  R mx_a = 0, mx_b = 0, mn_a = 0, mn_b = 0;
  R uu3_a = 0, ww3_a = 0;
  // the donor-cell part of the vertical flux statistic from stage 1, kept until stage 4 adds
  // the antidiffusive part three levels later (reading it back from memory would wait for the
  // prefetch issued before it: vector loads complete in issue order)
  R s1_a = 0, s1_b = 0, s1_c = 0;
  // One pipeline step; ST: steady state (5 <= K < nzm), where every stage is active and every
  // boundary case is the interior one, so the guards and level selects fold away.
  auto step = [&](const int K, auto steady) {
    constexpr bool ST = decltype(steady)::value;
    const R cf = nf1, cu = nu1, cw = nw1, cv = nv1;
    nf1 = nf2;
    nu1 = nu2;
    nw1 = nw2;
    nv1 = nv2;
    // unconditional, clamped at the top level: past it the ring already holds that level's
    // values, so the reload changes nothing, and a load that always issues can write the ring
    // register directly (a conditional one is preceded by a copy that waits for earlier loads)
    {
      const int kn = ::min(K + 2, nzm - 1);
      nf2 = F(f, kn);
      nu2 = F(d.u, kn);
      nw2 = Wv(kn);
      if constexpr (Y3) nv2 = F(d.v, kn);
    }
    R tq_new = R(0.0);
    if (want_cfl && (ST || (K >= 3 && K - 3 < nzm))) tq_new = F(d.tabs, K - 3);
    // stage 0: donor-cell fluxes at level K
    if (ST || K < nzm) {
      f_0 = cf;
      U_0 = cu;
      W_0 = cw;
      R fkb = (ST || K >= 1) ? f_1 : f_0;
      uu1_0 = fmax(R(0.0), U_0) * s.rot(f_0, -1) + fmin(R(0.0), U_0) * f_0;
      ww1_0 = fmax(R(0.0), W_0) * fkb + fmin(R(0.0), W_0) * f_0;
      if constexpr (Y3) {
        V_0 = cv;
        vv1_0 = fmax(R(0.0), V_0) * s.roty(f_0, -1) + fmin(R(0.0), V_0) * f_0;
      }
    } else {
      f_0 = U_0 = W_0 = uu1_0 = ww1_0 = R(0.0);
      V_0 = vv1_0 = R(0.0);
    }
    // stage 1: donor-cell update and first bounds at level K-1
    const int L1 = K - 1;
    R f1_new = 0, mx0_new = 0, mn0_new = 0, s1_new = 0;
    if (ST || (L1 >= 0 && L1 < nzm)) {
      R fkb = (ST || L1 >= 1) ? f_2 : f_1, fkc = (ST || L1 + 1 < nzm) ? f_0 : f_1;
      R fm = s.rot(f_1, -1), fp = s.rot(f_1, 1);
      if constexpr (Y3) {
        f1_new = f_1 - (s.rot(uu1_1, 1) - uu1_1 + s.roty(vv1_1, 1) - vv1_1 + (ww1_0 - ww1_1) * iadz[L1]) * irho[L1];
        R fjm = s.roty(f_1, -1), fjp = s.roty(f_1, 1);
        mx0_new = fmax(fmax(fmax(fmax(fmax(fmax(fm, fp), fjm), fjp), fkb), fkc), f_1);
        mn0_new = fmin(fmin(fmin(fmin(fmin(fmin(fm, fp), fjm), fjp), fkb), fkc), f_1);
      } else {
        f1_new = f_1 - (s.rot(uu1_1, 1) - uu1_1 + (ww1_0 - ww1_1) * iadz[L1]) * irho[L1];
        mx0_new = fmax(fmax(fmax(fmax(fm, fp), fkb), fkc), f_1);
        mn0_new = fmin(fmin(fmin(fmin(fm, fp), fkb), fkc), f_1);
      }
      if (flux) s1_new = s.sum(ww1_1);
    }
    // stage 2: antidiffusive velocities and second bounds at level K-2
    const int L2 = K - 2;
    R uu2_new = 0, ww2_new = 0, mx1_new = 0, mn1_new = 0, vv2_new = 0;
    if (ST || (L2 >= 0 && L2 < nzm)) {
      int kc = (ST || L2 + 1 < nzm) ? L2 + 1 : L2, kb = (ST || L2 >= 1) ? L2 - 1 : 0;
      R f1L = f1_a;
      R f1c = (kc == L2) ? f1_a : f1_new;
      R f1b = (kb == L2) ? f1_a : f1_b;
      R Uc = U_2, Ub = (kb == L2) ? U_2 : U_3;
      R Wc = W_2, Wkc = (kc == L2) ? W_2 : W_1;
      const R irho2 = irho[L2], irhow = irhowadz[L2];
      const R dd = ddadz[L2];
      R f1L_m = s.rot(f1L, -1), f1L_p = s.rot(f1L, 1);
      R f1c_m = s.rot(f1c, -1), f1b_m = s.rot(f1b, -1), f1b_p = s.rot(f1b, 1);
      R Wc_m = s.rot(Wc, -1), Wkc_m = s.rot(Wkc, -1);
      R Uc_p = s.rot(Uc, 1), Ub_p = s.rot(Ub, 1);
      if constexpr (Y3) {
        R Vc = V_2, Vb = (kb == L2) ? V_2 : V_3;
        R f1L_jm = s.roty(f1L, -1), f1L_jp = s.roty(f1L, 1);
        R f1L_m_jm = s.roty(f1L_m, -1), f1L_m_jp = s.roty(f1L_m, 1);
        R f1L_p_jm = s.roty(f1L_p, -1);
        R f1c_jm = s.roty(f1c, -1), f1b_jm = s.roty(f1b, -1), f1b_jp = s.roty(f1b, 1);
        R Vc_m = s.rot(Vc, -1), Vc_jp = s.roty(Vc, 1), Vc_m_jp = s.roty(Vc_m, 1), Vb_jp = s.roty(Vb, 1);
        R Uc_jm = s.roty(Uc, -1), Uc_p_jm = s.roty(Uc_p, -1);
        R Wc_jm = s.roty(Wc, -1), Wkc_jm = s.roty(Wkc, -1);
        uu2_new = andiff(f1L_m, f1L, Uc, irho2) -
                  (across(f1L_m_jp + f1L_jp - f1L_m_jm - f1L_jm, Uc, Vc_m + Vc_m_jp + Vc_jp + Vc) +
                   across(dd * (f1c_m + f1c - f1b_m - f1b), Uc, Wc_m + Wkc_m + Wc + Wkc)) * irho2;
        vv2_new = andiff(f1L_jm, f1L, Vc, irho2) -
                  (across(f1L_p_jm + f1L_p - f1L_m_jm - f1L_m, Vc, Uc_jm + Uc + Uc_p + Uc_p_jm) +
                   across(dd * (f1c_jm + f1c - f1b_jm - f1b), Vc, Wc_jm + Wc + Wkc + Wkc_jm)) * irho2;
        ww2_new = andiff(f1b, f1L, Wc, irhow) -
                  (across(f1b_p + f1L_p - f1b_m - f1L_m, Wc, Ub + Uc + Uc_p + Ub_p) +
                   across(f1L_jp + f1b_jp - f1L_jm - f1b_jm, Wc, Vb + Vb_jp + Vc_jp + Vc)) * irho2;
        mx1_new = fmax(fmax(fmax(fmax(fmax(fmax(fmax(f1L_m, f1L_p), f1L_jm), f1L_jp), f1b), f1c), f1L), mx0_a);
        mn1_new = fmin(fmin(fmin(fmin(fmin(fmin(fmin(f1L_m, f1L_p), f1L_jm), f1L_jp), f1b), f1c), f1L), mn0_a);
      } else {
        uu2_new = andiff(f1L_m, f1L, Uc, irho2) -
                  across(dd * (f1c_m + f1c - f1b_m - f1b), Uc, Wc_m + Wkc_m + Wc + Wkc) * irho2;
        ww2_new = andiff(f1b, f1L, Wc, irhow) -
                  across(f1b_p + f1L_p - f1b_m - f1L_m, Wc, Ub + Uc + Uc_p + Ub_p) * irho2;
        mx1_new = fmax(fmax(fmax(fmax(fmax(f1L_m, f1L_p), f1b), f1c), f1L), mx0_a);
        mn1_new = fmin(fmin(fmin(fmin(fmin(f1L_m, f1L_p), f1b), f1c), f1L), mn0_a);
      }
      if (!ST && L2 == 0) ww2_new = R(0.);
    }
    // stage 3: normalised limiters at level K-3
    const int L3 = K - 3;
    R mx_new = 0, mn_new = 0;
    if (ST || (L3 >= 0 && L3 < nzm)) {
      R uu2p = s.rot(uu2_a, 1);
      R ww2c = (ST || L3 + 1 < nzm) ? ww2_new : ww2_a;
      const R iadz3 = iadz[L3];
      if constexpr (Y3) {
        R vv2p = s.roty(vv2_a, 1);
        mx_new = rho[L3] * (mx1_a - f1_b) /
                 (pneg(uu2p) + ppos(uu2_a) + pneg(vv2p) + ppos(vv2_a) + iadz3 * (pneg(ww2c) + ppos(ww2_a)) + eps);
        mn_new = rho[L3] * (f1_b - mn1_a) /
                 (ppos(uu2p) + pneg(uu2_a) + ppos(vv2p) + pneg(vv2_a) + iadz3 * (ppos(ww2c) + pneg(ww2_a)) + eps);
      } else {
        mx_new = rho[L3] * (mx1_a - f1_b) / (pneg(uu2p) + ppos(uu2_a) + iadz3 * (pneg(ww2c) + ppos(ww2_a)) + eps);
        mn_new = rho[L3] * (f1_b - mn1_a) / (ppos(uu2p) + pneg(uu2_a) + iadz3 * (ppos(ww2c) + pneg(ww2_a)) + eps);
      }
    }
    // stage 4: limited antidiffusive fluxes at level K-4
    const int L4 = K - 4;
    R uu3_new = 0, ww3_new = 0, vv3_new = 0;
    if (ST || (L4 >= 0 && L4 < nzm)) {
      R mxm = s.rot(mx_a, -1), mnm = s.rot(mn_a, -1);
      R mxb = (ST || L4 >= 1) ? mx_b : mx_a, mnb = (ST || L4 >= 1) ? mn_b : mn_a;
      uu3_new = ppos(uu2_b) * fmin(fmin(R(1.0), mx_a), mnm) - pneg(uu2_b) * fmin(fmin(R(1.0), mxm), mn_a);
      ww3_new = ppos(ww2_b) * fmin(fmin(R(1.0), mx_a), mnb) - pneg(ww2_b) * fmin(fmin(R(1.0), mxb), mn_a);
      if constexpr (Y3) {
        R mxjm = s.roty(mx_a, -1), mnjm = s.roty(mn_a, -1);
        vv3_new = ppos(vv2_b) * fmin(fmin(R(1.0), mx_a), mnjm) - pneg(vv2_b) * fmin(fmin(R(1.0), mxjm), mn_a);
      }
      if (flux) {
        R s3 = s.sum(ww3_new);
        if (live && i == 0) flux[L4] = (double)s1_c + s3;
      }
    }
    // stage 5: corrected update at level K-5
    const int L5 = K - 5;
    if (ST || (L5 >= 0 && L5 < nzm)) {
      R ww3up = (ST || L5 + 1 < nzm) ? ww3_new : R(0.0);
      R fn;
      if constexpr (Y3)
        fn = fmax(R(0.0), f1_d - (s.rot(uu3_a, 1) - uu3_a + s.roty(vv3_a, 1) - vv3_a + (ww3up - ww3_a) * iadz[L5]) * irho[L5]);
      else
        fn = fmax(R(0.0), f1_d - (s.rot(uu3_a, 1) - uu3_a + (ww3up - ww3_a) * iadz[L5]) * irho[L5]);
      size_t o = ((size_t)c * nzm + L5) * NX + i;
      R adv = s.sum(fn - f_5);
      if (live) {
        f[o] = fn;
        if (fadv && i == 0) fadv[L5] = adv;
      }
      if (want_cfl) {
        R wmax = dz * adz[L5 >= 1 ? L5 - 1 : 0] / (R)d.dtn;
        if constexpr (std::is_same<R, double>::value) {
          // Only a Courant number above 0.9 changes nprec = max(1, ceiling(prec_cfl/0.9)), so
          // the exact FP64 value is needed only where the FP32 bound, with a 1e-3 margin, could
          // exceed 0.9 (or is not a number); elsewhere the bound, below 0.9, stands in for it
          // and nprec is the same
          R cfl = pratio[L5] * (R)term_vel_bound(d.tv, fn, rho[L5], tq_b) * (R(1.) / wmax);
          if (!(cfl * R(1.001) <= R4(0.9))) {
            R wp = pratio[L5] * term_vel(d.tv, fn, rho[L5], tq_b);
            cfl = wp * (R(1.) / wmax);
          }
          pcfl = fmax(pcfl, cfl);
        } else {
          R wp = pratio[L5] * term_vel(d.tv, fn, rho[L5], tq_b);
          pcfl = fmax(pcfl, wp * (R(1.) / wmax));
        }
      }
    }
    f_5 = f_4;
    f_4 = f_3;
    f_3 = f_2;
    f_2 = f_1;
    f_1 = f_0;
    tq_b = tq_a;
    tq_a = tq_new;
    U_3 = U_2;
    U_2 = U_1;
    U_1 = U_0;
    V_3 = V_2;
    V_2 = V_1;
    V_1 = V_0;
    vv1_1 = vv1_0;
    vv2_b = vv2_a;
    vv2_a = vv2_new;
    vv3_a = vv3_new;
    W_2 = W_1;
    W_1 = W_0;
    uu1_1 = uu1_0;
    ww1_1 = ww1_0;
    f1_d = f1_c;
    f1_c = f1_b;
    f1_b = f1_a;
    f1_a = f1_new;
    mx0_a = mx0_new;
    mn0_a = mn0_new;
    uu2_b = uu2_a;
    uu2_a = uu2_new;
    ww2_b = ww2_a;
    ww2_a = ww2_new;
    mx1_a = mx1_new;
    mn1_a = mn1_new;
    mx_b = mx_a;
    mx_a = mx_new;
    mn_b = mn_a;
    mn_a = mn_new;
    uu3_a = uu3_new;
    ww3_a = ww3_new;
    s1_c = s1_b;
    s1_b = s1_a;
    s1_a = s1_new;
  };
  int K = 0;
  // Peeled into fill, steady state and drain.  FP64 issues every instruction it keeps
  // (measured 27% faster peeled).  The FP32 sweep was measured faster as one loop until the
  // prefetch became unconditional; peeled since (2D 64-column 18% faster, the uneven classes
  // 4%).
  for (; K < 5 && K <= nzm + 4; K++) step(K, std::false_type{});
  // two levels per trip: fewer of the prefetched values are copied along the ring before
  // their use (a copy waits for its load); with the unconditional prefetch, measured 8%
  // faster at 64 x 58.  The other shapes hold more per lane and spilled more with it (the
  // uneven classes measured slower), so there it is the alternative variant
  if constexpr (PAIR) {
    for (; K + 1 < nzm; K += 2) {
      step(K, std::true_type{});
      step(K + 1, std::true_type{});
    }
  }
  for (; K < nzm; K++) step(K, std::true_type{});
  for (; K <= nzm + 4; K++) step(K, std::false_type{});
  if (want_cfl) {
    pcfl = s.max(pcfl);
    if (live && i == 0) atomic_max_nonneg(&d.gmax[GMAX_PCFL + (c % KURANT_SLOTS) * KURANT_STRIDE], pcfl);
  }
}

// advect_all_scalars.F90:7-94: t, then the two microphysics fields.
template <int NX, typename R, typename G, bool PAIR>
__device__ __forceinline__ void mpdata_all(const DevT<R>& d) {
  G s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  mpdata_one<NX, R, G, PAIR>(d, s, c, live, d.t, nullptr, nullptr, false);
  mpdata_one<NX, R, G, PAIR>(d, s, c, live, d.qt, prof(d, P_MKADV1, c), prof(d, P_MKWLE1, c), false);
  mpdata_one<NX, R, G, PAIR>(d, s, c, live, d.qp, prof(d, P_MKADV2, c), prof(d, P_MKWLE2, c), true);
}
template <int NX, typename R, typename G = SegFor<NX>, bool PAIR = (NX == 64 && !G::Y3)>
__global__ void CRM_LB k_mpdata(DevT<R> d) {
  CRM_STOP_CHECK(d);
  mpdata_all<NX, R, G, PAIR>(d);
}
// The same sweep with its registers bounded for three waves per SIMD (168 VGPRs; what does not
// fit is spilled): more waves to hide latency at the cost of the spills (launch variants)
template <int NX, typename R, typename G = SegFor<NX>, bool PAIR = (NX == 64 && !G::Y3)>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_mpdata_w3(DevT<R> d) {
  CRM_STOP_CHECK(d);
  mpdata_all<NX, R, G, PAIR>(d);
}
// The three scalars on separate waves (launch variants): the grid is three copies of the
// one-scalar grid of nb blocks, qp (the longest sweep, with its Courant number) first, then
// qt, then t.  The scalars are independent (each writes only its own field and statistics),
// so the results are those of the sequential sweep; the GPU holds three times as many sweeps,
// and a SIMD whose registers admit fewer waves than it has instances no longer ends with one
// wave running its three sweeps alone.
template <int NX, typename R, typename G = SegFor<NX>, bool PAIR = (NX == 64 && !G::Y3), bool W3 = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W3 ? 3 : CRM_WAVES_PER_EU)))
k_mpdata_split(DevT<R> d, int nb) {
  CRM_STOP_CHECK(d);
  G s;
  const int scalar = (int)blockIdx.x / nb, bl = (int)blockIdx.x % nb;
  int g = (bl * (int)blockDim.x + (int)threadIdx.x) / NX;
  if constexpr (NX >= 64) g = __builtin_amdgcn_readfirstlane(g);
  const int c = g < d.ncrms ? g : d.ncrms - 1;
  const bool live = g < d.ncrms;
  if (scalar == 0)
    mpdata_one<NX, R, G, PAIR>(d, s, c, live, d.qp, prof(d, P_MKADV2, c), prof(d, P_MKWLE2, c), true);
  else if (scalar == 1)
    mpdata_one<NX, R, G, PAIR>(d, s, c, live, d.qt, prof(d, P_MKADV1, c), prof(d, P_MKWLE1, c), false);
  else
    mpdata_one<NX, R, G, PAIR>(d, s, c, live, d.t, nullptr, nullptr, false);
}

// ------------------------------------------------------------------------------------------
// Precipitation sedimentation: precip_fall (microphysics.F90:529-833) with the batch-global
// sub-step count nprec read from the maximum Courant number gathered in k_mpdata.
// Per column, one pipelined sweep per sub-step.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_precip(DevT<R> d) {
  CRM_STOP_CHECK(d);
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, i = s.i;
  const R dz = inst(d, I_DZ, c), dtn = d.dtn;
  const cprof<R> adz = profc(d, P_ADZ, c);
  const cprof<R> rho = profc(d, P_RHO, c);
  const cprof<R> rhow = profc(d, P_RHOW, c);
  double* qpfall = prof(d, P_QPFALL, c);
  double* precflux = prof(d, P_PRECFLUX, c);
  const cprof<R> pratio = profc(d, P_PRATIO, c);
  const cprof<R> irhoadz_p = profc(d, P_IRHOADZ, c);
  const R eps = R4(1.e-10);
  const R a_pr = A_PR();
  // non-negative doubles order as their bit patterns
  unsigned long long pbits = 0;
  for (int j = 0; j < KURANT_SLOTS; j++) pbits = ::max(pbits, d.gmax[GMAX_PCFL + j * KURANT_STRIDE]);
  R prec_cfl = __longlong_as_double((long long)pbits);
  int nprec = 1;
  if (!(prec_cfl <= NPREC_MAX * R4(0.9))) {
    // non-finite or runaway precipitation Courant number: one sweep, and the host reports
    // the error after the call
    if (live && i == 0) atomicMax(&d.gmax[2], prec_cfl != prec_cfl ? ERR_FLAG_NAN : ERR_FLAG_PRECIP);
  } else if (prec_cfl > R4(0.9)) {
    nprec = (int)ceil(prec_cfl / R4(0.9));
  }
  // the call's largest count, for crm_hip_stats: nprec is batch-global, so one lane of one
  // instance keeps it (launches are stream-ordered)
  if (live && c == 0 && i == 0 && (unsigned long long)nprec > d.gmax[GMAX_NPREC])
    d.gmax[GMAX_NPREC] = (unsigned long long)nprec;
  const R rnprec = (R)nprec;
  auto Q = [&](R* p, int k) -> R& { return p[((size_t)c * nzm + k) * NX + i]; };
  auto lfac_of = [&](R tabs) {
    R om = fmax(R(0.0), fmin(R(1.0), (tabs - TPRMIN) * a_pr));
    return FAC_COND + (1 - om) * FAC_FUS;
  };
  const R omega0 = fmax(R(0.0), fmin(R(1.0), (Q(d.tabs, 0) - TPRMIN) * a_pr));
  for (int iprec = 1; iprec <= nprec; iprec++) {
    R tq_0 = 0, tq_1 = 0, tq_2 = 0, tq_3 = 0;
    R ta_0 = 0, ta_1 = 0, ta_2 = 0, ta_3 = 0;  // tabs ring
    R tt_a = 0, tt_b = 0;                       // t of level K-2, K-3
    // qp and tabs of the next two levels, loaded two iterations before use
    R nq1 = Q(d.qp, 0), na1 = Q(d.tabs, 0), nq2 = Q(d.qp, 1), na2 = Q(d.tabs, 1);
    R wp_0 = 0, wp_1 = 0, wp_2 = 0;
    R fz0_0 = 0, fz0_1 = 0, fz0_2 = 0;
    R tq1_a = 0, tq1_b = 0;
    R mx0_a = 0, mn0_a = 0;
    R www_a = 0;
    R mx_a = 0, mn_a = 0;
    R fz_a = 0;
    // one level step; ST: steady state (4 <= K < nzm), where every stage is active and every
    // boundary select is the interior one, so the guards fold away (as in MPDATA)
    auto step = [&](const int K, auto steady) {
      constexpr bool ST = decltype(steady)::value;
      const R cq = nq1, ca = na1;
      nq1 = nq2;
      na1 = na2;
      {  // unconditional, clamped at the top level: past it the loaded values are not used
         // (stage a is idle there), and a load that always issues writes the ring register
         // directly (a conditional one is preceded by a copy that waits for earlier loads).
         // Level nzm-1 is written only at step nzm+2, after the last of these reads.
        const int kn = ::min(K + 2, nzm - 1);
        nq2 = Q(d.qp, kn);
        na2 = Q(d.tabs, kn);
      }
      const R tt_new = (ST || (K >= 1 && K - 1 < nzm)) ? Q(d.t, K - 1) : R(0.0);
      // a: level K, upwind flux
      if (ST || K < nzm) {
        tq_0 = cq;
        ta_0 = ca;
        R v = pratio[K] * term_vel(d.tv, tq_0, rho[K], ta_0);
        wp_0 = -v * rhow[K] * dtn / dz / rnprec;
        fz0_0 = tq_0 * wp_0;
      } else {
        tq_0 = wp_0 = fz0_0 = R(0.0);
      }
      // b: level K-1, upwind update, first bounds, antidiffusive flux
      const int B = K - 1;
      R tq1_new = 0, mx0_new = 0, mn0_new = 0, www_new = 0;
      if (ST || (B >= 0 && B < nzm)) {
        R irhoadz = irhoadz_p[B];
        tq1_new = tq_1 - (fz0_0 - fz0_1) * irhoadz;
        R tkb = (ST || B >= 1) ? tq_2 : tq_1, tkc = (ST || B + 1 < nzm) ? tq_0 : tq_1;
        mx0_new = fmax(fmax(tkb, tkc), tq_1);
        mn0_new = fmin(fmin(tkb, tkc), tq_1);
        R t1b = (ST || B >= 1) ? tq1_a : tq1_new, wpb = (ST || B >= 1) ? wp_2 : wp_1;
        www_new = R(0.5) * (R(1.) + wp_1 * irhoadz) * (t1b * wpb - tq1_new * wp_1);
      }
      // d: level K-2, limiter and corrected flux
      const int D = K - 2;
      R mx_new = 0, mn_new = 0, fz_new = 0;
      if (ST || (D >= 0 && D < nzm)) {
        R t1b = (ST || D >= 1) ? tq1_b : tq1_a, t1c = (ST || D + 1 < nzm) ? tq1_new : tq1_a;
        R mx1 = fmax(fmax(fmax(t1b, t1c), tq1_a), mx0_a);
        R mn1 = fmin(fmin(fmin(t1b, t1c), tq1_a), mn0_a);
        R wwc = (ST || D + 1 < nzm) ? www_new : www_a;
        R ra = rho[D] * adz[D];
        mx_new = ra * (mx1 - tq1_a) / (pneg(wwc) + ppos(www_a) + eps);
        mn_new = ra * (tq1_a - mn1) / (ppos(wwc) + pneg(www_a) + eps);
        R mxb = (ST || D >= 1) ? mx_a : mx_new, mnb = (ST || D >= 1) ? mn_a : mn_new;
        fz_new = fz0_2 + ppos(www_a) * fmin(fmin(R(1.0), mx_new), mnb) - pneg(www_a) * fmin(fmin(R(1.0), mxb), mn_new);
      }
      // f: level K-3, update of qp and t, statistics
      const int Fl = K - 3;
      if (ST || (Fl >= 0 && Fl < nzm)) {
        R irhoadz = irhoadz_p[Fl];
        R fzF = fz_a, fzF1 = (ST || Fl + 1 < nzm) ? fz_new : R(0.0);
        R lf = lfac_of(ta_3), lf1 = (ST || Fl + 1 < nzm) ? lfac_of(ta_2) : R(0.0);
        R dq = -(fzF1 - fzF) * irhoadz;
        R lat_heat = -(lf1 * fzF1 - lf * fzF) * irhoadz;
        R sfall = s.sum(dq * R(1.));
        R sflux = s.sum(fzF * R(1.));
        if (live) {
          Q(d.qp, Fl) = tq_3 - (fzF1 - fzF) * irhoadz;
          Q(d.t, Fl) = tt_b - lat_heat;
          if (i == 0) {
            accum(&qpfall[Fl], sfall);
            accum(&precflux[Fl], -sflux);
          }
          if (!ST && Fl == 0) {
            d.precsfc[(size_t)c * NX + i] -= fzF * R(1.);
            d.precssfc[(size_t)c * NX + i] -= fzF * (R(1.) - omega0) * R(1.);
          }
        }
      }
      tq_3 = tq_2;
      tq_2 = tq_1;
      tq_1 = tq_0;
      ta_3 = ta_2;
      ta_2 = ta_1;
      ta_1 = ta_0;
      tt_b = tt_a;
      tt_a = tt_new;
      wp_2 = wp_1;
      wp_1 = wp_0;
      fz0_2 = fz0_1;
      fz0_1 = fz0_0;
      tq1_b = tq1_a;
      tq1_a = tq1_new;
      mx0_a = mx0_new;
      mn0_a = mn0_new;
      www_a = www_new;
      mx_a = mx_new;
      mn_a = mn_new;
      fz_a = fz_new;
    };
    int K = 0;
    for (; K < 4 && K <= nzm + 2; K++) step(K, std::false_type{});
    for (; K < nzm; K++) step(K, std::true_type{});
    for (; K <= nzm + 2; K++) step(K, std::false_type{});
  }
}

// ------------------------------------------------------------------------------------------
// SGS diffusion of t, qt, qp: diffuse_scalar.F90:8-72 and diffuse_scalar2D.F90:5-141 (3D:
// diffuse_scalar3D.F90:6-218, dy == dx; zero surface and top fluxes in the CRM).
// ------------------------------------------------------------------------------------------
// The three fields diffuse independently and share the eddy diffusivity, so one sweep carries
// all three: tkh and the level profiles are read once per level instead of once per field, and
// the three fields' chains interleave.  Each field's arithmetic is the one-field sweep's, so
// the results are the same bits.
template <int NX, typename R, typename G = SegFor<NX>>
__global__ void CRM_LB k_diffuse(DevT<R> d) {
  CRM_STOP_CHECK(d);
  G s;
  int c;
  const bool live = seg_instance<NX>(d.ncrms, c);
  constexpr int NF = 3;
  R* const fs[NF] = {d.t, d.qt, d.qp};
  double* const fdiffs[NF] = {nullptr, prof(d, P_MKDIFF1, c), prof(d, P_MKDIFF2, c)};
  double* const fluxes[NF] = {nullptr, prof(d, P_MKWSB1, c), prof(d, P_MKWSB2, c)};
  const int nzm = d.nzm, i = s.i;
  const R dz = inst(d, I_DZ, c), dtn = d.dtn;
  const cprof<R> adzw = profc(d, P_ADZW, c);
  const cprof<R> rhow = profc(d, P_RHOW, c);
  const cprof<R> grdf = profc(d, P_GRDFX, c);
  const cprof<R> irhoadz = profc(d, P_IRHOADZ, c);
  const cprof<R> drhoi = profc(d, P_DRHOI, c);
  const R rdx2 = (R)(R(1.) / (d.dx * d.dx)), rdz2 = (R)1. / (dz * dz);
  auto Q = [&](const R* p, int k) { return p[((size_t)c * nzm + k) * NX + i]; };
  R f_k[NF], f_n[NF], flx_prev[NF];
#pragma unroll
  for (int q = 0; q < NF; q++) {
    f_k[q] = Q(fs[q], 0);
    f_n[q] = Q(fs[q], 1);
    flx_prev[q] = R(0.0) * (R(1.) / dz) * rhow[0];
  }
  R tkh_k = Q(d.tkh, 0), tkh_n = Q(d.tkh, 1);
  for (int k = 0; k < nzm; k++) {
    // level k+2, clamped at the top: a value read past it is never used (level k+1's values
    // enter only inner levels, k+1 <= nzm-2), and a load that always issues needs no copy
    // that would wait for earlier loads
    const int k2 = ::min(k + 2, nzm - 1);
    R f_nn[NF];
#pragma unroll
    for (int q = 0; q < NF; q++) f_nn[q] = Q(fs[q], k2);
    const R tkh_nn = Q(d.tkh, k2);
    const R rdx5 = R(0.5) * rdx2 * grdf[k];
    const R tkx = rdx5 * (tkh_k + s.rot(tkh_k, 1));
    R tky = R(0.0);
    if constexpr (G::Y3) tky = rdx5 * (tkh_k + s.roty(tkh_k, 1));
    const bool inner = k <= nzm - 2;
    R tkz = R(0.0), rhoi_z = R(0.0);
    if (inner) {
      rhoi_z = drhoi[k];
      const R rdz5 = R(0.5) * rdz2 * R(1.0);
      tkz = rdz5 * (tkh_k + tkh_n);
    }
    const R rhoi = irhoadz[k];
    // FP64: the level sums of qt and qp (vertical SGS flux, change), [2 (q - 1)] and
    // [2 (q - 1) + 1], in one batched reduction
    R red[2 * (NF - 1)];
#pragma unroll
    for (int q = 0; q < NF; q++) {
      const R fk = f_k[q];
      const R flxh = -tkx * (s.rot(fk, 1) - fk);
      R dfh = R(0.) - (flxh - s.rot(flxh, -1));
      if constexpr (G::Y3) {
        const R flxy = -tky * (s.roty(fk, 1) - fk);
        dfh = dfh - (flxy - s.roty(flxy, -1));
      }
      R flxv;
      if (inner)
        flxv = -tkz * (f_n[q] - fk) * rhoi_z;
      else
        flxv = R(0.0) * (R(1.) / dz) * (R(1.) / adzw[nzm]) * rhow[nzm];
      const R dfdt = dtn * (dfh - (flxv - flx_prev[q]) * rhoi);
      const R fnew = fk + dfdt;
      if (live) fs[q][((size_t)c * nzm + k) * NX + i] = fnew;
      if (q > 0) {
        if constexpr (std::is_same<R, double>::value) {  // reduced together below
          red[2 * (q - 1)] = flxv;
          red[2 * (q - 1) + 1] = fnew - fk;
        } else {  // FP32: one DPP reduction each, in place (measured faster than batched)
          if (inner) {
            const R sv = s.sum(flxv);
            if (live && i == 0) fluxes[q][k + 1] = sv;
          }
          const R sd = s.sum(fnew - fk);
          if (live && i == 0) fdiffs[q][k] = sd;
        }
      }
      flx_prev[q] = flxv;
      f_k[q] = f_n[q];
      f_n[q] = f_nn[q];
    }
    if constexpr (std::is_same<R, double>::value) {
      s.sums(red);
      if (live) {
#pragma unroll
        for (int q = 1; q < NF; q++) {
          if (inner && s.template sum_owner<R>(2 * (q - 1))) fluxes[q][k + 1] = red[2 * (q - 1)];
          if (s.template sum_owner<R>(2 * (q - 1) + 1)) fdiffs[q][k] = red[2 * (q - 1) + 1];
        }
      }
    }
    tkh_k = tkh_n;
    tkh_n = tkh_nn;
  }
#pragma unroll
  for (int q = 0; q < NF; q++)
    if (fluxes[q] && live && i == 0) fluxes[q][0] = R(0.0);
}

// ------------------------------------------------------------------------------------------
// Bulk microphysics coefficients per level: precip_init.F90:47-91 (one thread per level).
// ------------------------------------------------------------------------------------------
template <typename R>
__global__ void k_precip_init(DevT<R> d) {
  CRM_STOP_CHECK(d);
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= d.ncrms * d.nzm) return;
  int c = idx / d.nzm, k = idx - c * d.nzm;
  double rho = prof(d, P_RHO, c)[k], t0 = prof(d, P_TABS0, c)[k], pres = prof(d, P_PRES, c)[k];
  auto C = [&](int q) -> R& { return d.coef[((size_t)q * d.ncrms + c) * d.nz + k]; };
  const Gam& g = d.gam;
  double pratio = sqrt(R4(1.29) / rho);
  double rrr1 = 393. / (t0 + 120.) * pow(t0 / 273., 1.5);
  double rrr2 = pow(t0 / 273., R4(1.94)) * (1000. / pres);
  double estw = 100. * esatw(t0);
  double esti = 100. * esati(t0);
  double coef1 = 0.25 * PI_R4 * NZEROS * A_SNOW * g.gams1 * pratio / pow(PI_R4 * RHOS * NZEROS / rho, (3 + B_SNOW) / 4.);
  double coef2 = exp(R4(0.025) * (t0 - R4(273.15)));
  C(C_ACCRSI) = coef1 * coef2 * ESICOEF;
  C(C_ACCRSC) = coef1 * 1.0;
  C(C_COEFICE) = coef2;
  coef1 = (LSUB / (t0 * RV) - 1.) * LSUB / (THERCO * rrr1 * t0);
  coef2 = RV * t0 / (DIFFELQ * rrr2 * esti);
  C(C_EVAPS1) = R4(0.65) * 4. * NZEROS / sqrt(PI_R4 * RHOS * NZEROS) / (coef1 + coef2) / sqrt(rho);
  C(C_EVAPS2) = R4(0.49) * 4. * NZEROS * g.gams2 * sqrt(A_SNOW / (MUELQ * rrr1)) / pow(PI_R4 * RHOS * NZEROS, (5 + B_SNOW) / 8.) /
                (coef1 + coef2) * pow(rho, (1 + B_SNOW) / 8.) * sqrt(pratio);
  coef1 = 0.25 * PI_R4 * NZEROG * A_GRAU * g.gamg1 * pratio / pow(PI_R4 * RHOG * NZEROG / rho, (3 + B_GRAU) / 4.);
  coef2 = exp(R4(0.025) * (t0 - R4(273.15)));
  C(C_ACCRGI) = coef1 * coef2 * EGICOEF;
  C(C_ACCRGC) = coef1 * 1.0;
  coef1 = (LSUB / (t0 * RV) - 1.) * LSUB / (THERCO * rrr1 * t0);
  coef2 = RV * t0 / (DIFFELQ * rrr2 * esti);
  C(C_EVAPG1) = R4(0.65) * 4. * NZEROG / sqrt(PI_R4 * RHOG * NZEROG) / (coef1 + coef2) / sqrt(rho);
  C(C_EVAPG2) = R4(0.49) * 4. * NZEROG * g.gamg2 * sqrt(A_GRAU / (MUELQ * rrr1)) / pow(PI_R4 * RHOG * NZEROG, (5 + B_GRAU) / 8.) /
                (coef1 + coef2) * pow(rho, (1 + B_GRAU) / 8.) * sqrt(pratio);
  C(C_ACCRRC) = 0.25 * PI_R4 * NZEROR * A_RAIN * g.gamr1 * pratio / pow(PI_R4 * RHOR * NZEROR / rho, (3 + B_RAIN) / 4.) * 1.0;
  coef1 = (LCOND / (t0 * RV) - 1.) * LCOND / (THERCO * rrr1 * t0);
  coef2 = RV * t0 / (DIFFELQ * rrr2 * estw);
  C(C_EVAPR1) = R4(0.78) * 2. * PI_R4 * NZEROR / sqrt(PI_R4 * RHOR * NZEROR) / (coef1 + coef2) / sqrt(rho);
  C(C_EVAPR2) = R4(0.31) * 2. * PI_R4 * NZEROR * g.gamr2 * R4(0.89) * sqrt(A_RAIN / (MUELQ * rrr1)) /
                pow(PI_R4 * RHOR * NZEROR, (5 + B_RAIN) / 8.) / (coef1 + coef2) * pow(rho, (1 + B_RAIN) / 8.) * sqrt(pratio);
}

// ------------------------------------------------------------------------------------------
// Saturation adjustment (cloud.F90:7-114), precipitation processes (precip_proc.F90:7-151),
// micro_diagnose (microphysics.F90:331-357) and diagnose (diagnose.F90:8-101), pointwise
// per level followed by the level means.  Also finds the cloud-ice level range for the next
// ice fall (ice_fall.F90:34-46).
// ------------------------------------------------------------------------------------------
// Inputs of the microphysics network at the entry of micro_proc, per grid point: T, p, rho,
// qv, qcl, qci, qpl, qpi, w, with the current total water and precipitation split by the
// temperature ramps of the last diagnose (microphysics.F90:331-357) and T from the static
// energy with that split (diagnose.F90).
template <int NX, typename R>
__global__ void k_mlp_inputs(DevT<R> d, float* X) {
  const size_t np = (size_t)d.ncrms * d.nzm * NX;
  const size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= np) return;
  const int i = (int)(o % NX);
  const size_t ck = o / NX;
  const int k = (int)(ck % d.nzm), c = (int)(ck / d.nzm);
  const double a_bg = A_BG(), a_pr = A_PR();
  const double q = fmax(0.0, (double)d.qt[o]), qp = d.qp[o], qn = d.qn[o], tabs = d.tabs[o];
  const double omn = fmax(0.0, fmin(1.0, (tabs - TBGMIN) * a_bg));
  const double omp = fmax(0.0, fmin(1.0, (tabs - TPRMIN) * a_pr));
  const double qcl = qn * omn, qci = qn * (1. - omn), qpl = qp * omp, qpi = qp * (1. - omp);
  const double T = (double)d.t[o] - prof(d, P_GAMAZ, c)[k] + FAC_COND_D * (qcl + qpl) + FAC_SUB_D * (qci + qpi);
  const size_t ow = ((size_t)c * d.nz + k) * NX + i;
  float* x = X + o * CRM_MLP_NIN;
  x[0] = (float)T;
  x[1] = (float)prof(d, P_PRES, c)[k];
  x[2] = (float)prof(d, P_RHO, c)[k];
  x[3] = (float)(q - qn);
  x[4] = (float)qcl;
  x[5] = (float)qci;
  x[6] = (float)qpl;
  x[7] = (float)qpi;
  x[8] = (float)(0.5 * ((double)d.w[ow] + (double)d.w[ow + NX]));
}

// diagnose (diagnose.F90:8-177) at level k: tabs from t and the condensate split, the level
// means, the cloud-ice level range of ice_fall; writes the microphysics fields and tabs.
// With src_evp (from micro_proc), the level sums of the precipitation source and evaporation
// join the same batched reduction and are written to P_QPSRC and P_QPEVP.
template <int NX, typename R, typename S_>
__device__ __forceinline__ void diagnose_level(const DevT<R>& d, const S_& s, int c, bool live, int k, size_t o,
                                               bool y3, R t, R q, R qp, R qn, R qv, R qcl, R qci, R qpl, R qpi,
                                               R u, R v, int& kmin, int& kmax, const R* src_evp = nullptr) {
  const R factor = R(1.) / (R)NX;
  const cprof<R> gamaz = profc(d, P_GAMAZ, c);
  R tabs = t - gamaz[k] + FAC_COND * (qcl + qpl) + FAC_SUB * (qci + qpi);
  // FP64: the level sums in one batched reduction, q0 and qn0 at 0 and 4, which one lane holds
  // together (P_QV0 is their difference); FP32: one DPP reduction each (measured faster)
  R sv[9] = {qv + qcl + qci, u, t, tabs, qcl + qci, qpl + qpi, y3 ? v : R(0.0),
             src_evp ? src_evp[0] : R(0.0), src_evp ? src_evp[1] : R(0.0)};
  if constexpr (std::is_same<R, double>::value) {
    s.sums(sv);
  } else {
    sv[1] = s.sum(u);
    if (y3) sv[6] = s.sum(v);
    sv[2] = s.sum(t);
    sv[3] = s.sum(tabs);
    sv[0] = s.sum(sv[0]);
    sv[4] = s.sum(sv[4]);
    sv[5] = s.sum(sv[5]);
    if (src_evp) {
      sv[7] = s.sum(sv[7]);
      sv[8] = s.sum(sv[8]);
    }
  }
  const R q0 = sv[0] * factor, u0 = sv[1] * factor, t0 = sv[2] * factor, tabs0 = sv[3] * factor;
  const R qn0 = sv[4] * factor, qp0 = sv[5] * factor, v0 = y3 ? sv[6] * factor : R(0.0);
  if (qcl + qci > R(0.) && tabs < R4(273.15)) {
    kmin = ::min(kmin, k);
    kmax = ::max(kmax, k);
  }
  auto own = [&](int j) { return s.template sum_owner<R>(j); };
  if (live && src_evp) {
    if (own(7)) prof(d, P_QPSRC, c)[k] = sv[7];
    if (own(8)) prof(d, P_QPEVP, c)[k] = sv[8];
  }
  if (live) {
    d.qt[o] = q;
    d.qp[o] = qp;
    d.qn[o] = qn;
    d.tabs[o] = tabs;
    d.qv[o] = qv;
    d.qcl[o] = qcl;
    d.qci[o] = qci;
    d.qpl[o] = qpl;
    d.qpi[o] = qpi;
    if (own(0)) {
      prof(d, P_Q0, c)[k] = q0;
      prof(d, P_QV0, c)[k] = q0 - qn0;
    }
    if (own(1)) prof(d, P_U0, c)[k] = u0;
    if (own(2)) prof(d, P_T0, c)[k] = t0;
    if (own(3)) prof(d, P_TABS0, c)[k] = tabs0;
    if (own(4)) prof(d, P_QN0, c)[k] = qn0;
    if (own(5)) prof(d, P_QP0, c)[k] = qp0;
    if (y3 && own(6)) prof(d, P_V0, c)[k] = v0;
    if constexpr (!std::is_same<R, double>::value) {  // FP32 copies read by the bulk kernels
      if (s.i == 0) {
        if (y3) profr(d, P_V0, c)[k] = v0;
        profr(d, P_U0, c)[k] = u0;
        profr(d, P_T0, c)[k] = t0;
        profr(d, P_TABS0, c)[k] = tabs0;
        profr(d, P_Q0, c)[k] = q0;
        profr(d, P_QN0, c)[k] = qn0;
        profr(d, P_QP0, c)[k] = qp0;
        profr(d, P_QV0, c)[k] = q0 - qn0;
      }
    }
  }
}

// The level means of accelerate_crm.F90:196-250 after micro_proc, minus those of the last
// diagnose; flags a mean t change above 5 K (ceaseflag) in gmax[3].
template <int NX, typename R, typename S_>
__device__ __forceinline__ void accel_means(const DevT<R>& d, const S_& s, int c, bool live, int k, size_t o,
                                            bool y3, R t, R qv, R qcl, R qci) {
  const double coef = 1. / (double)NX;
  double tb = s.sum((double)t * coef);
  double qtb = s.sum(((double)qcl + (double)qci + (double)qv) * coef);
  double ub = s.sum((double)d.u[o] * coef);
  double vb = y3 ? s.sum((double)d.v[o] * coef) : 0.0;
  if (live && s.i == 0) {
    double tt = tb - prof(d, P_T0, c)[k];
    prof(d, P_ACCT, c)[k] = tt;
    prof(d, P_ACCQ, c)[k] = qtb - prof(d, P_Q0, c)[k];
    prof(d, P_ACCU, c)[k] = ub - prof(d, P_U0, c)[k];
    if (y3) prof(d, P_ACCV, c)[k] = vb - prof(d, P_V0, c)[k];
    if (fabs(tt) > 5.) atomicMax(&d.gmax[3], 1ull);
  }
}

// KUR: also gather the kurant maximum (k_kurant's, kurant.F90:9-83 and sgs.F90:329-372) of
// the state the kernel leaves, for the next substep of a captured call: u, v, w and tkh are
// final by then (nothing after it in the substep changes them), and k_micro already reads u
// and v every level, so only w and tkh are added to its prefetch.
template <int NX, typename R, bool KUR = false>
__global__ void CRM_LB k_micro(DevT<R> d, int accel) {
  CRM_STOP_CHECK(d);
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const bool y3 = d.ny > 1;
  const R dtn = d.dtn, factor = R(1.) / (R)NX;
  const cprof<R> gamaz = profc(d, P_GAMAZ, c);
  const cprof<R> presp = profc(d, P_PRES, c);
  const R an = R(1.) / (TBGMAX - TBGMIN), bn = TBGMIN * an;
  const R ap = R(1.) / (TPRMAX - TPRMIN), bp = TPRMIN * ap;
  const R fac1 = FAC_COND + (1 + bp) * FAC_FUS, fac2 = FAC_FUS * ap;
  const R a_bg = A_BG(), a_pr = A_PR(), a_gr = A_GR();
  const R powr1 = (3 + B_RAIN) / R(4.), powr2 = (5 + B_RAIN) / R(8.);
  const R pows1 = (3 + B_SNOW) / R(4.), pows2 = (5 + B_SNOW) / R(8.);
  const R powg1 = (3 + B_GRAU) / R(4.), powg2 = (5 + B_GRAU) / R(8.);
  int kmin = nzm, kmax = -1;
  // t, qt, qp (and the winds for the level means) of the next level, read one level ahead
  const size_t o0 = ((size_t)c * nzm) * NX + i;
  R t_pre = d.t[o0], qt_pre = d.qt[o0], qp_pre = d.qp[o0], u_pre = d.u[o0];
  R v_pre = y3 ? d.v[o0] : R(0.0);
  // KUR: w of levels k and k+1 (the second read one level ahead) and tkh of level k
  auto Wk = [&](int k) { return d.w[((size_t)c * nz + k) * NX + i]; };
  R w_a = 0, w_b = 0, tkh_pre = 0;
  double cfl = 0.0;
  const double d_dz = inst(d, I_DZ, c);
  const cprof<double> kur_adzw = profcd(d, P_ADZW, c);
  const cprof<double> kur_grdf = profcd(d, P_GRDFX, c);
  const double kur_dxf = sqrt((1. / d.dx) * (1. / d.dx) + (y3 ? (1. / d.dx) * (1. / d.dx) : 0.0));
  if constexpr (KUR) {
    w_a = Wk(0);
    w_b = Wk(1);
    tkh_pre = d.tkh[o0];
  }
  for (int k = 0; k < nzm; k++) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    // the precip_init coefficients are constant during this kernel: scalar loads, which do
    // not wait for the vector prefetch of the next level (a vector load here would)
    auto CO = [&](int q) { return ((cprof<R>)d.coef)[((size_t)q * d.ncrms + c) * d.nz + k]; };
    const R pres = presp[k];
    R t = t_pre;
    R q = fmax(R(0.0), qt_pre);
    R qp = qp_pre;
    const R q_in = q, qp_in = qp;  // the entry state, for recorded training pairs
    const R u_k = u_pre, v_k = v_pre;
    const R w_k = w_a, w_k1 = w_b, tkh_k = tkh_pre;
    {  // clamped at the top, where the values are not used again (an unconditional load needs
       // no copy that waits for earlier loads)
      const size_t on = ((size_t)c * nzm + ::min(k + 1, nzm - 1)) * NX + i;
      t_pre = d.t[on];
      qt_pre = d.qt[on];
      qp_pre = d.qp[on];
      u_pre = d.u[on];
      if (y3) v_pre = d.v[on];
      if constexpr (KUR) {
        w_a = w_b;
        w_b = Wk(::min(k + 2, nzm));
        tkh_pre = d.tkh[on];
      }
    }
    // the network's rates where it applies (inside its training envelope): condensate and
    // precipitation stepped by the rates, clipped at zero, total water q + qp kept, the
    // vapour taking the remainder (condensate scaled down should the vapour go negative); t
    // (the static energy) is untouched, so the later diagnose conserves energy as the scheme
    // does; T by its rate sets the temperature of the condensate split
    const bool net = d.mlp_y != nullptr && d.mlp_y[o * 8 + 6] > 0.5f;
    R tabs;
    R src = R(0.0), evp = R(0.0);
    R qn;
    if (net) {
      const float* y = d.mlp_y + o * 8;
      const float* x = d.mlp_x + o * CRM_MLP_NIN;
      const R qn_in = (R)x[4] + (R)x[5];
      R qn_n = fmax(R(0.0), qn_in + ((R)y[1] + (R)y[2]) * dtn);
      R qp_n = fmax(R(0.0), qp_in + ((R)y[3] + (R)y[4]) * dtn);
      const R wtot = q + qp_in;
      R qv_n = wtot - qn_n - qp_n;
      if (qv_n < R(0.)) {
        const R f = (qn_n + qp_n) > R(0.) ? fmax(R(0.0), wtot) / (qn_n + qp_n) : R(0.);
        qn_n = qn_n * f;
        qp_n = qp_n * f;
        qv_n = R(0.);
      }
      const R dqp = qp_n - qp_in;
      if (dqp > R(0.))
        src = dqp;
      else
        evp = dqp;
      tabs = (R)x[0] + (R)y[5] * dtn;
      q = qv_n + qn_n;
      qn = qn_n;
      qp = qp_n;
    } else {
    // cloud(): saturation adjustment with Newton iterations
    tabs = t - gamaz[k];
    R tabs1 = (tabs + fac1 * qp) / (R(1.) + fac2 * qp);
    R qsatt, om;
    if (tabs1 >= TBGMAX) {
      tabs1 = tabs + FAC_COND * qp;
      qsatt = qsatw(tabs1, pres);
    } else if (tabs1 <= TBGMIN) {
      tabs1 = tabs + FAC_SUB * qp;
      qsatt = qsati(tabs1, pres);
    } else {
      om = an * tabs1 - bn;
      qsatt = om * qsatw(tabs1, pres) + (R(1.) - om) * qsati(tabs1, pres);
    }
    if (q > qsatt) {
      int niter = 0;
      R dtabs = R(100.), dqsat = R(0.);
      while (fabs(dtabs) > R4(0.01) && niter < 10) {
        R lstarn, dlstarn, lstarp, dlstarp;
        if (tabs1 >= TBGMAX) {
          lstarn = FAC_COND;
          dlstarn = R(0.);
          qsatt = qsatw(tabs1, pres);
          dqsat = dtqsatw(tabs1, pres);
        } else if (tabs1 <= TBGMIN) {
          lstarn = FAC_SUB;
          dlstarn = R(0.);
          qsatt = qsati(tabs1, pres);
          dqsat = dtqsati(tabs1, pres);
        } else {
          om = an * tabs1 - bn;
          lstarn = FAC_COND + (R(1.) - om) * FAC_FUS;
          dlstarn = an * FAC_FUS;
          qsatt = om * qsatw(tabs1, pres) + (R(1.) - om) * qsati(tabs1, pres);
          dqsat = om * dtqsatw(tabs1, pres) + (R(1.) - om) * dtqsati(tabs1, pres);
        }
        if (tabs1 >= TPRMAX) {
          lstarp = FAC_COND;
          dlstarp = R(0.);
        } else if (tabs1 <= TPRMIN) {
          lstarp = FAC_SUB;
          dlstarp = R(0.);
        } else {
          R omp = ap * tabs1 - bp;
          lstarp = FAC_COND + (R(1.) - omp) * FAC_FUS;
          dlstarp = ap * FAC_FUS;
        }
        R fff = tabs - tabs1 + lstarn * (q - qsatt) + lstarp * qp;
        R dfff = dlstarn * (q - qsatt) + dlstarp * qp - lstarn * dqsat - R(1.);
        dtabs = -fff / dfff;
        niter++;
        tabs1 = tabs1 + dtabs;
      }
      qsatt = qsatt + dqsat * dtabs;
      qn = fmax(R(0.0), q - qsatt);
    } else {
      qn = R(0.);
    }
    tabs = tabs1;
    qp = fmax(R(0.0), qp);
    // precip_proc()
    if (qn + qp > R(0.)) {
      R omn = fmax(R(0.0), fmin(R(1.0), (tabs - TBGMIN) * a_bg));
      R omp = fmax(R(0.0), fmin(R(1.0), (tabs - TPRMIN) * a_pr));
      R omg = fmax(R(0.0), fmin(R(1.0), (tabs - TGRMIN) * a_gr));
      if (qn > R(0.)) {
        R qcc = qn * omn, qii = qn * (R(1.) - omn);
        R autor = (qcc > QCW0) ? ALPHAELQ : R(0.);
        R autos = (qii > QCI0) ? BETAELQ * CO(C_COEFICE) : R(0.);
        R accrr = R(0.);
        if (omp > R4(0.001)) accrr = CO(C_ACCRRC) * fpow<R>(qp * omp, powr1);
        R accrcs = R(0.), accris = R(0.);
        if (omp < R4(0.999) && omg < R4(0.999)) {
          R tmp = fpow<R>(qp * (R(1.) - omp) * (R(1.) - omg), pows1);
          accrcs = CO(C_ACCRSC) * tmp;
          accris = CO(C_ACCRSI) * tmp;
        }
        R accrcg = R(0.), accrig = R(0.);
        if (omp < R4(0.999) && omg > R4(0.001)) {
          R tmp = fpow<R>(qp * (R(1.) - omp) * omg, powg1);
          accrcg = CO(C_ACCRGC) * tmp;
          accrig = CO(C_ACCRGI) * tmp;
        }
        qcc = (qcc + dtn * autor * QCW0) / (R(1.) + dtn * (accrr + accrcs + accrcg + autor));
        qii = (qii + dtn * autos * QCI0) / (R(1.) + dtn * (accris + accrig + autos));
        R dq = dtn * (accrr * qcc + autor * (qcc - QCW0) + (accris + accrig) * qii + (accrcs + accrcg) * qcc + autos * (qii - QCI0));
        dq = fmin(dq, qn);
        qp = qp + dq;
        q = q - dq;
        qn = qn - dq;
        src = dq;
      } else if (qp > QP_THRESHOLD && qn == R(0.)) {
        R qs = R(0.);
        if (omn > R4(0.001)) qs = qs + omn * qsatw(tabs, pres);
        if (omn < R4(0.999)) qs = qs + (R(1.) - omn) * qsati(tabs, pres);
        R dq = R(0.);
        if (omp > R4(0.001)) {
          R qrr = qp * omp;
          dq = dq + CO(C_EVAPR1) * sqrt(qrr) + CO(C_EVAPR2) * fpow<R>(qrr, powr2);
        }
        if (omp < R4(0.999) && omg < R4(0.999)) {
          R qss = qp * (R(1.) - omp) * (R(1.) - omg);
          dq = dq + CO(C_EVAPS1) * sqrt(qss) + CO(C_EVAPS2) * fpow<R>(qss, pows2);
        }
        if (omp < R4(0.999) && omg > R4(0.001)) {
          R qgg = qp * (R(1.) - omp) * omg;
          dq = dq + CO(C_EVAPG1) * sqrt(qgg) + CO(C_EVAPG2) * fpow<R>(qgg, powg2);
        }
        dq = dq * dtn * (q / qs - R(1.));
        dq = fmax(-R(0.5) * qp, dq);
        qp = qp + dq;
        q = q - dq;
        evp = dq;
      } else {
        q = q + qp;
        evp = -qp;
        qp = R(0.);
      }
    }
    {
      R dq = qp;
      qp = fmax(R(0.0), qp);
      q = q + (dq - qp);
    }
    }  // plain microphysics
    // FP64: the level sums of src and evp are taken with diagnose's batched reduction; here
    // in FP32, or before the means of the mean-state acceleration
    constexpr bool batched = std::is_same<R, double>::value;
    if (accel || !batched) {
      R ssrc = s.sum(src), sevp = s.sum(evp);
      if (live && i == 0) {
        prof(d, P_QPSRC, c)[k] = ssrc;
        prof(d, P_QPEVP, c)[k] = sevp;
      }
    }
    // micro_diagnose, then diagnose (or, with mean-state acceleration pending, the means it
    // needs; k_accel then applies it and diagnoses)
    R qv, qcl, qci, qpl, qpi;
    micro_diag(q, qn, qp, tabs, a_bg, a_pr, qv, qcl, qci, qpl, qpi);
    // (input, target) pairs of the plain scheme for training the network
    if (d.samp_x && !net && live && o % (size_t)d.samp_stride == 0) {
      const float* xin = d.mlp_x + o * CRM_MLP_NIN;
      const unsigned long long slot = atomicAdd(d.samp_n, 1ull);
      if ((long long)slot < d.samp_cap) {
        float* xs = d.samp_x + slot * CRM_MLP_NIN;
        float* ts = d.samp_t + slot * CRM_MLP_NOUT;
        for (int f = 0; f < CRM_MLP_NIN; f++) xs[f] = xin[f];
        // the entry split again in FP64 (as k_mlp_inputs forms it; the fields still hold
        // the previous diagnose), so that the rates are differences of FP64 values
        const double qn0 = d.qn[o], tb = d.tabs[o];
        const double omn = fmax(0.0, fmin(1.0, (tb - TBGMIN) * (double)a_bg));
        const double omp = fmax(0.0, fmin(1.0, (tb - TPRMIN) * (double)a_pr));
        const double b_qv = (double)q_in - qn0, b_qcl = qn0 * omn, b_qci = qn0 * (1. - omn);
        const double b_qpl = (double)qp_in * omp, b_qpi = (double)qp_in * (1. - omp);
        const double b_T = (double)t - (double)gamaz[k] + FAC_COND_D * (b_qcl + b_qpl) + FAC_SUB_D * (b_qci + b_qpi);
        const double idtn = 1.0 / (double)dtn;
        const double tdiag = (double)t - (double)gamaz[k] + FAC_COND_D * ((double)qcl + (double)qpl) +
                             FAC_SUB_D * ((double)qci + (double)qpi);
        ts[0] = (float)(((double)qv - b_qv) * idtn);
        ts[1] = (float)(((double)qcl - b_qcl) * idtn);
        ts[2] = (float)(((double)qci - b_qci) * idtn);
        ts[3] = (float)(((double)qpl - b_qpl) * idtn);
        ts[4] = (float)(((double)qpi - b_qpi) * idtn);
        ts[5] = (float)((tdiag - b_T) * idtn);
      }
    }
    if (accel) {
      accel_means<NX, R>(d, s, c, live, k, o, y3, t, qv, qcl, qci);
      if (live) {
        d.qt[o] = q;
        d.qp[o] = qp;
        d.qn[o] = qn;
        d.qv[o] = qv;
        d.qcl[o] = qcl;
        d.qci[o] = qci;
        d.qpl[o] = qpl;
        d.qpi[o] = qpi;
      }
    } else {
      const R se[2] = {src, evp};
      diagnose_level<NX, R>(d, s, c, live, k, o, y3, t, q, qp, qn, qv, qcl, qci, qpl, qpi, u_k, v_k, kmin, kmax,
                            batched ? se : nullptr);
    }
    if constexpr (KUR) {  // k_kurant's level terms, from the same values (the lid's w counts as zero)
      const double uu = u_k, vv = y3 ? (double)v_k : 0.0, tkh = tkh_k;
      const double wm = fabs((double)w_k), wm_up = k + 1 < nzm ? fabs((double)w_k1) : 0.0;
      const double uhm = sqrt(uu * uu + vv * vv);
      const double dzz = d_dz * kur_adzw[k];
      const double sgs = fmax(fmax(0.5 * tkh * 1.0 * d.dt / (dzz * dzz), 0.5 * tkh * kur_grdf[k] * d.dt / (d.dx * d.dx)), 0.0);
      cfl = fmax(cfl, sgs);
      cfl = fmax(cfl, fmax(uhm * d.dt * kur_dxf, fmax(wm, wm_up) * d.dt / (d_dz * kur_adzw[k])));
    }
  }
  if constexpr (KUR) {
    cfl = s.max(cfl);
    if (live && i == 0) atomic_max_nonneg(&d.gmax[GMAX_FLAGS + (c % KURANT_SLOTS) * KURANT_STRIDE], cfl);
  }
  if (accel) return;
  kmin = s.imin(kmin);
  kmax = s.imax(kmax);
  if (live && i == 0) {
    inst(d, I_KMIN, c) = kmin;
    inst(d, I_KMAX, c) = kmax;
  }
}

// ------------------------------------------------------------------------------------------
// Mean-state acceleration: accelerate_crm.F90:109-366 (called between micro_proc and
// diagnose, crm_module.F90:872-876), then diagnose.  Unless a mean t change above 5 K was
// flagged in this call (gmax[3]; accelerate_crm.F90:255-278), t and qt (and u, v with
// crm_accel_uv) get crm_accel_factor times the change of their level mean since the last
// diagnose (computed by k_micro, profiles P_ACCT..P_ACCV); negative qt is clipped level by
// level with the excess removed from the positive values, and qv, qcl, qci re-split.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_accel(DevT<R> d, double factor, int uv) {
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, i = s.i;
  const bool y3 = d.ny > 1;
  const bool apply = *(volatile unsigned long long*)&d.gmax[3] == 0;
  const double tmin = 50.;
  int kmin = nzm, kmax = -1;
  for (int k = 0; k < nzm; k++) {
    size_t o = ((size_t)c * nzm + k) * NX + i;
    R t = d.t[o], q = d.qt[o], qp = d.qp[o], qn = d.qn[o];
    R qv = d.qv[o], qcl = d.qcl[o], qci = d.qci[o], qpl = d.qpl[o], qpi = d.qpi[o];
    R u = d.u[o], v = y3 ? d.v[o] : R(0.0);
    if (apply) {
      t = (R)fmax(tmin, (double)t + factor * prof(d, P_ACCT, c)[k]);
      if (uv) {
        u = (R)((double)u + factor * prof(d, P_ACCU, c)[k]);
        if (y3) v = (R)((double)v + factor * prof(d, P_ACCV, c)[k]);
      }
      double qt = (double)q + factor * prof(d, P_ACCQ, c)[k];
      double qneg = s.sum(qt < 0. ? qt : 0.0), qpoz = s.sum(qt < 0. ? 0.0 : qt);
      if (qpoz + qneg <= 0.) {
        qt = 0.;
        qv = qcl = qci = R(0.);
      } else {
        double f = 1. + qneg / qpoz;
        qt = fmax(0., qt * f);
        if (qt <= 0.) {
          qv = qcl = qci = R(0.);
        } else {
          double qt_res = qt - (double)qcl - (double)qci;
          qv = (R)fmax(0., qt_res);
          if (qt_res < 0.) {
            f = 1. + qt_res / ((double)qcl + (double)qci);
            qcl = (R)((double)qcl * f);
            qci = (R)((double)qci * f);
          }
        }
      }
      q = (R)qt;
      if (live) {
        d.t[o] = t;
        d.u[o] = u;
        if (y3) d.v[o] = v;
      }
    }
    diagnose_level<NX, R>(d, s, c, live, k, o, y3, t, q, qp, qn, qv, qcl, qci, qpl, qpi, u, v, kmin, kmax);
  }
  kmin = s.imin(kmin);
  kmax = s.imax(kmax);
  if (live && i == 0) {
    inst(d, I_KMIN, c) = kmin;
    inst(d, I_KMAX, c) = kmax;
  }
}

// ------------------------------------------------------------------------------------------
// Per-GCM-substep statistics: crm_module.F90:896-1073.  Accumulated on the device in
// profiles indexed by CRM level (w level for the mass fluxes) and per-instance sums.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_stats(DevT<R> d) {
  CRM_STOP_CHECK(d);
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX, false>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, i = s.i;
  const double dz = inst(d, I_DZ, c);
  const double* rho = prof(d, P_RHO, c);
  const double* rhow = prof(d, P_RHOW, c);
  const double* adz = prof(d, P_ADZ, c);
  const double* pres = prof(d, P_PRES, c);
  const double thr = R4(0.001), wmin = 2.;
  const int nxr = d.nx_rad, g = NX / nxr, ir = rad_col(d, i);
  auto Q = [&](const R* p, int k) -> double { return p[((size_t)c * nzm + k) * NX + i]; };
  auto Wv = [&](int k) -> double { return d.w[((size_t)c * nz + k) * NX + i]; };
  // radiation sums: exactly one lane owns each (instance, level, radiation column), so they
  // are plain read-modify-writes
  const bool rad_owner = live && rad_lead(d, i);
  const size_t rad_blk = (size_t)d.ncrms * nzm * nxr;
  auto rad_at = [&](int k) { return (size_t)c * nzm * nxr + (size_t)k * nxr + ir; };
  // the fields, profiles and radiation sums of one level; the sweep runs down from the top and
  // reads level k-1 while it works on level k.  Everything the level needs from memory is in
  // this read: vector loads complete in issue order, so a read issued after the prefetch (a
  // profile, or the radiation sums the level adds to) would wait for it.  The sums are R: in
  // the mixed engine FP32, which halves the bytes of this kernel's largest traffic (each
  // addition rounds once, from the FP64 sum of the two terms)
  struct Lvl {
    double qcl, qci, qpl, qpi, tabs, qv, w, rho, rhow, adz, pres;
    R ra[5];
  };
  auto load = [&](int k, Lvl& L) {
    L.qcl = Q(d.qcl, k);
    L.qci = Q(d.qci, k);
    L.qpl = Q(d.qpl, k);
    L.qpi = Q(d.qpi, k);
    L.tabs = Q(d.tabs, k);
    L.qv = Q(d.qv, k);
    L.w = Wv(k);
    L.rho = rho[k];
    L.rhow = rhow[k];
    L.adz = adz[k];
    L.pres = pres[k];
    for (int q = 0; q < 5; q++) L.ra[q] = R(0.0);
    if (rad_owner)
      for (int q = 0; q < 5; q++) L.ra[q] = d.rad[q * rad_blk + rad_at(k)];
  };
  double cwp = 0, cwph = 0, cwpm = 0, cwpl = 0, clt = 0, cmt = 0, cht = 0, ctt = 0;
  bool flag_top = true;
  double w_up = Wv(nzm);  // lid
  Lvl up2{}, up{}, cur, nxt{};
  load(nzm - 1, cur);
  // mass flux through a w level with vertical velocity wj (crm_module.F90:1017-1049): the level
  // whose condensate decides is Xu (the level below, or the surface level) for an updraft and
  // the level above (Xa, or Xb, the top level, when a is false) for a downdraft; one saturation
  // evaluation serves both branches
  // (the levels are passed as values, selected field by field: a reference chosen at run time
  // between two records would put them in memory)
  auto mass_flux = [&](double wj, double rhowj, const Lvl& Xu, const Lvl& Xa, const Lvl& Xb, bool a,
                       double& mui, double& mdi) {
    if (wj > 0. || wj < 0.) {
      const bool upd = wj > 0.;
      const double dt = a ? Xa.tabs : Xb.tabs, dp = a ? Xa.pres : Xb.pres;
      const double dn = a ? Xa.qcl + Xa.qci : Xb.qcl + Xb.qci, dq = a ? Xa.qpl + Xa.qpi : Xb.qpl + Xb.qpi;
      const double xt = upd ? Xu.tabs : dt, xp = upd ? Xu.pres : dp;
      const double xn = upd ? Xu.qcl + Xu.qci : dn;
      const double qs = qsatw(xt, xp);
      const bool cloudy = xn > fmin(R4(1.e-5), 0.01 * qs);
      // selects rather than stores under branches (a store to one of two locals chosen at run
      // time would put them in memory)
      const double f = rhowj * wj;
      mui = (upd && cloudy) ? f : 0.0;
      mdi = (!upd && (cloudy || dq > R4(1.0e-4))) ? f : 0.0;
    }
  };
  for (int k = nzm - 1; k >= 0; k--) {
    // at k = 0 level 0 again, unused (unconditional, as above); its radiation sums are read
    // before this level's are written, and never written from that read
    load(::max(k - 1, 0), nxt);
    const size_t rad_base = rad_at(k);
    const R* ra = cur.ra;
    const double qcl = cur.qcl, qci = cur.qci, qpl = cur.qpl, qpi = cur.qpi, tabs = cur.tabs;
    const double wk = cur.w;
    // column water paths from the top (the reference's level nz-k)
    double tmp1 = cur.rho * cur.adz * dz * (qcl + qci);
    cwp += tmp1;
    ctt = fmax(1.0, ctt);
    double top_hit = 0.0;
    if (cwp > thr && flag_top) {
      top_hit = 1.0;
      flag_top = false;
    }
    if (cur.pres >= 700.) {
      cwpl += tmp1;
      clt = fmax(1.0, clt);
    } else if (cur.pres < 400.) {
      cwph += tmp1;
      cht = fmax(1.0, cht);
    } else {
      cwpm += tmp1;
      cmt = fmax(1.0, cmt);
    }
    // level statistics (the reference's level k)
    double mass = cur.rho * cur.adz * dz;
    double ws = w_up + wk;
    double rw = cur.rho * 0.5 * ws;
    double cld = 0, mcup = 0, mcdn = 0, mcuup = 0, mcudn = 0;
    if (mass * (qcl + qci) > thr) {
      cld = 1.0;
      if (ws > 2 * wmin) {
        mcup = rw * 1.0;
        mcuup = rw * (1.0 - 1.0);
      }
      if (ws < -2 * wmin) {
        mcdn = rw * 1.0;
        mcudn = rw * (1. - 1.0);
      }
    } else {
      if (ws > 2 * wmin) mcuup = rw;
      if (ws < -2 * wmin) mcudn = rw;
    }
    // mass fluxes through w level k+1 (w level nzm, the lid, has none): its levels are all read
    // by now, so the sweep does not wait here for the read of level k-1 it just issued
    double mui = 0, mdi = 0;
    if (k + 1 < nzm) mass_flux(w_up, up.rhow, cur, up2, up, k + 2 < nzm, mui, mdi);
    double sv[10] = {top_hit, cld, mcup, mcdn, mcuup, mcudn, qcl, qci, mui, mdi};
    s.multi_sum(sv);
    // radiation-column sums (crm_module.F90:979-1015)
    double rs[5] = {tabs, fmax(0.0, cur.qv), qcl, qci, (qcl + qci > 0) ? 1.0 : 0.0};
    // the group's columns: contiguous lanes in 2D, a block of rows in 3D
    if (g > 1) s.gsums2(rs, d.rad_gx, d.rad_row, d.rad_gy);
    const double r_t = rs[0], r_qv = rs[1], r_qc = rs[2], r_qi = rs[3], r_cld = rs[4];
    if (live) {
      // the reference's index mix for the cloud top (crm_module.F90:918-925)
      if (s.owns(0)) accum(&prof(d, P_CLDTOP, c)[nzm - 1 - k], sv[0]);
      if (s.owns(1)) accum(&prof(d, P_CLD, c)[k], sv[1]);
      if (s.owns(2)) accum(&prof(d, P_MCUP, c)[k], sv[2]);
      if (s.owns(3)) accum(&prof(d, P_MCDN, c)[k], sv[3]);
      if (s.owns(4)) accum(&prof(d, P_MCUUP, c)[k], sv[4]);
      if (s.owns(5)) accum(&prof(d, P_MCUDN, c)[k], sv[5]);
      if (s.owns(6)) accum(&prof(d, P_GLIQ, c)[k], sv[6]);
      if (s.owns(7)) accum(&prof(d, P_GICE, c)[k], sv[7]);
      if (k + 1 < nzm && s.owns(8)) accum(&prof(d, P_MUI, c)[k + 1], sv[8]);
      if (k + 1 < nzm && s.owns(9)) accum(&prof(d, P_MDI, c)[k + 1], sv[9]);
      if (rad_owner) {
        d.rad[0 * rad_blk + rad_base] = (R)((double)ra[0] + r_t);
        d.rad[1 * rad_blk + rad_base] = (R)((double)ra[1] + r_qv);
        d.rad[2 * rad_blk + rad_base] = (R)((double)ra[2] + r_qc);
        d.rad[3 * rad_blk + rad_base] = (R)((double)ra[3] + r_qi);
        d.rad[4 * rad_blk + rad_base] = (R)((double)ra[4] + r_cld);
      }
    }
    w_up = wk;
    up2 = up;
    up = cur;
    cur = nxt;
  }
  {  // w level 0, with level 0 (now up) below and level 1 (up2) above
    double mui = 0, mdi = 0;
    mass_flux(w_up, up.rhow, up, up2, up, nzm > 1, mui, mdi);
    double sv[10] = {0, 0, 0, 0, 0, 0, 0, 0, mui, mdi};
    s.multi_sum(sv);
    if (live) {
      if (s.owns(8)) accum(&prof(d, P_MUI, c)[0], sv[8]);
      if (s.owns(9)) accum(&prof(d, P_MDI, c)[0], sv[9]);
    }
  }
  double cv[4] = {cwp > thr ? ctt : 0.0, cwph > thr ? cht : 0.0, cwpm > thr ? cmt : 0.0, cwpl > thr ? clt : 0.0};
  s.multi_sum(cv);
  if (live) {
    if (s.owns(0)) accum(&inst(d, I_CLTOT, c), cv[0]);
    if (s.owns(1)) accum(&inst(d, I_CLHGH, c), cv[1]);
    if (s.owns(2)) accum(&inst(d, I_CLMED, c), cv[2]);
    if (s.owns(3)) accum(&inst(d, I_CLLOW, c), cv[3]);
  }
}

// ------------------------------------------------------------------------------------------
// End of call: crm_module.F90:1078-1659.  State write-back and every crm_output member.
// ------------------------------------------------------------------------------------------
template <int NX, typename R>
__global__ void CRM_LB k_final(DevT<R> d, crm_input_view in, crm_state_view st, crm_rad_view rad, crm_output_view out,
                        double dt_gl, int nstop) {
  SegFor<NX> s;
  int c;
  bool live = seg_instance<NX>(d.ncrms, c);
  const int nzm = d.nzm, nz = d.nz, ld = d.ld, plev = d.plev, i = s.i;
  const double factor_xy = 1. / (double)NX, factor_xyt = factor_xy / (double)nstop;
  const double icrm_run_time = 1. / dt_gl, dz = inst(d, I_DZ, c);
  const cprof<double> rho = profcd(d, P_RHO, c);
  const cprof<double> rhow = profcd(d, P_RHOW, c);
  const cprof<double> u0 = profcd(d, P_U0, c);
  const cprof<double> v0 = profcd(d, P_V0, c);
  const bool y3 = d.ny > 1;
  const int ptop0 = plev - nzm;  // 0-based GCM index of the top CRM level
  auto IN = [&](const double* p, int l) { return p[(size_t)c + (size_t)ld * l]; };
  auto O2 = [&](double* p, int l) -> double& { return p[(size_t)c + (size_t)ld * l]; };
  auto G4 = [&](double* p, int k) -> double& { return p[(size_t)c + (size_t)ld * ((size_t)i + (size_t)NX * k)]; };
  auto Q = [&](const R* p, int k) -> double { return p[((size_t)c * nzm + k) * NX + i]; };
  auto Wv = [&](int k) { return d.w[((size_t)c * nz + k) * NX + i]; };
  const double a_gr = A_GR();
  double colprec = 0, colprecs = 0;
  for (int k = 0; k < nzm; k++) {
    int l = plev - 1 - k;
    double tabs = Q(d.tabs, k), qv = Q(d.qv, k), qcl = Q(d.qcl, k), qci = Q(d.qci, k);
    double qpl = Q(d.qpl, k), qpi = Q(d.qpi, k), u = Q(d.u, k), wk = Wv(k), wk1 = Wv(k + 1);
    double pd = IN(in.pdel, l);
    colprec += s.sum((qpl + qpi) * pd);
    colprecs += s.sum(qpi * pd);
    double tln = s.sum(tabs) * factor_xy, qln = s.sum(qv) * factor_xy;
    double qccln = s.sum(qcl) * factor_xy, qiiln = s.sum(qci) * factor_xy;
    double omg = fmax(0.0, fmin(1.0, (tabs - TGRMIN) * a_gr));
    double qc_mean = s.sum(qcl) * factor_xy, qi_mean = s.sum(qci) * factor_xy, qr_mean = s.sum(qpl) * factor_xy;
    double qg_mean = s.sum(qpi * omg) * factor_xy, qs_mean = s.sum(qpi * (1. - omg)) * factor_xy;
    double du = u - u0[k];
    double u2z = s.sum(du * du), w2z = s.sum(0.5 * (wk1 * wk1 + wk * wk));
    double v = 0.0, v2z = 0.0;
    if (y3) {
      v = Q(d.v, k);
      double dv = v - v0[k];
      v2z = s.sum(dv * dv);
    }
    double tke = s.sum(Q(d.tke, k)), tkz = s.sum(Q(d.tk, k));
    if (live) {
      // every prognostic value written back must be finite (CRM_ERR_NAN otherwise)
      const double chk = u + v + wk + tabs + Q(d.qt, k) + Q(d.qp, k) + Q(d.qn, k);
      if (!isfinite(chk)) atomicMax(&d.gmax[2], ERR_FLAG_NAN);
      G4(st.u_wind, k) = u;
      G4(st.v_wind, k) = v;
      G4(st.w_wind, k) = wk;
      G4(st.temperature, k) = tabs;
      G4(st.qt, k) = Q(d.qt, k);
      G4(st.qp, k) = Q(d.qp, k);
      G4(st.qn, k) = Q(d.qn, k);
      G4(out.tk, k) = Q(d.tk, k);
      G4(out.tkh, k) = Q(d.tkh, k);
      G4(out.qcl, k) = qcl;
      G4(out.qci, k) = qci;
      G4(out.qpl, k) = qpl;
      G4(out.qpi, k) = qpi;
      if (i == 0) {
        bool top2 = (l == ptop0 || l == ptop0 + 1);
        O2(out.sltend, l) = top2 ? 0.0 : CP_D * (tln - IN(in.tl, l)) * icrm_run_time;
        O2(out.qltend, l) = top2 ? 0.0 : (qln - IN(in.ql, l)) * icrm_run_time;
        O2(out.qcltend, l) = top2 ? 0.0 : (qccln - IN(in.qccl, l)) * icrm_run_time;
        O2(out.qiltend, l) = top2 ? 0.0 : (qiiln - IN(in.qiil, l)) * icrm_run_time;
        O2(out.qc_mean, l) = qc_mean;
        O2(out.qi_mean, l) = qi_mean;
        O2(out.qr_mean, l) = qr_mean;
        O2(out.qg_mean, l) = qg_mean;
        O2(out.qs_mean, l) = qs_mean;
        O2(out.cld, l) = fmin(1.0, prof(d, P_CLD, c)[k] * factor_xyt);
        O2(out.cldtop, l) = fmin(1.0, prof(d, P_CLDTOP, c)[k] * factor_xyt);
        O2(out.gicewp, l) = prof(d, P_GICE, c)[k] * pd * 1000. / GGR_D * factor_xyt;
        O2(out.gliqwp, l) = prof(d, P_GLIQ, c)[k] * pd * 1000. / GGR_D * factor_xyt;
        double mcup = prof(d, P_MCUP, c)[k] * factor_xyt, mcdn = prof(d, P_MCDN, c)[k] * factor_xyt;
        double mcuup = prof(d, P_MCUUP, c)[k] * factor_xyt, mcudn = prof(d, P_MCUDN, c)[k] * factor_xyt;
        O2(out.mcup, l) = mcup;
        O2(out.mcdn, l) = mcdn;
        O2(out.mcuup, l) = mcuup;
        O2(out.mcudn, l) = mcudn;
        O2(out.mctot, l) = mcup + mcdn + mcuup + mcudn;
        double t1 = dz / rhow[k], t2 = t1 / d.dtn;
        double mkwsb1 = prof(d, P_MKWSB1, c)[k] * t1 * rhow[k] * factor_xy / nstop;
        double mkwsb2 = prof(d, P_MKWSB2, c)[k] * t1 * rhow[k] * factor_xy / nstop;
        double mkwle1 = prof(d, P_MKWLE1, c)[k] * t2 * rhow[k] * factor_xy / nstop;
        double mkwle2 = prof(d, P_MKWLE2, c)[k] * t2 * rhow[k] * factor_xy / nstop;
        double mkadv1 = prof(d, P_MKADV1, c)[k] * factor_xy * icrm_run_time;
        double mkadv2 = prof(d, P_MKADV2, c)[k] * factor_xy * icrm_run_time;
        double mkdiff1 = prof(d, P_MKDIFF1, c)[k] * factor_xy * icrm_run_time;
        double mkdiff2 = prof(d, P_MKDIFF2, c)[k] * factor_xy * icrm_run_time;
        O2(out.flux_u, l) = (prof(d, P_UWLE, c)[k] + prof(d, P_UWSB, c)[k]) * t1 * factor_xy / nstop;
        O2(out.flux_v, l) = (y3 ? prof(d, P_VWLE, c)[k] + prof(d, P_VWSB, c)[k] : 0.0) * t1 * factor_xy / nstop;
        O2(out.flux_qt, l) = mkwle1 + mkwsb1;
        O2(out.fluxsgs_qt, l) = mkwsb1;
        O2(out.flux_qp, l) = mkwle2 + mkwsb2;
        O2(out.qt_trans, l) = mkadv1 + mkdiff1;
        O2(out.qp_trans, l) = mkadv2 + mkdiff2;
        double tkesgsz = rho[k] * tke * factor_xy;
        O2(out.tkesgsz, l) = tkesgsz;
        O2(out.tkez, l) = rho[k] * 0.5 * (u2z + v2z + w2z) * factor_xy + tkesgsz;
        O2(out.tkz, l) = tkz * factor_xy;
        O2(out.precflux, l) = prof(d, P_PRECFLUX, c)[k] * factor_xy * dz / d.dt / nstop / 1000.;
        O2(out.qp_fall, l) = prof(d, P_QPFALL, c)[k] * factor_xy * icrm_run_time;
        O2(out.qp_evp, l) = prof(d, P_QPEVP, c)[k] * factor_xy * icrm_run_time;
        O2(out.qp_src, l) = prof(d, P_QPSRC, c)[k] * factor_xy * icrm_run_time;
        O2(out.qt_ls, l) = prof(d, P_QTEND, c)[k];
        O2(out.t_ls, l) = prof(d, P_TTEND, c)[k];
      }
    }
  }
  // GCM levels above the CRM
  if (live && i == 0) {
    for (int l = 0; l < ptop0; l++) {
      O2(out.sltend, l) = 0.0;
      O2(out.qltend, l) = 0.0;
      O2(out.qcltend, l) = 0.0;
      O2(out.qiltend, l) = 0.0;
      double* zero2[] = {out.qc_mean, out.qi_mean, out.qr_mean, out.qg_mean, out.qs_mean, out.cld, out.cldtop,
                         out.gicewp, out.gliqwp, out.mcup, out.mcdn, out.mcuup, out.mcudn, out.mctot,
                         out.flux_u, out.flux_v, out.flux_qt, out.fluxsgs_qt, out.flux_qp, out.qt_trans,
                         out.qp_trans, out.tkesgsz, out.tkez, out.tkz, out.precflux, out.qp_fall, out.qp_evp,
                         out.qp_src, out.qt_ls, out.t_ls};
      for (double* p : zero2) O2(p, l) = 0.0;
    }
    // mass fluxes: interface l (0..plev) is w level plev-l
    double jt = plev * 1.0, mx = 1.0;
    const cprof<double> mui = profcd(d, P_MUI, c);
    const cprof<double> mdi = profcd(d, P_MDI, c);
    auto MU = [&](int lw) { int kw = plev - lw; return (kw >= 0 && kw <= nzm - 1) ? mui[kw] : 0.0; };
    auto MD = [&](int lw) { int kw = plev - lw; return (kw >= 0 && kw <= nzm - 1) ? mdi[kw] : 0.0; };
    for (int l = 0; l < plev; l++) {
      double pd = IN(in.pdel, l);
      double mu_a = MU(l), mu_b = MU(l + 1), md_a = MD(l), md_b = MD(l + 1);
      double mu = 0.5 * (mu_a + mu_b) * GGR_D / 100.;
      double md = 0.5 * (md_a + md_b) * GGR_D / 100.;
      O2(out.mu_crm, l) = mu;
      O2(out.md_crm, l) = md;
      O2(out.eu_crm, l) = 0.;
      O2(out.du_crm, l) = 0.;
      O2(out.ed_crm, l) = 0.;
      if (mu_a - mu_b > 0)
        O2(out.eu_crm, l) = (mu_a - mu_b) * GGR_D / pd;
      else
        O2(out.du_crm, l) = -1.0 * (mu_a - mu_b) * GGR_D / pd;
      if (md_b - md_a < 0) O2(out.ed_crm, l) = (md_a - md_b) * GGR_D / pd;
      if (fabs(mu) > R4(1.0e-15) || fabs(md) > R4(1.0e-15)) {
        jt = fmin((double)(l + 1), jt);
        mx = fmax((double)(l + 1), mx);
      }
    }
    out.jt_crm[c] = jt;
    out.mx_crm[c] = mx;
    out.prectend[c] = (colprec - inst(d, I_COLPREC, c)) / GGR_D * factor_xy * icrm_run_time;
    out.precstend[c] = (colprecs - inst(d, I_COLPRECS, c)) / GGR_D * factor_xy * icrm_run_time;
    out.cltot[c] = inst(d, I_CLTOT, c) * factor_xyt;
    out.clhgh[c] = inst(d, I_CLHGH, c) * factor_xyt;
    out.clmed[c] = inst(d, I_CLMED, c) * factor_xyt;
    out.cllow[c] = inst(d, I_CLLOW, c) * factor_xyt;
    out.z0m[c] = inst(d, I_Z0, c);
    out.taux[c] = 0.0 / (double)nstop;
    out.tauy[c] = 0.0 / (double)nstop;
    out.timing_factor[c] = (double)nstop / nstop;
  }
  // surface precipitation (crm_module.F90:1417-1481)
  double ps = d.precsfc[(size_t)c * NX + i] * dz / d.dt / (double)nstop;
  double pss = d.precssfc[(size_t)c * NX + i] * dz / d.dt / (double)nstop;
  bool conv = ps > (double)(10.f / 86400.f);
  double precc = s.sum(conv ? ps : 0.0), precsc = s.sum(conv ? pss : 0.0);
  double precl = s.sum(conv ? 0.0 : ps), precsl = s.sum(conv ? 0.0 : pss);
  if (live) {
    out.prec_crm[(size_t)c + (size_t)ld * i] = ps / 1000.;
    if (i == 0) {
      out.precc[c] = precc * factor_xy / 1000.;
      out.precl[c] = precl * factor_xy / 1000.;
      out.precsc[c] = precsc * factor_xy / 1000.;
      out.precsl[c] = precsl * factor_xy / 1000.;
    }
  }
  // time-averaged radiation groups
  if (live) {
    const int nxr = d.nx_rad;
    const double f = ((double)nxr / (double)NX) * 1.0 / (double)nstop;
    size_t blk = (size_t)d.ncrms * nzm * nxr;
    double* outs[5] = {rad.temperature, rad.qv, rad.qc, rad.qi, rad.cld};
    for (int idx = i; idx < nzm * nxr; idx += NX) {
      int k = idx / nxr, ir = idx - k * nxr;
      size_t src = (size_t)c * nzm * nxr + idx;
      size_t dst = (size_t)c + (size_t)ld * ((size_t)ir + (size_t)nxr * k);
      for (int q = 0; q < 5; q++) outs[q][dst] = (double)d.rad[q * blk + src] * f;
    }
  }
}

}  // namespace CRM_NS
