/*
 * crm_oracle.c -- CPU restatement of the reference cloud-resolving model crm()
 * (components/cam/src/physics/crm/crm_module.F90:45-1718) for the 1-moment microphysics,
 * MPDATA scalar advection and Smagorinsky SGS, 2D (ny == 1).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  It is never part of
 * the product path.
 *
 * Semantics kept on purpose (each is a property of the reference that parity depends on):
 *  - Default-kind literals in the reference are REAL(4) and are widened when they meet
 *    REAL(8) operands; R4(x) below reproduces (double)(float)x.
 *  - The x-transform of the pressure solve is the reference's fft991_crm, whose helper
 *    routines have implicit REAL(4) twiddles (fft.F90:247-295, 382-466, 469-808); it is
 *    restated here pass by pass.
 *  - ncycle (kurant.F90:53-64) and nprec (microphysics.F90:602-662) are batch-global maxima;
 *    the first-substep wind clamp is decided by instance 0 (crm_module.F90:399-417).
 *  - Halos are filled by explicit periodic copies with the reference's widths
 *    (periodic.F90, bound_exchange.F90) so that every stencil reads what the reference reads.
 *
 * Parity status per function is recorded in the test files: every phase is pinned against
 * the reference phase routines compiled from the reference tree (oracle/_ref), end to end
 * after one substep and after a full GCM step.
 *
 * Layout: field[c][k][i] with 1-based k and Fortran-style i offsets (see FLD below).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/crm_hip.h"

#define R4(x) ((double)(x##f))
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#define MIN(a, b) ((a) < (b) ? (a) : (b))

/* ---- constants: crm/params.F90:49-69, cime shr_const_mod.F90:22-55 ------------------- */
static const double SHR_RGAS = 6.02214e26 * 1.38065e-23;
#define CP 1.00464e3
#define GGR 9.80616
#define LCOND 2.501e6
#define LFUS 3.337e5
#define LSUB (LCOND + LFUS)
#define RGAS (SHR_RGAS / 28.966)
#define RV (SHR_RGAS / 18.016)
#define FAC_COND (LCOND / CP)
#define FAC_FUS (LFUS / CP)
#define FAC_SUB (LSUB / CP)
#define PI_R4 R4(3.141592653589793)
#define DIFFELQ R4(2.21e-05)
#define THERCO R4(2.40e-02)
#define MUELQ R4(1.717e-05)
#define EPSV R4(0.61)
/* micro_params.F90:12-61 */
#define RHOR 1000.0
#define RHOS 100.0
#define RHOG 400.0
#define TBGMIN R4(253.16)
#define TBGMAX R4(273.16)
#define TPRMIN R4(268.16)
#define TPRMAX R4(283.16)
#define TGRMIN R4(223.16)
#define TGRMAX R4(283.16)
#define A_RAIN 842.0
#define B_RAIN R4(0.8)
#define A_SNOW R4(4.84)
#define B_SNOW 0.25
#define A_GRAU 94.5
#define B_GRAU 0.5
#define QCW0 R4(1.e-3)
#define QCI0 R4(1.e-4)
#define ALPHAELQ R4(1.e-3)
#define BETAELQ R4(1.e-3)
#define ERCCOEF 1.0
#define ESCCOEF 1.0
#define ESICOEF R4(0.1)
#define EGCCOEF 1.0
#define EGICOEF R4(0.1)
#define NZEROR 8.e6
#define NZEROS 3.e6
#define NZEROG 4.e6
#define QP_THRESHOLD R4(1.e-8)

/* ---- storage ---------------------------------------------------------------------------- */
typedef struct {
  double *a;
  int i0, ni, nk;
} fld;
#define FLD(f, c, i, k) ((f).a[((size_t)(c) * (f).nk + ((k)-1)) * (f).ni + ((i) - (f).i0)])

static fld fld_new(int ncrms, int i0, int i1, int nk) {
  fld f;
  f.i0 = i0;
  f.ni = i1 - i0 + 1;
  f.nk = nk;
  f.a = (double *)calloc((size_t)ncrms * f.ni * nk, sizeof(double));
  return f;
}

typedef struct {
  int live, ncrms, nx, ny, nz, nzm, plev;
  double dx, dy, dt, dtn, dtfactor;
  int nstep, nstop, ncycle, icycle, na, nb, nc;
  double at, bt, ct, dt3[3];
  /* profiles, [c][k], k = 1..nz */
  double *z, *zi, *pres, *presi, *adz, *adzw, *rho, *rhow, *bet, *gamaz, *prespot;
  double *t0, *q0, *qv0, *tabs0, *u0, *v0, *tke0, *qn0, *qp0, *p0;
  double *ttend, *qtend, *utend, *vtend, *ug0, *vg0, *tg0, *qg0;
  double *grdf_x, *grdf_z;
  double *uwle, *uwsb, *mkwle[2], *mkwsb[2], *mkadv[2], *mkdiff[2];
  double *qpsrc, *qpevp, *qpfall, *precflux, *tlat, *qifall, *tlatqi;
  /* micro coefficients (precip_init.F90) [c][k] */
  double *accrsc, *accrsi, *accrrc, *coefice, *accrgc, *accrgi;
  double *evaps1, *evaps2, *evapr1, *evapr2, *evapg1, *evapg2;
  float gam3, gams1, gams2, gams3, gamg1, gamg2, gamg3, gamr1, gamr2, gamr3;
  double a_bg, a_pr, a_gr;
  double *dz; /* [c] */
  double *z0, *taux0, *tauy0;
  /* grid fields */
  fld u, w, t, qt, qp, tke, tk, tkh, p, tabs, qv, qcl, qci, qpl, qpi, qn, misc;
  fld dudt[3], dwdt[3];
  fld precsfc, precssfc; /* nk = 1 */
} oracle_state;

static oracle_state S;

static void state_free(void) {
  if (!S.live) return;
  double **profs[] = {&S.z, &S.zi, &S.pres, &S.presi, &S.adz, &S.adzw, &S.rho, &S.rhow,
                      &S.bet, &S.gamaz, &S.prespot, &S.t0, &S.q0, &S.qv0, &S.tabs0, &S.u0,
                      &S.v0, &S.tke0, &S.qn0, &S.qp0, &S.p0, &S.ttend, &S.qtend, &S.utend,
                      &S.vtend, &S.ug0, &S.vg0, &S.tg0, &S.qg0, &S.grdf_x, &S.grdf_z,
                      &S.uwle, &S.uwsb, &S.mkwle[0], &S.mkwle[1], &S.mkwsb[0], &S.mkwsb[1],
                      &S.mkadv[0], &S.mkadv[1], &S.mkdiff[0], &S.mkdiff[1], &S.qpsrc,
                      &S.qpevp, &S.qpfall, &S.precflux, &S.tlat, &S.qifall, &S.tlatqi,
                      &S.accrsc, &S.accrsi, &S.accrrc, &S.coefice, &S.accrgc, &S.accrgi,
                      &S.evaps1, &S.evaps2, &S.evapr1, &S.evapr2, &S.evapg1, &S.evapg2,
                      &S.dz, &S.z0, &S.taux0, &S.tauy0};
  for (size_t n = 0; n < sizeof(profs) / sizeof(profs[0]); n++) {
    free(*profs[n]);
    *profs[n] = NULL;
  }
  fld *fs[] = {&S.u, &S.w, &S.t, &S.qt, &S.qp, &S.tke, &S.tk, &S.tkh, &S.p, &S.tabs,
               &S.qv, &S.qcl, &S.qci, &S.qpl, &S.qpi, &S.qn, &S.misc, &S.dudt[0],
               &S.dudt[1], &S.dudt[2], &S.dwdt[0], &S.dwdt[1], &S.dwdt[2], &S.precsfc,
               &S.precssfc};
  for (size_t n = 0; n < sizeof(fs) / sizeof(fs[0]); n++) {
    free(fs[n]->a);
    fs[n]->a = NULL;
  }
  S.live = 0;
}

static void state_alloc(int ncrms, int nx, int nzm) {
  int nz = nzm + 1;
  S.ncrms = ncrms;
  S.nx = nx;
  S.ny = 1;
  S.nzm = nzm;
  S.nz = nz;
  double **profs[] = {&S.z, &S.zi, &S.pres, &S.presi, &S.adz, &S.adzw, &S.rho, &S.rhow,
                      &S.bet, &S.gamaz, &S.prespot, &S.t0, &S.q0, &S.qv0, &S.tabs0, &S.u0,
                      &S.v0, &S.tke0, &S.qn0, &S.qp0, &S.p0, &S.ttend, &S.qtend, &S.utend,
                      &S.vtend, &S.ug0, &S.vg0, &S.tg0, &S.qg0, &S.grdf_x, &S.grdf_z,
                      &S.uwle, &S.uwsb, &S.mkwle[0], &S.mkwle[1], &S.mkwsb[0], &S.mkwsb[1],
                      &S.mkadv[0], &S.mkadv[1], &S.mkdiff[0], &S.mkdiff[1], &S.qpsrc,
                      &S.qpevp, &S.qpfall, &S.precflux, &S.tlat, &S.qifall, &S.tlatqi,
                      &S.accrsc, &S.accrsi, &S.accrrc, &S.coefice, &S.accrgc, &S.accrgi,
                      &S.evaps1, &S.evaps2, &S.evapr1, &S.evapr2, &S.evapg1, &S.evapg2};
  for (size_t n = 0; n < sizeof(profs) / sizeof(profs[0]); n++)
    *profs[n] = (double *)calloc((size_t)ncrms * nz, sizeof(double));
  S.dz = (double *)calloc(ncrms, sizeof(double));
  S.z0 = (double *)calloc(ncrms, sizeof(double));
  S.taux0 = (double *)calloc(ncrms, sizeof(double));
  S.tauy0 = (double *)calloc(ncrms, sizeof(double));
  /* halo extents of grid.F90:31-46 with NADV = NADVS = 0 (2D) */
  S.u = fld_new(ncrms, -1, nx + 3, nzm);
  S.w = fld_new(ncrms, -1, nx + 2, nz);
  S.t = fld_new(ncrms, -2, nx + 3, nzm);
  S.qt = fld_new(ncrms, -2, nx + 3, nzm);
  S.qp = fld_new(ncrms, -2, nx + 3, nzm);
  S.tke = fld_new(ncrms, -2, nx + 3, nzm);
  S.tk = fld_new(ncrms, 0, nx + 1, nzm);
  S.tkh = fld_new(ncrms, 0, nx + 1, nzm);
  S.p = fld_new(ncrms, 0, nx, nzm);
  S.tabs = fld_new(ncrms, 1, nx, nzm);
  S.qv = fld_new(ncrms, 1, nx, nzm);
  S.qcl = fld_new(ncrms, 1, nx, nzm);
  S.qci = fld_new(ncrms, 1, nx, nzm);
  S.qpl = fld_new(ncrms, 1, nx, nzm);
  S.qpi = fld_new(ncrms, 1, nx, nzm);
  S.qn = fld_new(ncrms, 1, nx, nzm);
  S.misc = fld_new(ncrms, 1, nx, nz);
  for (int s = 0; s < 3; s++) {
    S.dudt[s] = fld_new(ncrms, 1, nx + 1, nzm);
    S.dwdt[s] = fld_new(ncrms, 1, nx, nz);
  }
  S.precsfc = fld_new(ncrms, 1, nx, 1);
  S.precssfc = fld_new(ncrms, 1, nx, 1);
  S.live = 1;
}

#define PR(a, c, k) ((a)[(size_t)(c) * S.nz + (k)-1])

/* ---- saturation functions: crm/sat.F90:12-139 ------------------------------------------- */
static double esatw(double t) {
  static const float a[9] = {6.105851f, 0.4440316f, 0.1430341e-1f, 0.2641412e-3f,
                             0.2995057e-5f, 0.2031998e-7f, 0.6936113e-10f, 0.2564861e-13f,
                             -0.3704404e-15f};
  double dt = t - R4(273.16);
  if (dt > -80.0)
    return a[0] + dt * (a[1] + dt * (a[2] + dt * (a[3] + dt * (a[4] + dt * (a[5] + dt * (a[6] + dt * (a[7] + a[8] * dt)))))));
  return R4(0.01) * 2.0 * exp(R4(9.550426) - R4(5723.265) / t + R4(3.53068) * log(t) - R4(0.00728332) * t);
}
static double qsatw(double t, double p) {
  double e = esatw(t);
  return R4(0.622) * e / MAX(e, p - e);
}
static double dtesatw(double t) {
  static const float a[9] = {0.443956472f, 0.285976452e-1f, 0.794747212e-3f, 0.121167162e-4f,
                             0.103167413e-6f, 0.385208005e-9f, -0.604119582e-12f,
                             -0.792933209e-14f, -0.599634321e-17f};
  double dt = t - R4(273.16);
  if (dt > -80.0)
    return a[0] + dt * (a[1] + dt * (a[2] + dt * (a[3] + dt * (a[4] + dt * (a[5] + dt * (a[6] + dt * (a[7] + a[8] * dt)))))));
  return esatw(t + 1) - esatw(t);
}
static double dtqsatw(double t, double p) { return R4(0.622) * dtesatw(t) / p; }
static double esati(double t) {
  static const float a[9] = {6.11147274f, 0.503160820f, 0.188439774e-1f, 0.420895665e-3f,
                             0.615021634e-5f, 0.602588177e-7f, 0.385852041e-9f,
                             0.146898966e-11f, 0.252751365e-14f};
  double dt = t - R4(273.16);
  if (dt > -80.0)
    return a[0] + dt * (a[1] + dt * (a[2] + dt * (a[3] + dt * (a[4] + dt * (a[5] + dt * (a[6] + dt * (a[7] + a[8] * dt)))))));
  return R4(0.01) * exp(R4(9.550426) - R4(5723.265) / t + R4(3.53068) * log(t) - R4(0.00728332) * t);
}
static double qsati(double t, double p) {
  double e = esati(t);
  return R4(0.622) * e / MAX(e, p - e);
}
static double dtesati(double t) {
  static const float a[9] = {0.503223089f, 0.377174432e-1f, 0.126710138e-2f, 0.249065913e-4f,
                             0.312668753e-6f, 0.255653718e-8f, 0.132073448e-10f,
                             0.390204672e-13f, 0.497275778e-16f};
  double dt = t - R4(273.16);
  if (dt > -80.0)
    return a[0] + dt * (a[1] + dt * (a[2] + dt * (a[3] + dt * (a[4] + dt * (a[5] + dt * (a[6] + dt * (a[7] + a[8] * dt)))))));
  return esati(t + 1.0) - esati(t);
}
static double dtqsati(double t, double p) { return R4(0.622) * dtesati(t) / p; }

/* exposed for the closed-form tests of the saturation polynomials */
double oracle_qsatw(double t, double p) { return qsatw(t, p); }
double oracle_qsati(double t, double p) { return qsati(t, p); }
double oracle_esatw(double t) { return esatw(t); }
double oracle_esati(double t) { return esati(t); }
double oracle_dtesatw(double t) { return dtesatw(t); }
double oracle_dtesati(double t) { return dtesati(t); }

/* gammafff (crm/gammafff.c:13): REAL(4) Gamma via exp(lgamma) */
static float gammafff(float x) { return (float)exp(lgamma((double)x)); }

/* ---- MT19937 as in cam/RNG_MT.F90:60-139 (note UMASK = -1147483648 there) -------------- */
static int32_t mt[624];
static int mti = 625;
static void mt_seed(int32_t seed) {
  mt[0] = seed;
  for (mti = 1; mti < 624; mti++) mt[mti] = (int32_t)((uint32_t)69069 * (uint32_t)mt[mti - 1]);
}
static double mt_rand(void) {
  const uint32_t UMASK = (uint32_t)-1147483648LL, LMASK = 2147483647u;
  const uint32_t mag01[2] = {0u, (uint32_t)-1727483681LL};
  uint32_t y;
  if (mti >= 624) {
    if (mti == 625) mt_seed(4357);
    int kk;
    for (kk = 0; kk < 624 - 397; kk++) {
      y = ((uint32_t)mt[kk] & UMASK) | ((uint32_t)mt[kk + 1] & LMASK);
      mt[kk] = (int32_t)((uint32_t)mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u]);
    }
    for (; kk < 623; kk++) {
      y = ((uint32_t)mt[kk] & UMASK) | ((uint32_t)mt[kk + 1] & LMASK);
      mt[kk] = (int32_t)((uint32_t)mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u]);
    }
    y = ((uint32_t)mt[623] & UMASK) | ((uint32_t)mt[0] & LMASK);
    mt[623] = (int32_t)((uint32_t)mt[396] ^ (y >> 1) ^ mag01[y & 1u]);
    mti = 0;
  }
  y = (uint32_t)mt[mti++];
  y ^= y >> 11;
  y ^= (y << 7) & (uint32_t)-1658038656LL;
  y ^= (y << 15) & (uint32_t)-272236544LL;
  y ^= y >> 18;
  int32_t ys = (int32_t)y;
  if (ys < 0) return ((double)ys + 4294967296.0) / 4294967296.0;
  return ((double)ys + 0.5) / 4294967296.0;
}
/* exposed for the generator test */
void oracle_mt_seed(int seed) { mt_seed(seed); }
double oracle_mt_rand(void) { return mt_rand(); }

/* ---- fft991_crm (crm/fft.F90:5-808), REAL(4) helper variables kept -------------------- */
static void fax(int *ifax, int n) {
  int nn = n / 2, k = 1;
  if (nn + nn != n) {
    ifax[0] = -99;
    return;
  }
  while (nn % 4 == 0) {
    ifax[k++] = 4;
    nn /= 4;
    if (nn == 1) goto done;
  }
  if (nn % 2 == 0) {
    ifax[k++] = 2;
    nn /= 2;
    if (nn == 1) goto done;
  }
  while (nn % 3 == 0) {
    ifax[k++] = 3;
    nn /= 3;
    if (nn == 1) goto done;
  }
  {
    int l = 5, inc = 2;
    for (;;) {
      while (nn % l == 0) {
        ifax[k++] = l;
        nn /= l;
        if (nn == 1) goto done;
      }
      l += inc;
      inc = 6 - inc;
    }
  }
done:
  ifax[0] = k - 1;
  int nfax = ifax[0];
  for (int ii = 2; ii <= nfax; ii++)
    for (int i = 1; i <= nfax + 1 - ii; i++)
      if (ifax[i + 1] < ifax[i]) {
        int t = ifax[i];
        ifax[i] = ifax[i + 1];
        ifax[i + 1] = t;
      }
}

/* fftrig_crm mode 3: every helper is implicit REAL(4) (fft.F90:247-295) */
static void fftrig(double *trigs, int n) {
  float pi = 2.0f * asinf(1.0f);
  int nn = n / 2;
  float del = (float)((double)(pi + pi) / (double)nn);
  int l = nn + nn;
  for (int i = 1; i <= l; i += 2) {
    float angle = (float)(0.5 * (double)(i - 1) * (double)del);
    trigs[i - 1] = cosf(angle);
    trigs[i] = sinf(angle);
  }
  del = 0.5f * del;
  int nh = (nn + 1) / 2;
  l = nh + nh;
  int la = nn + nn;
  for (int i = 1; i <= l; i += 2) {
    float angle = (float)(0.5 * (double)(i - 1) * (double)del);
    trigs[la + i - 1] = cosf(angle);
    trigs[la + i] = sinf(angle);
  }
}

/* vpassm_crm (fft.F90:469-808).  Arrays are 1-based in the reference; a,b,c,d here are
 * pointers to the element the reference calls a(1) etc. */
static void vpassm(const double *a, const double *b, double *c, double *d, const double *trigs,
                   int inc1, int inc2, int inc3, int inc4, int lot, int n, int ifac, int la) {
  const float sin36 = 0.587785252292473f, cos36 = 0.809016994374947f;
  const float sin72 = 0.951056516295154f, cos72 = 0.309016994374947f;
  const float sin60 = 0.866025403784437f;
  int m = n / ifac, iink = m * inc1, jink = la * inc2, jump = (ifac - 1) * jink;
  int ibase = 0, jbase = 0;
#define A(x) a[(x)-1]
#define B(x) b[(x)-1]
#define Cc(x) c[(x)-1]
#define D(x) d[(x)-1]
  if (ifac == 2) {
    int ia = 1, ja = 1, ib = ia + iink, jb = ja + jink;
    for (int l = 1; l <= la; l++) {
      int i = ibase, j = jbase;
      for (int ijk = 1; ijk <= lot; ijk++) {
        Cc(ja + j) = A(ia + i) + A(ib + i);
        D(ja + j) = B(ia + i) + B(ib + i);
        Cc(jb + j) = A(ia + i) - A(ib + i);
        D(jb + j) = B(ia + i) - B(ib + i);
        i += inc3;
        j += inc4;
      }
      ibase += inc1;
      jbase += inc2;
    }
    if (la == m) return;
    jbase += jump;
    for (int k = la + 1; k <= m; k += la) {
      int kb = k + k - 2;
      float c1 = (float)trigs[kb], s1 = (float)trigs[kb + 1];
      for (int l = 1; l <= la; l++) {
        int i = ibase, j = jbase;
        for (int ijk = 1; ijk <= lot; ijk++) {
          Cc(ja + j) = A(ia + i) + A(ib + i);
          D(ja + j) = B(ia + i) + B(ib + i);
          Cc(jb + j) = c1 * (A(ia + i) - A(ib + i)) - s1 * (B(ia + i) - B(ib + i));
          D(jb + j) = s1 * (A(ia + i) - A(ib + i)) + c1 * (B(ia + i) - B(ib + i));
          i += inc3;
          j += inc4;
        }
        ibase += inc1;
        jbase += inc2;
      }
      jbase += jump;
    }
    return;
  }
  if (ifac == 3) {
    int ia = 1, ja = 1, ib = ia + iink, jb = ja + jink, ic = ib + iink, jc = jb + jink;
    for (int l = 1; l <= la; l++) {
      int i = ibase, j = jbase;
      for (int ijk = 1; ijk <= lot; ijk++) {
        Cc(ja + j) = A(ia + i) + (A(ib + i) + A(ic + i));
        D(ja + j) = B(ia + i) + (B(ib + i) + B(ic + i));
        Cc(jb + j) = (A(ia + i) - 0.5 * (A(ib + i) + A(ic + i))) - (sin60 * (B(ib + i) - B(ic + i)));
        Cc(jc + j) = (A(ia + i) - 0.5 * (A(ib + i) + A(ic + i))) + (sin60 * (B(ib + i) - B(ic + i)));
        D(jb + j) = (B(ia + i) - 0.5 * (B(ib + i) + B(ic + i))) + (sin60 * (A(ib + i) - A(ic + i)));
        D(jc + j) = (B(ia + i) - 0.5 * (B(ib + i) + B(ic + i))) - (sin60 * (A(ib + i) - A(ic + i)));
        i += inc3;
        j += inc4;
      }
      ibase += inc1;
      jbase += inc2;
    }
    if (la == m) return;
    jbase += jump;
    for (int k = la + 1; k <= m; k += la) {
      int kb = k + k - 2, kc = kb + kb;
      float c1 = (float)trigs[kb], s1 = (float)trigs[kb + 1];
      float c2 = (float)trigs[kc], s2 = (float)trigs[kc + 1];
      for (int l = 1; l <= la; l++) {
        int i = ibase, j = jbase;
        for (int ijk = 1; ijk <= lot; ijk++) {
          double p1 = (A(ia + i) - 0.5 * (A(ib + i) + A(ic + i)));
          double q1 = (sin60 * (B(ib + i) - B(ic + i)));
          double p2 = (B(ia + i) - 0.5 * (B(ib + i) + B(ic + i)));
          double q2 = (sin60 * (A(ib + i) - A(ic + i)));
          Cc(ja + j) = A(ia + i) + (A(ib + i) + A(ic + i));
          D(ja + j) = B(ia + i) + (B(ib + i) + B(ic + i));
          Cc(jb + j) = c1 * (p1 - q1) - s1 * (p2 + q2);
          D(jb + j) = s1 * (p1 - q1) + c1 * (p2 + q2);
          Cc(jc + j) = c2 * (p1 + q1) - s2 * (p2 - q2);
          D(jc + j) = s2 * (p1 + q1) + c2 * (p2 - q2);
          i += inc3;
          j += inc4;
        }
        ibase += inc1;
        jbase += inc2;
      }
      jbase += jump;
    }
    return;
  }
  if (ifac == 4) {
    int ia = 1, ja = 1, ib = ia + iink, jb = ja + jink, ic = ib + iink, jc = jb + jink;
    int id = ic + iink, jd = jc + jink;
    for (int l = 1; l <= la; l++) {
      int i = ibase, j = jbase;
      for (int ijk = 1; ijk <= lot; ijk++) {
        Cc(ja + j) = (A(ia + i) + A(ic + i)) + (A(ib + i) + A(id + i));
        Cc(jc + j) = (A(ia + i) + A(ic + i)) - (A(ib + i) + A(id + i));
        D(ja + j) = (B(ia + i) + B(ic + i)) + (B(ib + i) + B(id + i));
        D(jc + j) = (B(ia + i) + B(ic + i)) - (B(ib + i) + B(id + i));
        Cc(jb + j) = (A(ia + i) - A(ic + i)) - (B(ib + i) - B(id + i));
        Cc(jd + j) = (A(ia + i) - A(ic + i)) + (B(ib + i) - B(id + i));
        D(jb + j) = (B(ia + i) - B(ic + i)) + (A(ib + i) - A(id + i));
        D(jd + j) = (B(ia + i) - B(ic + i)) - (A(ib + i) - A(id + i));
        i += inc3;
        j += inc4;
      }
      ibase += inc1;
      jbase += inc2;
    }
    if (la == m) return;
    jbase += jump;
    for (int k = la + 1; k <= m; k += la) {
      int kb = k + k - 2, kc = kb + kb, kd = kc + kb;
      float c1 = (float)trigs[kb], s1 = (float)trigs[kb + 1];
      float c2 = (float)trigs[kc], s2 = (float)trigs[kc + 1];
      float c3 = (float)trigs[kd], s3 = (float)trigs[kd + 1];
      for (int l = 1; l <= la; l++) {
        int i = ibase, j = jbase;
        for (int ijk = 1; ijk <= lot; ijk++) {
          double ap = A(ia + i) + A(ic + i), am = A(ia + i) - A(ic + i);
          double bp = B(ia + i) + B(ic + i), bm = B(ia + i) - B(ic + i);
          double cp = A(ib + i) + A(id + i), cm = A(ib + i) - A(id + i);
          double dp = B(ib + i) + B(id + i), dm = B(ib + i) - B(id + i);
          Cc(ja + j) = ap + cp;
          D(ja + j) = bp + dp;
          Cc(jc + j) = c2 * (ap - cp) - s2 * (bp - dp);
          D(jc + j) = s2 * (ap - cp) + c2 * (bp - dp);
          Cc(jb + j) = c1 * (am - dm) - s1 * (bm + cm);
          D(jb + j) = s1 * (am - dm) + c1 * (bm + cm);
          Cc(jd + j) = c3 * (am + dm) - s3 * (bm - cm);
          D(jd + j) = s3 * (am + dm) + c3 * (bm - cm);
          i += inc3;
          j += inc4;
        }
        ibase += inc1;
        jbase += inc2;
      }
      jbase += jump;
    }
    return;
  }
  if (ifac == 5) {
    int ia = 1, ja = 1, ib = ia + iink, jb = ja + jink, ic = ib + iink, jc = jb + jink;
    int id = ic + iink, jd = jc + jink, ie = id + iink, je = jd + jink;
    for (int pass = 0;; pass++) {
      float c1 = 1, s1 = 0, c2 = 1, s2 = 0, c3 = 1, s3 = 0, c4 = 1, s4 = 0;
      int kk = la + 1 + (pass - 1) * la;
      if (pass > 0) {
        if (kk > m) return;
        int kb = kk + kk - 2, kc = kb + kb, kd = kc + kb, ke = kd + kb;
        c1 = (float)trigs[kb];
        s1 = (float)trigs[kb + 1];
        c2 = (float)trigs[kc];
        s2 = (float)trigs[kc + 1];
        c3 = (float)trigs[kd];
        s3 = (float)trigs[kd + 1];
        c4 = (float)trigs[ke];
        s4 = (float)trigs[ke + 1];
      }
      for (int l = 1; l <= la; l++) {
        int i = ibase, j = jbase;
        for (int ijk = 1; ijk <= lot; ijk++) {
          double a1 = A(ia + i) + cos72 * (A(ib + i) + A(ie + i)) - cos36 * (A(ic + i) + A(id + i));
          double b1 = sin72 * (B(ib + i) - B(ie + i)) + sin36 * (B(ic + i) - B(id + i));
          double a2 = B(ia + i) + cos72 * (B(ib + i) + B(ie + i)) - cos36 * (B(ic + i) + B(id + i));
          double b2 = sin72 * (A(ib + i) - A(ie + i)) + sin36 * (A(ic + i) - A(id + i));
          double a3 = A(ia + i) - cos36 * (A(ib + i) + A(ie + i)) + cos72 * (A(ic + i) + A(id + i));
          double b3 = sin36 * (B(ib + i) - B(ie + i)) - sin72 * (B(ic + i) - B(id + i));
          double a4 = B(ia + i) - cos36 * (B(ib + i) + B(ie + i)) + cos72 * (B(ic + i) + B(id + i));
          double b4 = sin36 * (A(ib + i) - A(ie + i)) - sin72 * (A(ic + i) - A(id + i));
          Cc(ja + j) = A(ia + i) + (A(ib + i) + A(ie + i)) + (A(ic + i) + A(id + i));
          D(ja + j) = B(ia + i) + (B(ib + i) + B(ie + i)) + (B(ic + i) + B(id + i));
          if (pass == 0) {
            Cc(jb + j) = a1 - b1;
            Cc(je + j) = a1 + b1;
            D(jb + j) = a2 + b2;
            D(je + j) = a2 - b2;
            Cc(jc + j) = a3 - b3;
            Cc(jd + j) = a3 + b3;
            D(jc + j) = a4 + b4;
            D(jd + j) = a4 - b4;
          } else {
            Cc(jb + j) = c1 * (a1 - b1) - s1 * (a2 + b2);
            D(jb + j) = s1 * (a1 - b1) + c1 * (a2 + b2);
            Cc(je + j) = c4 * (a1 + b1) - s4 * (a2 - b2);
            D(je + j) = s4 * (a1 + b1) + c4 * (a2 - b2);
            Cc(jc + j) = c2 * (a3 - b3) - s2 * (a4 + b4);
            D(jc + j) = s2 * (a3 - b3) + c2 * (a4 + b4);
            Cc(jd + j) = c3 * (a3 + b3) - s3 * (a4 - b4);
            D(jd + j) = s3 * (a3 + b3) + c3 * (a4 - b4);
          }
          i += inc3;
          j += inc4;
        }
        ibase += inc1;
        jbase += inc2;
      }
      if (pass == 0 && la == m) return;
      jbase += jump;
    }
  }
#undef A
#undef B
#undef Cc
#undef D
}

/* fft99a_crm (fft.F90:297-378), lot = 1, inc = 1 */
static void fft99a(double *a, double *work, const double *trigs, int n) {
  int nh = n / 2;
  work[0] = a[0] + a[n];
  work[1] = a[0] - a[n];
  int iabase = 3, ibbase = n - 1, jabase = 3, jbbase = n - 1;
  for (int k = 3; k <= nh; k += 2) {
    float c = (float)trigs[n + k - 1], s = (float)trigs[n + k];
    int ia = iabase, ib = ibbase, ja = jabase, jb = jbbase;
    double A1 = a[ia - 1], B1 = a[ib - 1], A2 = a[ia], B2 = a[ib];
    work[ja - 1] = (A1 + B1) - (s * (A1 - B1) + c * (A2 + B2));
    work[jb - 1] = (A1 + B1) + (s * (A1 - B1) + c * (A2 + B2));
    work[ja] = (c * (A1 - B1) - s * (A2 + B2)) + (A2 - B2);
    work[jb] = (c * (A1 - B1) - s * (A2 + B2)) - (A2 - B2);
    iabase += 2;
    ibbase -= 2;
    jabase += 2;
    jbbase -= 2;
  }
  if (iabase == ibbase) {
    work[jabase - 1] = 2.0 * a[iabase - 1];
    work[jabase] = -2.0 * a[iabase];
  }
}

/* fft99b_crm (fft.F90:380-466), lot = 1, inc = 1; scale, c, s are implicit REAL(4) */
static void fft99b(const double *work, double *a, const double *trigs, int n) {
  int nh = n / 2;
  float scale = (float)(1.0 / (double)n);
  a[0] = scale * (work[0] + work[1]);
  a[n] = scale * (work[0] - work[1]);
  a[1] = 0.0;
  a[n + 1] = 0.0;
  scale = 0.5f * scale;
  int iabase = 3, ibbase = n - 1, jabase = 3, jbbase = n - 1;
  for (int k = 3; k <= nh; k += 2) {
    float c = (float)trigs[n + k - 1], s = (float)trigs[n + k];
    int ia = iabase, ib = ibbase, ja = jabase, jb = jbbase;
    double W1 = work[ia - 1], W2 = work[ib - 1], W3 = work[ia], W4 = work[ib];
    a[ja - 1] = scale * ((W1 + W2) + (c * (W3 + W4) + s * (W1 - W2)));
    a[jb - 1] = scale * ((W1 + W2) - (c * (W3 + W4) + s * (W1 - W2)));
    a[ja] = scale * ((c * (W1 - W2) - s * (W3 + W4)) + (W4 - W3));
    a[jb] = scale * ((c * (W1 - W2) - s * (W3 + W4)) - (W4 - W3));
    iabase += 2;
    ibbase -= 2;
    jabase += 2;
    jbbase -= 2;
  }
  if (iabase == ibbase) {
    scale = 2.0f * scale;
    a[jabase - 1] = scale * work[iabase - 1];
    a[jabase] = -scale * work[iabase];
  }
}

/* fft991_crm (fft.F90:5-153), inc = 1, lot = 1, jump = n+2.  a has n+2 entries. */
static void fft991(double *a, double *work, const double *trigs, const int *ifax, int n, int isign) {
  int nfax = ifax[0], nx = n + 1, nh = n / 2, ink = 2;
  int igo;
  (void)nx;
  if (isign == 1) {
    fft99a(a, work, trigs, n);
    igo = 60;
  } else {
    igo = 50;
    if (nfax % 2 != 1) {
      for (int m = 0; m < n; m++) work[m] = a[m];
      igo = 60;
    }
  }
  int la = 1;
  for (int k = 1; k <= nfax; k++) {
    if (igo == 60) {
      vpassm(&work[0], &work[1], &a[0], &a[1], trigs, 2, ink, n + 1, n + 2, 1, nh, ifax[k], la);
      igo = 50;
    } else {
      vpassm(&a[0], &a[1], &work[0], &work[1], trigs, ink, 2, n + 2, n + 1, 1, nh, ifax[k], la);
      igo = 60;
    }
    la *= ifax[k];
  }
  if (isign == -1) {
    fft99b(work, a, trigs, n);
    return;
  }
  if (nfax % 2 != 1)
    for (int m = 0; m < n; m++) a[m] = work[m];
  a[n] = 0.0;
  a[n + 1] = 0.0;
}

/* exposed for the transform tests */
void oracle_fft991(double *a, int n, int isign) {
  int ifax[20];
  double *trigs = (double *)calloc(3 * n + 8, sizeof(double));
  double *work = (double *)calloc(n + 8, sizeof(double));
  fax(ifax, n);
  fftrig(trigs, n);
  fft991(a, work, trigs, ifax, n, isign);
  free(trigs);
  free(work);
}

/* ---- halo exchange: bound_exchange.F90:189-237 (2D: east/west only) ---------------------- */
static void bound_x(fld *f, int i_1, int i_2) {
  int i1 = i_1 - 1, i2 = i_2 - 1, nx = S.nx;
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= f->nk; k++) {
      for (int i = -i1; i <= 0; i++) FLD(*f, c, i, k) = FLD(*f, c, nx + i, k);
      for (int i = nx + 1; i <= nx + 1 + i2; i++) FLD(*f, c, i, k) = FLD(*f, c, i - nx, k);
    }
}

/* periodic.F90:8-136 */
static void boundaries(int flag) {
  if (flag == 0) {
    bound_x(&S.u, 1, 1);
    /* w(:,:,nz) carries sstxy (zero in the CRM) through the exchange, then is reset */
    for (int c = 0; c < S.ncrms; c++)
      for (int i = 1; i <= S.nx; i++) FLD(S.w, c, i, S.nz) = 0.0;
    bound_x(&S.w, 1, 1);
    for (int c = 0; c < S.ncrms; c++)
      for (int i = 0; i <= S.nx + 1; i++) FLD(S.w, c, i, S.nz) = 0.0;
  } else if (flag == 2) {
    bound_x(&S.u, 2, 3);
    bound_x(&S.w, 2, 2);
    bound_x(&S.t, 3, 3);
    bound_x(&S.qt, 3, 3);
    bound_x(&S.qp, 3, 3);
  } else if (flag == 3) {
    bound_x(&S.t, 1, 1);
    bound_x(&S.qt, 1, 1);
    bound_x(&S.qp, 1, 1);
  } else if (flag == 4) {
    bound_x(&S.tk, 1, 1);
    bound_x(&S.tkh, 1, 1);
  }
}

/* bound_duvdt.F90:8-37 */
static void bound_duvdt(void) {
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++) FLD(S.dudt[S.na], c, S.nx + 1, k) = FLD(S.dudt[S.na], c, 1, k);
}

/* ---- kurant.F90:9-83 with kurant_sgs (SGS_TKE/sgs.F90:329-372) -------------------------- */
static int kurant(void) {
  int nzm = S.nzm, nx = S.nx;
  double cfl = 0.0;
  double *wm = (double *)calloc(S.nz + 1, sizeof(double));
  double *uhm = (double *)calloc(S.nz + 1, sizeof(double));
  for (int c = 0; c < S.ncrms; c++) {
    for (int k = 1; k <= S.nz; k++) wm[k] = uhm[k] = 0.0;
    for (int k = 1; k <= nzm; k++)
      for (int i = 1; i <= nx; i++) {
        wm[k] = MAX(wm[k], fabs(FLD(S.w, c, i, k)));
        double uu = FLD(S.u, c, i, k);
        uhm[k] = MAX(uhm[k], sqrt(uu * uu + 0.0));
      }
    double dxf = sqrt((1.0 / S.dx) * (1.0 / S.dx) + 0.0);
    for (int k = 1; k <= nzm; k++) {
      double tmp = MAX(uhm[k] * S.dt * dxf,
                       MAX(wm[k], wm[k + 1]) * S.dt / (S.dz[c] * PR(S.adzw, c, k)));
      cfl = MAX(cfl, tmp);
    }
  }
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      double tkhmax = 0.0;
      for (int i = 1; i <= nx; i++) tkhmax = MAX(tkhmax, FLD(S.tkh, c, i, k));
      double dzz = S.dz[c] * PR(S.adzw, c, k);
      double tmp = MAX(MAX(0.5 * tkhmax * PR(S.grdf_z, c, k) * S.dt / (dzz * dzz),
                           0.5 * tkhmax * PR(S.grdf_x, c, k) * S.dt / (S.dx * S.dx)),
                       0.0);
      cfl = MAX(cfl, tmp);
    }
  free(wm);
  free(uhm);
  int nc = (int)ceil(cfl / R4(0.7));
  S.ncycle = MAX(1, MAX(1, nc));
  return S.ncycle > 16 ? CRM_ERR_CFL : CRM_OK;
}

/* abcoefs.F90:7-33 */
static void abcoefs(void) {
  if (S.nstep >= 3) {
    double alpha = S.dt3[S.nb] / S.dt3[S.na];
    double beta = S.dt3[S.nc] / S.dt3[S.na];
    S.ct = (2. + 3. * alpha) / (6. * (alpha + beta) * beta);
    S.bt = -(1. + 2. * (alpha + beta) * S.ct) / (2. * alpha);
    S.at = 1. - S.bt - S.ct;
  } else if (S.nstep >= 2) {
    S.at = 3. / 2.;
    S.bt = -1. / 2.;
    S.ct = 0.;
  } else {
    S.at = 1.;
    S.bt = 0.;
    S.ct = 0.;
  }
}

/* zero.F90:8-25 */
static void zero(void) {
  for (int c = 0; c < S.ncrms; c++) {
    for (int k = 1; k <= S.nzm; k++)
      for (int i = 1; i <= S.nx + 1; i++) FLD(S.dudt[S.na], c, i, k) = 0.0;
    for (int k = 1; k <= S.nz; k++)
      for (int i = 1; i <= S.nx; i++) {
        FLD(S.dwdt[S.na], c, i, k) = 0.0;
        FLD(S.misc, c, i, k) = 0.0;
      }
  }
}

/* buoyancy.F90:7-38 */
static void buoyancy(void) {
  fld *dw = &S.dwdt[S.na];
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 2; k <= S.nzm; k++) {
      int kb = k - 1;
      double betu = PR(S.adz, c, kb) / (PR(S.adz, c, k) + PR(S.adz, c, kb));
      double betd = PR(S.adz, c, k) / (PR(S.adz, c, k) + PR(S.adz, c, kb));
      for (int i = 1; i <= S.nx; i++) {
        double up = PR(S.tabs0, c, k) * (EPSV * (FLD(S.qv, c, i, k) - PR(S.qv0, c, k)) -
                                         (FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k) - PR(S.qn0, c, k) +
                                          FLD(S.qpl, c, i, k) + FLD(S.qpi, c, i, k) - PR(S.qp0, c, k))) +
                    (FLD(S.tabs, c, i, k) - PR(S.tabs0, c, k)) *
                        (1. + EPSV * PR(S.qv0, c, k) - PR(S.qn0, c, k) - PR(S.qp0, c, k));
        double dn = PR(S.tabs0, c, kb) * (EPSV * (FLD(S.qv, c, i, kb) - PR(S.qv0, c, kb)) -
                                          (FLD(S.qcl, c, i, kb) + FLD(S.qci, c, i, kb) - PR(S.qn0, c, kb) +
                                           FLD(S.qpl, c, i, kb) + FLD(S.qpi, c, i, kb) - PR(S.qp0, c, kb))) +
                    (FLD(S.tabs, c, i, kb) - PR(S.tabs0, c, kb)) *
                        (1. + EPSV * PR(S.qv0, c, kb) - PR(S.qn0, c, kb) - PR(S.qp0, c, kb));
        FLD(*dw, c, i, k) = FLD(*dw, c, i, k) + PR(S.bet, c, k) * betu * up + PR(S.bet, c, kb) * betd * dn;
      }
    }
}

/* forcing.F90:7-79 */
static void forcing(void) {
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++) {
      double qneg = 0.0, qpoz = 0.0;
      int nneg = 0;
      for (int i = 1; i <= S.nx; i++) {
        FLD(S.t, c, i, k) = FLD(S.t, c, i, k) + PR(S.ttend, c, k) * S.dtn;
        FLD(S.qt, c, i, k) = FLD(S.qt, c, i, k) + PR(S.qtend, c, k) * S.dtn;
        if (FLD(S.qt, c, i, k) < 0.) {
          nneg++;
          qneg += FLD(S.qt, c, i, k);
        } else {
          qpoz += FLD(S.qt, c, i, k);
        }
        FLD(S.dudt[S.na], c, i, k) = FLD(S.dudt[S.na], c, i, k) + PR(S.utend, c, k);
      }
      if (nneg > 0 && qpoz + qneg > 0.) {
        double factor = 1. + qneg / qpoz;
        for (int i = 1; i <= S.nx; i++) FLD(S.qt, c, i, k) = MAX(0.0, FLD(S.qt, c, i, k) * factor);
      }
    }
}

/* damping.F90:9-117 */
static void damping(void) {
  const double tau_min = 60., tau_max = 450., damp_depth = R4(0.4);
  int nzm = S.nzm, nx = S.nx;
  for (int c = 0; c < S.ncrms; c++) {
    int n_damp = 0;
    for (int k = nzm; k >= 1; k--)
      if (PR(S.z, c, nzm) - PR(S.z, c, k) < damp_depth * PR(S.z, c, nzm)) n_damp = nzm - k + 1;
    for (int k = 1; k <= nzm; k++) {
      if (!(k >= nzm - n_damp)) continue;
      double tau = tau_min * pow(tau_max / tau_min, (PR(S.z, c, nzm) - PR(S.z, c, k)) /
                                                      (PR(S.z, c, nzm) - PR(S.z, c, nzm - n_damp)));
      tau = 1. / tau;
      double u0l = 0.0, t0l = 0.0;
      for (int i = 1; i <= nx; i++) {
        u0l += FLD(S.u, c, i, k) / (double)(nx);
        t0l += FLD(S.t, c, i, k) / (double)(nx);
      }
      for (int i = 1; i <= nx; i++) {
        FLD(S.dudt[S.na], c, i, k) = FLD(S.dudt[S.na], c, i, k) - (FLD(S.u, c, i, k) - u0l) * tau;
        FLD(S.dwdt[S.na], c, i, k) = FLD(S.dwdt[S.na], c, i, k) - FLD(S.w, c, i, k) * tau;
        FLD(S.t, c, i, k) = FLD(S.t, c, i, k) - S.dtn * (FLD(S.t, c, i, k) - t0l) * tau;
        FLD(S.qt, c, i, k) = FLD(S.qt, c, i, k) - S.dtn * (FLD(S.qv, c, i, k) - PR(S.qv0, c, k)) * tau;
      }
    }
  }
}

/* ice_fall.F90:6-173 */
static void ice_fall(void) {
  int nzm = S.nzm, nz = S.nz, nx = S.nx;
  double *fz = (double *)calloc((size_t)(nz + 2), sizeof(double));
  for (int c = 0; c < S.ncrms; c++) {
    int kmax = 0, kmin = nzm + 1;
    for (int i = 1; i <= nx; i++)
      for (int k = 1; k <= nzm; k++)
        if (FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k) > 0. && FLD(S.tabs, c, i, k) < R4(273.15)) {
          kmin = MIN(kmin, k);
          kmax = MAX(kmax, k);
        }
    for (int k = 1; k <= nzm; k++) PR(S.qifall, c, k) = PR(S.tlatqi, c, k) = 0.0;
    for (int i = 1; i <= nx; i++) {
      for (int k = 1; k <= nz; k++) fz[k] = 0.0;
      for (int k = MAX(1, kmin - 1); k <= kmax; k++) {
        int kc = MIN(nzm, k + 1), kb = MAX(1, k - 1);
        double coef = S.dtn / (0.5 * (PR(S.adz, c, kb) + PR(S.adz, c, k)) * S.dz[c]);
        double qiu = PR(S.rho, c, kc) * FLD(S.qci, c, i, kc);
        double qic = PR(S.rho, c, k) * FLD(S.qci, c, i, k);
        double qid = PR(S.rho, c, kb) * FLD(S.qci, c, i, kb);
        double vt_ice = MIN(R4(0.4), R4(8.66) * pow(MAX(0.0, qic) + R4(1.e-10), R4(0.24)));
        double tmp_phi;
        if (fabs(qic - qid) < R4(1.0e-25)) {
          tmp_phi = 0.;
        } else {
          double th = (qiu - qic) / (qic - qid);
          tmp_phi = MAX(0.0, MIN(MIN(0.5 * (1. + th), 2.0), 2. * th));
        }
        fz[k] = -vt_ice * (qic - 0.5 * (1. - coef * vt_ice) * tmp_phi * (qic - qid));
      }
      fz[nz] = 0.;
      for (int k = MAX(1, kmin - 2); k <= kmax; k++) {
        double coef = S.dtn / (S.dz[c] * PR(S.adz, c, k) * PR(S.rho, c, k));
        double dqi = coef * (fz[k] - fz[k + 1]);
        FLD(S.qt, c, i, k) = FLD(S.qt, c, i, k) + dqi;
        PR(S.qifall, c, k) += dqi;
        double lat_heat = (FAC_COND + FAC_FUS) * dqi;
        FLD(S.t, c, i, k) = FLD(S.t, c, i, k) - lat_heat;
        PR(S.tlatqi, c, k) -= lat_heat;
      }
      double coef = S.dtn / S.dz[c];
      double dqi = -coef * fz[1];
      FLD(S.precsfc, c, i, 1) += dqi;
      FLD(S.precssfc, c, i, 1) += dqi;
    }
  }
  free(fz);
}

/* ---- SGS: shear_prod2D.F90:7-96, tke_full.F90:11-317 (Smagorinsky branch) ------------- */
static void tke_full(void) {
  int nzm = S.nzm, nx = S.nx;
  const double Cs = R4(0.15), Ck = R4(0.1), Pr = 1.;
  const double Ce = Ck * Ck * Ck / (Cs * Cs * Cs * Cs);
  double *def2 = (double *)calloc((size_t)(nx + 2) * (nzm + 2), sizeof(double));
  double *bsv = (double *)calloc((size_t)(nx + 2) * (nzm + 2), sizeof(double));
  double *apb = (double *)calloc((size_t)(nx + 2) * (nzm + 2), sizeof(double));
#define D2(i, k) def2[(size_t)(k) * (nx + 2) + (i)]
#define BSV(i, k) bsv[(size_t)(k) * (nx + 2) + (i)]
#define APB(i, k) apb[(size_t)(k) * (nx + 2) + (i)]
  double rdx0 = 1. / S.dx;
  for (int c = 0; c < S.ncrms; c++) {
    double dz = S.dz[c];
    /* shear production (2D; v == 0 so its terms vanish exactly) */
    for (int k = 1; k <= nzm; k++) {
      int kb = k - 1, kc = k + 1;
      double rdz = 1. / (dz * PR(S.adz, c, k));
      double rdx = rdx0 * sqrt(S.dx * rdz);
      for (int i = 1; i <= nx; i++) {
        int ib = i - 1, ic = i + 1;
        double du = (FLD(S.u, c, ic, k) - FLD(S.u, c, i, k)) * rdx;
        double dw = (FLD(S.w, c, i, kc) - FLD(S.w, c, i, k)) * rdz;
        double base = 2. * (du * du + dw * dw);
        if (k >= 2 && k <= nzm - 1) {
          double rdzw_up = 1. / (dz * PR(S.adzw, c, kc));
          double rdzw_dn = 1. / (dz * PR(S.adzw, c, k));
          double rdx_up = rdx0 * sqrt(S.dx * rdzw_up);
          double rdx_dn = rdx0 * sqrt(S.dx * rdzw_dn);
          double a1 = (FLD(S.u, c, ic, kc) - PR(S.u0, c, kc) - FLD(S.u, c, ic, k) + PR(S.u0, c, k)) * rdzw_up +
                      (FLD(S.w, c, ic, kc) - FLD(S.w, c, i, kc)) * rdx_up;
          double a2 = (FLD(S.u, c, i, kc) - PR(S.u0, c, kc) - FLD(S.u, c, i, k) + PR(S.u0, c, k)) * rdzw_up +
                      (FLD(S.w, c, i, kc) - FLD(S.w, c, ib, kc)) * rdx_up;
          double a3 = (FLD(S.u, c, ic, k) - PR(S.u0, c, k) - FLD(S.u, c, ic, kb) + PR(S.u0, c, kb)) * rdzw_dn +
                      (FLD(S.w, c, ic, k) - FLD(S.w, c, i, k)) * rdx_dn;
          double a4 = (FLD(S.u, c, i, k) - PR(S.u0, c, k) - FLD(S.u, c, i, kb) + PR(S.u0, c, kb)) * rdzw_dn +
                      (FLD(S.w, c, i, k) - FLD(S.w, c, ib, k)) * rdx_dn;
          D2(i, k) = base + 0.5 * (a1 * a1 + a2 * a2 + a3 * a3 + a4 * a4);
        } else if (k == 1) {
          double rdzw_up = 1. / (dz * PR(S.adzw, c, kc));
          double rdx_up = rdx0 * sqrt(S.dx * rdzw_up);
          double a1 = (FLD(S.u, c, ic, kc) - PR(S.u0, c, kc) - FLD(S.u, c, ic, k) + PR(S.u0, c, k)) * rdzw_up +
                      (FLD(S.w, c, ic, kc) - FLD(S.w, c, i, kc)) * rdx_up;
          double a2 = (FLD(S.u, c, i, kc) - PR(S.u0, c, kc) - FLD(S.u, c, i, k) + PR(S.u0, c, k)) * rdzw_up +
                      (FLD(S.w, c, i, kc) - FLD(S.w, c, ib, kc)) * rdx_up;
          D2(i, k) = base + 0.5 * (a1 * a1 + a2 * a2);
        } else {
          double rdzw_dn = 1. / (dz * PR(S.adzw, c, k));
          double rdx_dn = rdx0 * sqrt(S.dx * rdzw_dn);
          double a3 = (FLD(S.u, c, ic, k) - PR(S.u0, c, k) - FLD(S.u, c, ic, kb) + PR(S.u0, c, kb)) * rdzw_dn +
                      (FLD(S.w, c, ic, k) - FLD(S.w, c, i, k)) * rdx_dn;
          double a4 = (FLD(S.u, c, i, k) - PR(S.u0, c, k) - FLD(S.u, c, i, kb) + PR(S.u0, c, kb)) * rdzw_dn +
                      (FLD(S.w, c, i, k) - FLD(S.w, c, ib, k)) * rdx_dn;
          D2(i, k) = base + 0.5 * (a3 * a3 + a4 * a4);
        }
      }
    }
    /* SGS buoyancy flux on interfaces, tke_full.F90:120-237 */
    for (int i = 1; i <= nx; i++) {
      BSV(i, 0) = APB(i, 0) = 0.0;
      BSV(i, nzm) = APB(i, nzm) = 0.0;
    }
    for (int k = 1; k <= nzm - 1; k++) {
      int kc = k + 1, kb = k;
      double betdz = 0.5 * (PR(S.bet, c, kc) + PR(S.bet, c, kb)) / dz / PR(S.adzw, c, k + 1);
      for (int i = 1; i <= nx; i++) {
        double tabs_kc = FLD(S.tabs, c, i, kc), tabs_kb = FLD(S.tabs, c, i, kb);
        double qcl_kc = FLD(S.qcl, c, i, kc), qcl_kb = FLD(S.qcl, c, i, kb);
        double qci_kc = FLD(S.qci, c, i, kc), qci_kb = FLD(S.qci, c, i, kb);
        double qv_kc = FLD(S.qv, c, i, kc), qv_kb = FLD(S.qv, c, i, kb);
        double qpl_kc = FLD(S.qpl, c, i, kc), qpl_kb = FLD(S.qpl, c, i, kb);
        double qpi_kc = FLD(S.qpi, c, i, kc), qpi_kb = FLD(S.qpi, c, i, kb);
        double tabs_if = 0.5 * (tabs_kc + FAC_COND * qcl_kc + FAC_SUB * qci_kc + tabs_kb +
                                FAC_COND * qcl_kb + FAC_SUB * qci_kb);
        double qtot_if = 0.5 * (qv_kc + qcl_kc + qci_kc + qv_kb + qcl_kb + qci_kb);
        double qp_if = 0.5 * (qpl_kc + qpi_kc + qpl_kb + qpi_kb);
        double bbb = 1. + EPSV * qtot_if - qp_if;
        double dtt = FLD(S.t, c, i, kc) - FLD(S.t, c, i, kb);
        double dqt = qv_kc + qcl_kc + qci_kc - qv_kb - qcl_kb - qci_kb;
        double buoy = betdz * (bbb * dtt + EPSV * tabs_if * dqt + (bbb * FAC_COND - tabs_if) * (qpl_kc - qpl_kb) +
                               (bbb * FAC_SUB - tabs_if) * (qpi_kc - qpi_kb));
        double tkhs = FLD(S.tkh, c, i, kc) + FLD(S.tkh, c, i, kb) + R4(0.002);
        BSV(i, k) = buoy;
        APB(i, k) = -0.5 * tkhs * buoy;
        double qctot = qcl_kc + qci_kc + qcl_kb + qci_kb;
        if (qctot > 0.) {
          double omn = (qcl_kc + qcl_kb) / (qctot + R4(1.e-20));
          double pp = PR(S.presi, c, k + 1);
          double qsat_check = omn * qsatw(tabs_if, pp) + (1. - omn) * qsati(tabs_if, pp);
          if (qtot_if > qsat_check) {
            double lstarn = FAC_COND + (1. - omn) * FAC_FUS;
            double tif = 0.5 * (tabs_kc + tabs_kb);
            double dqsat = omn * dtqsatw(tif, pp) + (1. - omn) * dtqsati(tif, pp);
            double qsatt = omn * qsatw(tif, pp) + (1. - omn) * qsati(tif, pp);
            double b2 = 1. + EPSV * qsatt + qsatt - qtot_if - qp_if + R4(1.61) * tif * dqsat;
            b2 = b2 / (1. + lstarn * dqsat);
            double tk_ = FLD(S.tabs, c, i, k);
            buoy = betdz * (b2 * dtt + (b2 * lstarn - (1. + lstarn * dqsat) * tif) * dqt +
                            (b2 * FAC_COND - (1. + FAC_COND * dqsat) * tk_) * (qpl_kc - qpl_kb) +
                            (b2 * FAC_SUB - (1. + FAC_SUB * dqsat) * tk_) * (qpi_kc - qpi_kb));
            BSV(i, k) = buoy;
            APB(i, k) = -0.5 * tkhs * buoy;
          }
        }
      }
    }
    /* eddy viscosity, tke_full.F90:246-308 */
    for (int k = 1; k <= nzm - 1; k++) {
      double grd = dz * PR(S.adz, c, k);
      double Ce1 = Ce / R4(0.7) * R4(0.19);
      double Ce2 = Ce / R4(0.7) * R4(0.51);
      double cx = S.dx * S.dx / S.dt / PR(S.grdf_x, c, k);
      double cy = S.dx * S.dx / S.dt / PR(S.grdf_x, c, k);
      double dzm = dz * MIN(PR(S.adzw, c, k), PR(S.adzw, c, k + 1));
      double cz = dzm * dzm / S.dt / PR(S.grdf_z, c, k);
      double tkmax = R4(0.09) / (1. / cx + 1. / cy + 1. / cz);
      for (int i = 1; i <= nx; i++) {
        double buoy = 0.5 * (BSV(i, k - 1) + BSV(i, k));
        double smix;
        if (buoy <= 0.)
          smix = grd;
        else
          smix = MIN(grd, MAX(R4(0.1) * grd, sqrt(R4(0.76) * FLD(S.tk, c, i, k) / Ck / sqrt(buoy + R4(1.e-10)))));
        double ratio = smix / grd;
        double Cee = Ce1 + Ce2 * ratio;
        double tk = sqrt(Ck * Ck * Ck / Cee * MAX(0.0, D2(i, k) - Pr * buoy)) * (smix * smix);
        double tt = tk / (Ck * smix);
        FLD(S.tke, c, i, k) = tt * tt;
        tk = MIN(tk, tkmax);
        FLD(S.tk, c, i, k) = tk;
        FLD(S.tkh, c, i, k) = Pr * tk;
      }
    }
  }
#undef D2
#undef BSV
#undef APB
  free(def2);
  free(bsv);
  free(apb);
}

/* ---- momentum advection: advect2_mom_xy.F90:60-84 (2D), advect2_mom_z.F90:7-118 ------- */
static void advect_mom(void) {
  int nzm = S.nzm, nx = S.nx, nz = S.nz;
  fld *du = &S.dudt[S.na], *dw = &S.dwdt[S.na];
  double dx25 = 0.25 / S.dx;
  double *fuz = (double *)calloc(nz + 2, sizeof(double));
  double *fwz = (double *)calloc(nz + 2, sizeof(double));
  for (int c = 0; c < S.ncrms; c++) {
    for (int k = 1; k <= nzm; k++) {
      int kc = k + 1, kcu = MIN(kc, nzm);
      double irho = 1. / (PR(S.rhow, c, kc) * PR(S.adzw, c, kc));
      double ra = PR(S.rho, c, k) * PR(S.adz, c, k), rau = PR(S.rho, c, kcu) * PR(S.adz, c, kcu);
      for (int i = 1; i <= nx; i++) {
        double u_i = FLD(S.u, c, i, k), u_m = FLD(S.u, c, i - 1, k), u_p = FLD(S.u, c, i + 1, k);
        double fu1 = dx25 * (u_i + u_m) * (u_m + u_i);
        double fu2 = dx25 * (u_p + u_i) * (u_i + u_p);
        FLD(*du, c, i, k) = FLD(*du, c, i, k) - (fu2 - fu1);
        double fw1 = dx25 * (u_i * PR(S.rho, c, k) * PR(S.adz, c, k) + FLD(S.u, c, i, kcu) * PR(S.rho, c, kcu) * PR(S.adz, c, kcu)) *
                     (FLD(S.w, c, i - 1, kc) + FLD(S.w, c, i, kc));
        double fw2 = dx25 * (u_p * PR(S.rho, c, k) * PR(S.adz, c, k) + FLD(S.u, c, i + 1, kcu) * PR(S.rho, c, kcu) * PR(S.adz, c, kcu)) *
                     (FLD(S.w, c, i, kc) + FLD(S.w, c, i + 1, kc));
        (void)ra;
        (void)rau;
        FLD(*dw, c, i, kc) = FLD(*dw, c, i, kc) - irho * (fw2 - fw1);
      }
    }
    for (int k = 1; k <= nz; k++) PR(S.uwle, c, k) = 0.0;
    double dz25 = 1. / (4. * S.dz[c]);
    for (int i = 1; i <= nx; i++) {
      fuz[1] = 0.;
      fuz[nz] = 0.;
      for (int k = 2; k <= nzm; k++) {
        double rhoi = dz25 * PR(S.rhow, c, k);
        double www = rhoi * (FLD(S.w, c, i, k) + FLD(S.w, c, i - 1, k));
        fuz[k] = www * (FLD(S.u, c, i, k) + FLD(S.u, c, i, k - 1));
        PR(S.uwle, c, k) += fuz[k];
      }
      for (int k = 1; k <= nzm; k++) {
        int kc = k + 1;
        double rhoi = 1. / (PR(S.rho, c, k) * PR(S.adz, c, k));
        FLD(*du, c, i, k) = FLD(*du, c, i, k) - (fuz[kc] - fuz[k]) * rhoi;
        fwz[k] = dz25 * (FLD(S.w, c, i, kc) * PR(S.rhow, c, kc) + FLD(S.w, c, i, k) * PR(S.rhow, c, k)) *
                 (FLD(S.w, c, i, kc) + FLD(S.w, c, i, k));
      }
      for (int k = 2; k <= nzm; k++) {
        double rhoi = 1. / (PR(S.rhow, c, k) * PR(S.adzw, c, k));
        FLD(*dw, c, i, k) = FLD(*dw, c, i, k) - (fwz[k] - fwz[k - 1]) * rhoi;
      }
    }
  }
  free(fuz);
  free(fwz);
}

/* ---- SGS momentum diffusion: diffuse_mom2D.F90:7-161 --------------------------------- */
static void diffuse_mom(void) {
  int nzm = S.nzm, nx = S.nx, nz = S.nz;
  fld *du = &S.dudt[S.na], *dw = &S.dwdt[S.na];
  double rdx2 = 1. / S.dx / S.dx, rdx25 = 0.25 * rdx2;
  size_t ni = nx + 1;
  double *fu = (double *)calloc(ni * (nz + 1), sizeof(double));
  double *fw = (double *)calloc(ni * (nz + 1), sizeof(double));
#define FU(i, k) fu[(size_t)(k) * ni + (i)]
#define FW(i, k) fw[(size_t)(k) * ni + (i)]
  for (int c = 0; c < S.ncrms; c++) {
    double dz = S.dz[c];
    for (int k = 1; k <= nzm; k++) {
      int kc = k + 1, kcu = MIN(kc, nzm);
      double dxz = S.dx / (dz * PR(S.adzw, c, kc));
      double rdx21 = rdx2 * PR(S.grdf_x, c, k), rdx251 = rdx25 * PR(S.grdf_x, c, k);
      for (int i = 0; i <= nx; i++) {
        int ic = i + 1;
        double tkx = rdx21 * FLD(S.tk, c, i, k);
        FU(i, k) = -2. * tkx * (FLD(S.u, c, ic, k) - FLD(S.u, c, i, k));
        tkx = rdx251 * (FLD(S.tk, c, i, k) + FLD(S.tk, c, ic, k) + FLD(S.tk, c, i, kcu) + FLD(S.tk, c, ic, kcu));
        FW(i, k) = -tkx * (FLD(S.w, c, ic, kc) - FLD(S.w, c, i, kc) + (FLD(S.u, c, ic, kcu) - FLD(S.u, c, ic, k)) * dxz);
      }
      for (int i = 1; i <= nx; i++) {
        FLD(*du, c, i, k) = FLD(*du, c, i, k) - (FU(i, k) - FU(i - 1, k));
        FLD(*dw, c, i, kc) = FLD(*dw, c, i, kc) - (FW(i, k) - FW(i - 1, k));
      }
    }
    for (int k = 1; k <= nzm; k++) PR(S.uwsb, c, k) = 0.0;
    for (int k = 1; k <= nzm - 1; k++) {
      int kc = k + 1;
      double rdz = 1. / dz;
      double rdz2 = rdz * rdz * PR(S.grdf_z, c, k);
      double rdz25 = 0.25 * rdz2;
      double dzx = dz / S.dx;
      double iadz = 1. / PR(S.adz, c, k), iadzw = 1. / PR(S.adzw, c, kc);
      for (int i = 1; i <= nx; i++) {
        int ib = i - 1;
        double tkz = rdz2 * FLD(S.tk, c, i, k);
        FW(i, kc) = -2. * tkz * (FLD(S.w, c, i, kc) - FLD(S.w, c, i, k)) * PR(S.rho, c, k) * iadz;
        tkz = rdz25 * (FLD(S.tk, c, i, k) + FLD(S.tk, c, ib, k) + FLD(S.tk, c, i, kc) + FLD(S.tk, c, ib, kc));
        FU(i, kc) = -tkz * ((FLD(S.u, c, i, kc) - FLD(S.u, c, i, k)) * iadzw + (FLD(S.w, c, i, kc) - FLD(S.w, c, ib, kc)) * dzx) *
                    PR(S.rhow, c, kc);
        PR(S.uwsb, c, kc) += FU(i, kc);
      }
    }
    for (int i = 1; i <= nx; i++) {
      double rdz = 1. / dz;
      double rdz2 = rdz * rdz * PR(S.grdf_z, c, nzm - 1);
      double tkz = rdz2 * PR(S.grdf_z, c, nzm) * FLD(S.tk, c, i, nzm);
      FW(i, nz) = -2. * tkz * (FLD(S.w, c, i, nz) - FLD(S.w, c, i, nzm)) / PR(S.adz, c, nzm) * PR(S.rho, c, nzm);
      FU(i, 1) = 0.0 * rdz * PR(S.rhow, c, 1);
      FU(i, nz) = 0.0 * rdz * PR(S.rhow, c, nz);
      PR(S.uwsb, c, 1) += FU(i, 1);
    }
    for (int k = 1; k <= nzm; k++) {
      double rhoi = 1. / (PR(S.rho, c, k) * PR(S.adz, c, k));
      for (int i = 1; i <= nx; i++) FLD(*du, c, i, k) = FLD(*du, c, i, k) - (FU(i, k + 1) - FU(i, k)) * rhoi;
    }
    for (int k = 2; k <= nzm; k++) {
      double rhoi = 1. / (PR(S.rhow, c, k) * PR(S.adzw, c, k));
      for (int i = 1; i <= nx; i++) FLD(*dw, c, i, k) = FLD(*dw, c, i, k) - (FW(i, k + 1) - FW(i, k)) * rhoi;
    }
  }
#undef FU
#undef FW
  free(fu);
  free(fw);
}

/* ---- pressure: pressure.F90:8-323, press_rhs.F90:8-106, press_grad.F90:8-76 --------- */
static void pressure(void) {
  int nzm = S.nzm, nx = S.nx, nz = S.nz;
  int na = S.na, nb = S.nb, nc = S.nc;
  int nx2 = nx + 2;
  bound_duvdt();
  double rdx = 1. / S.dx, btat = S.bt / S.at, ctat = S.ct / S.at;
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      int kc = k + 1;
      double rdz = 1. / (PR(S.adz, c, k) * S.dz[c]);
      double rup = PR(S.rhow, c, kc) / PR(S.rho, c, k) * rdz;
      double rdn = PR(S.rhow, c, k) / PR(S.rho, c, k) * rdz;
      double dta = 1. / S.dt3[na] / S.at;
      for (int i = 1; i <= nx; i++) {
        int ic = i + 1;
        double pv = (rdx * (FLD(S.u, c, ic, k) - FLD(S.u, c, i, k)) + (FLD(S.w, c, i, kc) * rup - FLD(S.w, c, i, k) * rdn)) * dta +
                    (rdx * (FLD(S.dudt[na], c, ic, k) - FLD(S.dudt[na], c, i, k)) +
                     (FLD(S.dwdt[na], c, i, kc) * rup - FLD(S.dwdt[na], c, i, k) * rdn)) +
                    btat * (rdx * (FLD(S.dudt[nb], c, ic, k) - FLD(S.dudt[nb], c, i, k)) +
                            (FLD(S.dwdt[nb], c, i, kc) * rup - FLD(S.dwdt[nb], c, i, k) * rdn)) +
                    ctat * (rdx * (FLD(S.dudt[nc], c, ic, k) - FLD(S.dudt[nc], c, i, k)) +
                            (FLD(S.dwdt[nc], c, i, kc) * rup - FLD(S.dwdt[nc], c, i, k) * rdn));
        FLD(S.p, c, i, k) = pv * PR(S.rho, c, k);
      }
    }
  int ifax[20];
  double *trigs = (double *)calloc(3 * nx + 8, sizeof(double));
  double *work = (double *)calloc(nx + 8, sizeof(double));
  double *f = (double *)calloc((size_t)nx2 * (nzm + 1), sizeof(double));
  double *a = (double *)calloc(nzm + 2, sizeof(double));
  double *cc = (double *)calloc(nzm + 2, sizeof(double));
  double *alfa = (double *)calloc(nzm + 2, sizeof(double));
  double *beta = (double *)calloc(nzm + 2, sizeof(double));
  double *eign = (double *)calloc(nx + 2, sizeof(double));
  fax(ifax, nx);
  fftrig(trigs, nx);
  {
    double ddx2 = 1.0 / (S.dx * S.dx), pii = 3.14159265358979323846, xnx = pii / nx;
    for (int i = 1; i <= nx + 1; i++) {
      int id = (int)(((float)i - 0.1f) / 2.0f);
      double xi = id;
      eign[i] = (2.0 * cos(2.0 * xnx * xi) - 2.0) * ddx2;
    }
  }
#define FF(i, k) f[(size_t)((k)-1) * nx2 + (i)-1]
  for (int c = 0; c < S.ncrms; c++) {
    double dz = S.dz[c];
    for (int k = 1; k <= nzm; k++) {
      for (int i = 1; i <= nx; i++) FF(i, k) = FLD(S.p, c, i, k);
      FF(nx + 1, k) = 0.0;
      FF(nx + 2, k) = 0.0;
      fft991(&FF(1, k), work, trigs, ifax, nx, -1);
    }
    for (int k = 1; k <= nzm; k++) {
      a[k] = PR(S.rhow, c, k) / (PR(S.adz, c, k) * PR(S.adzw, c, k) * dz * dz);
      cc[k] = PR(S.rhow, c, k + 1) / (PR(S.adz, c, k) * PR(S.adzw, c, k + 1) * dz * dz);
    }
    for (int i = 1; i <= nx + 1; i++) {
      int id = (int)(((float)i - 0.1f) / 2.0f);
      double b, e;
      if (id == 0)
        b = 1.0 / (eign[i] * PR(S.rho, c, 1) - a[1] - cc[1]);
      else
        b = 1.0 / (eign[i] * PR(S.rho, c, 1) - cc[1]);
      alfa[1] = -cc[1] * b;
      beta[1] = FF(i, 1) * b;
      for (int k = 2; k <= nzm - 1; k++) {
        e = 1.0 / (eign[i] * PR(S.rho, c, k) - a[k] - cc[k] + a[k] * alfa[k - 1]);
        alfa[k] = -cc[k] * e;
        beta[k] = (FF(i, k) - a[k] * beta[k - 1]) * e;
      }
      FF(i, nzm) = (FF(i, nzm) - a[nzm] * beta[nzm - 1]) /
                   (eign[i] * PR(S.rho, c, nzm) - a[nzm] + a[nzm] * alfa[nzm - 1]);
      for (int k = nzm - 1; k >= 1; k--) FF(i, k) = alfa[k] * FF(i, k + 1) + beta[k];
    }
    for (int k = 1; k <= nzm; k++) {
      fft991(&FF(1, k), work, trigs, ifax, nx, 1);
      for (int i = 1; i <= nx; i++) FLD(S.p, c, i, k) = FF(i, k);
      FLD(S.p, c, 0, k) = FF(nx, k);
    }
  }
#undef FF
  free(trigs);
  free(work);
  free(f);
  free(a);
  free(cc);
  free(alfa);
  free(beta);
  free(eign);
  /* press_grad */
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      int kb = MAX(1, k - 1);
      double rdz = 1. / (S.dz[c] * PR(S.adzw, c, k));
      for (int i = 1; i <= nx; i++) {
        FLD(S.dudt[na], c, i, k) = FLD(S.dudt[na], c, i, k) - (FLD(S.p, c, i, k) - FLD(S.p, c, i - 1, k)) * rdx;
        FLD(S.dwdt[na], c, i, k) = FLD(S.dwdt[na], c, i, k) - (FLD(S.p, c, i, k) - FLD(S.p, c, i, kb)) * rdz;
      }
    }
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= nzm; k++)
      for (int i = 0; i <= nx; i++) FLD(S.p, c, i, k) = FLD(S.p, c, i, k) * PR(S.rho, c, k);
  bound_duvdt();
  (void)nz;
}

/* adams.F90:7-40 */
static void adams(void) {
  double dtdx = S.dtn / S.dx;
  int na = S.na, nb = S.nb, nc = S.nc;
  for (int c = 0; c < S.ncrms; c++) {
    double dtdz = S.dtn / S.dz[c];
    for (int k = 1; k <= S.nzm; k++) {
      double rhox = PR(S.rho, c, k) * dtdx, rhoz = PR(S.rhow, c, k) * dtdz;
      for (int i = 1; i <= S.nx; i++) {
        double un = FLD(S.u, c, i, k) + S.dt3[na] * (S.at * FLD(S.dudt[na], c, i, k) + S.bt * FLD(S.dudt[nb], c, i, k) +
                                                      S.ct * FLD(S.dudt[nc], c, i, k));
        double wn = FLD(S.w, c, i, k) + S.dt3[na] * (S.at * FLD(S.dwdt[na], c, i, k) + S.bt * FLD(S.dwdt[nb], c, i, k) +
                                                      S.ct * FLD(S.dwdt[nc], c, i, k));
        FLD(S.dudt[nc], c, i, k) = un;
        FLD(S.dwdt[nc], c, i, k) = wn;
        FLD(S.u, c, i, k) = 0.5 * (FLD(S.u, c, i, k) + un) * rhox;
        FLD(S.misc, c, i, k) = 0.5 * (FLD(S.w, c, i, k) + wn);
        FLD(S.w, c, i, k) = 0.5 * (FLD(S.w, c, i, k) + wn) * rhoz;
      }
    }
  }
}

/* uvw.F90:6-27 */
static void uvw(void) {
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++)
      for (int i = 1; i <= S.nx; i++) {
        FLD(S.u, c, i, k) = FLD(S.dudt[S.nc], c, i, k);
        FLD(S.w, c, i, k) = FLD(S.dwdt[S.nc], c, i, k);
      }
}

/* ---- MPDATA: ADV_MPDATA/advect_scalar2D.F90:8-244 and advect_scalar.F90:9-76 --------- */
static double andiff(double x1, double x2, double a, double b) { return (fabs(a) - a * a * b) * 0.5 * (x2 - x1); }
static double across(double x1, double a1, double a2) { return 0.03125 * a1 * a2 * x1; }
static double pp(double y) { return MAX(0.0, y); }
static double pn(double y) { return -MIN(0.0, y); }

static void advect_scalar(fld *fp, double *fadv, double *flux) {
  int nzm = S.nzm, nx = S.nx, nz = S.nz;
  const double eps = R4(1.e-10);
  int nm = nx + 2;      /* mx, mn: i = 0..nx+1 */
  int nu = nx + 5;      /* uuu: i = -1..nx+3 */
  int nw = nx + 4;      /* www: i = -1..nx+2 */
  double *mx = (double *)malloc(sizeof(double) * nm * (nzm + 1));
  double *mn = (double *)malloc(sizeof(double) * nm * (nzm + 1));
  double *uuu = (double *)malloc(sizeof(double) * nu * (nzm + 1));
  double *www = (double *)malloc(sizeof(double) * nw * (nz + 1));
  double *f0 = (double *)malloc(sizeof(double) * nx * (nzm + 1));
#define MX(i, k) mx[(size_t)(k) * nm + (i)]
#define MN(i, k) mn[(size_t)(k) * nm + (i)]
#define UU(i, k) uuu[(size_t)(k) * nu + (i) + 1]
#define WW(i, k) www[(size_t)(k) * nw + (i) + 1]
#define FV(i, k) FLD(*fp, c, i, k)
#define UV(i, k) FLD(S.u, c, i, k)
#define WV(i, k) FLD(S.w, c, i, k)
  for (int c = 0; c < S.ncrms; c++) {
    for (int k = 1; k <= nzm; k++)
      for (int i = 1; i <= nx; i++) f0[(size_t)k * nx + i - 1] = FV(i, k);
    for (int i = -1; i <= nx + 2; i++) WW(i, nz) = 0.;
    for (int k = 1; k <= nzm; k++) {
      int kc = MIN(nzm, k + 1), kb = MAX(1, k - 1);
      for (int i = 0; i <= nx + 1; i++) {
        MX(i, k) = MAX(MAX(MAX(MAX(FV(i - 1, k), FV(i + 1, k)), FV(i, kb)), FV(i, kc)), FV(i, k));
        MN(i, k) = MIN(MIN(MIN(MIN(FV(i - 1, k), FV(i + 1, k)), FV(i, kb)), FV(i, kc)), FV(i, k));
      }
    }
    for (int k = 1; k <= nzm; k++) {
      int kb = MAX(1, k - 1);
      for (int i = -1; i <= nx + 3; i++) {
        UU(i, k) = MAX(0.0, UV(i, k)) * FV(i - 1, k) + MIN(0.0, UV(i, k)) * FV(i, k);
        if (i <= nx + 2) WW(i, k) = MAX(0.0, WV(i, k)) * FV(i, kb) + MIN(0.0, WV(i, k)) * FV(i, k);
      }
      flux[k - 1] = 0.;
    }
    for (int k = 1; k <= nzm; k++) {
      double irho = 1. / PR(S.rho, c, k), iadz = 1. / PR(S.adz, c, k);
      for (int i = -1; i <= nx + 2; i++) {
        if (i >= 1 && i <= nx) flux[k - 1] += WW(i, k);
        FV(i, k) = FV(i, k) - (UU(i + 1, k) - UU(i, k) + (WW(i, k + 1) - WW(i, k)) * iadz) * irho;
      }
    }
    for (int k = 1; k <= nzm; k++) {
      int kc = MIN(nzm, k + 1), kb = MAX(1, k - 1);
      double dd = (double)(2.0f / (float)(kc - kb)) / PR(S.adz, c, k);
      double irho = 1. / PR(S.rho, c, k);
      double irhow = 1. / (PR(S.rhow, c, k) * PR(S.adz, c, k));
      for (int i = 0; i <= nx + 2; i++) {
        int ib = i - 1;
        UU(i, k) = andiff(FV(ib, k), FV(i, k), UV(i, k), irho) -
                   across(dd * (FV(ib, kc) + FV(i, kc) - FV(ib, kb) - FV(i, kb)), UV(i, k),
                          WV(ib, k) + WV(ib, kc) + WV(i, k) + WV(i, kc)) * irho;
        if (i <= nx + 1) {
          int ic = i + 1;
          WW(i, k) = andiff(FV(i, kb), FV(i, k), WV(i, k), irhow) -
                     across(FV(ic, kb) + FV(ic, k) - FV(ib, kb) - FV(ib, k), WV(i, k),
                            UV(i, kb) + UV(i, k) + UV(ic, k) + UV(ic, kb)) * irho;
        }
      }
    }
    for (int i = -1; i <= nx + 2; i++) WW(i, 1) = 0.;
    for (int k = 1; k <= nzm; k++) {
      int kc = MIN(nzm, k + 1), kb = MAX(1, k - 1);
      for (int i = 0; i <= nx + 1; i++) {
        MX(i, k) = MAX(MAX(MAX(MAX(MAX(FV(i - 1, k), FV(i + 1, k)), FV(i, kb)), FV(i, kc)), FV(i, k)), MX(i, k));
        MN(i, k) = MIN(MIN(MIN(MIN(MIN(FV(i - 1, k), FV(i + 1, k)), FV(i, kb)), FV(i, kc)), FV(i, k)), MN(i, k));
      }
    }
    for (int k = 1; k <= nzm; k++) {
      int kc = MIN(nzm, k + 1);
      double iadz = 1. / PR(S.adz, c, k), rho = PR(S.rho, c, k);
      for (int i = 0; i <= nx + 1; i++) {
        int ic = i + 1;
        MX(i, k) = rho * (MX(i, k) - FV(i, k)) / (pn(UU(ic, k)) + pp(UU(i, k)) + iadz * (pn(WW(i, kc)) + pp(WW(i, k))) + eps);
        MN(i, k) = rho * (FV(i, k) - MN(i, k)) / (pp(UU(ic, k)) + pn(UU(i, k)) + iadz * (pp(WW(i, kc)) + pn(WW(i, k))) + eps);
      }
    }
    for (int k = 1; k <= nzm; k++) {
      int kb = MAX(1, k - 1);
      for (int i = 1; i <= nx + 1; i++) {
        int ib = i - 1;
        UU(i, k) = pp(UU(i, k)) * MIN(MIN(1.0, MX(i, k)), MN(ib, k)) - pn(UU(i, k)) * MIN(MIN(1.0, MX(ib, k)), MN(i, k));
        if (i <= nx) {
          WW(i, k) = pp(WW(i, k)) * MIN(MIN(1.0, MX(i, k)), MN(i, kb)) - pn(WW(i, k)) * MIN(MIN(1.0, MX(i, kb)), MN(i, k));
          flux[k - 1] += WW(i, k);
        }
      }
    }
    for (int k = 1; k <= nzm; k++) {
      double irho = 1. / PR(S.rho, c, k), iadz = 1. / PR(S.adz, c, k);
      for (int i = 1; i <= nx; i++)
        FV(i, k) = MAX(0.0, FV(i, k) - (UU(i + 1, k) - UU(i, k) + (WW(i, k + 1) - WW(i, k)) * iadz) * irho);
    }
    for (int k = 1; k <= nzm; k++) {
      fadv[k - 1] = 0.;
      for (int i = 1; i <= nx; i++) fadv[k - 1] += FV(i, k) - f0[(size_t)k * nx + i - 1];
    }
    fadv += S.nz;
    flux += S.nz;
  }
#undef MX
#undef MN
#undef UU
#undef WW
#undef FV
#undef UV
#undef WV
  free(mx);
  free(mn);
  free(uuu);
  free(www);
  free(f0);
}

/* ---- precipitation sedimentation: MICRO_SAM1MOM/microphysics.F90:453-833 ------------- */
typedef struct {
  double vrain, vsnow, vgrau, crain, csnow, cgrau;
} termv_t;

static double term_vel_qp(const termv_t *tv, double qploc, double rho, double tabs) {
  double term_vel = 0.;
  if (qploc > QP_THRESHOLD) {
    double omp = MAX(0.0, MIN(1.0, (tabs - TPRMIN) * S.a_pr));
    if (omp == 1.) {
      term_vel = tv->vrain * pow(rho * qploc, tv->crain);
    } else if (omp == 0.) {
      double omg = MAX(0.0, MIN(1.0, (tabs - TGRMIN) * S.a_gr));
      double qgg = omg * qploc;
      double qss = qploc - qgg;
      term_vel = (omg * tv->vgrau * pow(rho * qgg, tv->cgrau) + (1. - omg) * tv->vsnow * pow(rho * qss, tv->csnow));
    } else {
      double omg = MAX(0.0, MIN(1.0, (tabs - TGRMIN) * S.a_gr));
      double qrr = omp * qploc;
      double qss = qploc - qrr;
      double qgg = omg * qss;
      qss = qss - qgg;
      term_vel = (omp * tv->vrain * pow(rho * qrr, tv->crain) +
                  (1. - omp) * (omg * tv->vgrau * pow(rho * qgg, tv->cgrau) + (1. - omg) * tv->vsnow * pow(rho * qss, tv->csnow)));
    }
  }
  return term_vel;
}

/* largest sedimentation sub-step count of the current call (reported as "nprec_max") */
static int oracle_nprec_max = 1;

static void precip_fall(void) {
  int nzm = S.nzm, nz = S.nz, nx = S.nx, ncrms = S.ncrms;
  termv_t tv;
  tv.crain = B_RAIN / 4.;
  tv.csnow = B_SNOW / 4.;
  tv.cgrau = B_GRAU / 4.;
  tv.vrain = A_RAIN * (double)S.gamr3 / 6. / pow(PI_R4 * RHOR * NZEROR, tv.crain);
  tv.vsnow = A_SNOW * (double)S.gams3 / 6. / pow(PI_R4 * RHOS * NZEROS, tv.csnow);
  tv.vgrau = A_GRAU * (double)S.gamg3 / 6. / pow(PI_R4 * RHOG * NZEROG, tv.cgrau);
  const double eps = R4(1.e-10);
  size_t ncol = (size_t)ncrms * nx;
  size_t nk = nz + 1;
  double *omega = (double *)calloc(ncol * nk, sizeof(double));
  double *lfac = (double *)calloc(ncol * nk, sizeof(double));
  double *wp = (double *)calloc(ncol * nk, sizeof(double));
  double *fz = (double *)calloc(ncol * nk, sizeof(double));
  double *www = (double *)calloc(ncol * nk, sizeof(double));
  double *mx = (double *)calloc(ncol * nk, sizeof(double));
  double *mn = (double *)calloc(ncol * nk, sizeof(double));
  double *tq = (double *)calloc(ncol * nk, sizeof(double));
  double *rhofac = (double *)calloc((size_t)ncrms * nk, sizeof(double));
  double *irhoadz = (double *)calloc((size_t)ncrms * nk, sizeof(double));
  double *iwmax = (double *)calloc((size_t)ncrms * nk, sizeof(double));
#define CI(c, i) ((size_t)(c) * nx + (i)-1)
#define A3(arr, c, i, k) arr[CI(c, i) * nk + (k)]
#define A2(arr, c, k) arr[(size_t)(c) * nk + (k)]
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      A2(rhofac, c, k) = sqrt(R4(1.29) / PR(S.rho, c, k));
      A2(irhoadz, c, k) = 1. / (PR(S.rho, c, k) * PR(S.adz, c, k));
      int kb = MAX(1, k - 1);
      double wmax = S.dz[c] * PR(S.adz, c, kb) / S.dtn;
      A2(iwmax, c, k) = 1. / wmax;
    }
  double prec_cfl = 0.;
  for (int c = 0; c < ncrms; c++)
    for (int i = 1; i <= nx; i++)
      for (int k = 1; k <= nzm; k++) {
        double om = MAX(0.0, MIN(1.0, (FLD(S.tabs, c, i, k) - TPRMIN) * S.a_pr));
        A3(omega, c, i, k) = om;
        A3(lfac, c, i, k) = FAC_COND + (1 - om) * FAC_FUS;
        double tmp = term_vel_qp(&tv, FLD(S.qp, c, i, k), PR(S.rho, c, k), FLD(S.tabs, c, i, k));
        double w_ = A2(rhofac, c, k) * tmp;
        prec_cfl = MAX(prec_cfl, w_ * A2(iwmax, c, k));
        A3(wp, c, i, k) = -w_ * PR(S.rhow, c, k) * S.dtn / S.dz[c];
        A3(fz, c, i, nz) = 0.;
        A3(www, c, i, nz) = 0.;
        A3(lfac, c, i, nz) = 0.;
      }
  int nprec = 1;
  if (prec_cfl > R4(0.9)) {
    nprec = (int)ceil(prec_cfl / R4(0.9));
    if (nprec > oracle_nprec_max) oracle_nprec_max = nprec;
    for (int c = 0; c < ncrms; c++)
      for (int i = 1; i <= nx; i++)
        for (int k = 1; k <= nzm; k++) A3(wp, c, i, k) = A3(wp, c, i, k) / (double)nprec;
  }
  for (int iprec = 1; iprec <= nprec; iprec++) {
    for (int c = 0; c < ncrms; c++)
      for (int i = 1; i <= nx; i++) {
        for (int k = 1; k <= nzm; k++) A3(tq, c, i, k) = FLD(S.qp, c, i, k);
        for (int k = 1; k <= nzm; k++) {
          int kc = MIN(nzm, k + 1), kb = MAX(1, k - 1);
          A3(mx, c, i, k) = MAX(MAX(A3(tq, c, i, kb), A3(tq, c, i, kc)), A3(tq, c, i, k));
          A3(mn, c, i, k) = MIN(MIN(A3(tq, c, i, kb), A3(tq, c, i, kc)), A3(tq, c, i, k));
          A3(fz, c, i, k) = A3(tq, c, i, k) * A3(wp, c, i, k);
        }
        for (int k = 1; k <= nzm; k++)
          A3(tq, c, i, k) = A3(tq, c, i, k) - (A3(fz, c, i, k + 1) - A3(fz, c, i, k)) * A2(irhoadz, c, k);
        for (int k = 1; k <= nzm; k++) {
          int kb = MAX(1, k - 1);
          A3(www, c, i, k) = 0.5 * (1. + A3(wp, c, i, k) * A2(irhoadz, c, k)) *
                             (A3(tq, c, i, kb) * A3(wp, c, i, kb) - A3(tq, c, i, k) * A3(wp, c, i, k));
        }
        for (int k = 1; k <= nzm; k++) {
          int kc = MIN(nzm, k + 1), kb = MAX(1, k - 1);
          A3(mx, c, i, k) = MAX(MAX(MAX(A3(tq, c, i, kb), A3(tq, c, i, kc)), A3(tq, c, i, k)), A3(mx, c, i, k));
          A3(mn, c, i, k) = MIN(MIN(MIN(A3(tq, c, i, kb), A3(tq, c, i, kc)), A3(tq, c, i, k)), A3(mn, c, i, k));
          double ra = PR(S.rho, c, k) * PR(S.adz, c, k);
          A3(mx, c, i, k) = ra * (A3(mx, c, i, k) - A3(tq, c, i, k)) / (pn(A3(www, c, i, kc)) + pp(A3(www, c, i, k)) + eps);
          A3(mn, c, i, k) = ra * (A3(tq, c, i, k) - A3(mn, c, i, k)) / (pp(A3(www, c, i, kc)) + pn(A3(www, c, i, k)) + eps);
        }
        for (int k = 1; k <= nzm; k++) {
          int kb = MAX(1, k - 1);
          A3(fz, c, i, k) = A3(fz, c, i, k) + pp(A3(www, c, i, k)) * MIN(MIN(1.0, A3(mx, c, i, k)), A3(mn, c, i, kb)) -
                            pn(A3(www, c, i, k)) * MIN(MIN(1.0, A3(mx, c, i, kb)), A3(mn, c, i, k));
        }
      }
    for (int c = 0; c < ncrms; c++)
      for (int i = 1; i <= nx; i++)
        for (int k = 1; k <= nzm; k++) {
          int kc = k + 1;
          double dfz = A3(fz, c, i, kc) - A3(fz, c, i, k);
          FLD(S.qp, c, i, k) = FLD(S.qp, c, i, k) - dfz * A2(irhoadz, c, k);
          PR(S.qpfall, c, k) += -dfz * A2(irhoadz, c, k) * 1.;
          double lat_heat = -(A3(lfac, c, i, kc) * A3(fz, c, i, kc) - A3(lfac, c, i, k) * A3(fz, c, i, k)) * A2(irhoadz, c, k);
          FLD(S.t, c, i, k) = FLD(S.t, c, i, k) - lat_heat;
          PR(S.tlat, c, k) -= lat_heat;
          PR(S.precflux, c, k) -= A3(fz, c, i, k) * 1.;
          if (k == 1) {
            FLD(S.precsfc, c, i, 1) = FLD(S.precsfc, c, i, 1) - A3(fz, c, i, 1) * 1.;
            FLD(S.precssfc, c, i, 1) = FLD(S.precssfc, c, i, 1) - A3(fz, c, i, 1) * (1. - A3(omega, c, i, 1)) * 1.;
          }
        }
    if (iprec < nprec)
      for (int c = 0; c < ncrms; c++)
        for (int i = 1; i <= nx; i++)
          for (int k = 1; k <= nzm; k++) {
            double tmp = term_vel_qp(&tv, FLD(S.qp, c, i, k), PR(S.rho, c, k), FLD(S.tabs, c, i, k));
            double w_ = A2(rhofac, c, k) * tmp;
            A3(wp, c, i, k) = -w_ * PR(S.rhow, c, k) * S.dtn / S.dz[c] / (double)nprec;
            A3(fz, c, i, nz) = 0.;
            A3(www, c, i, nz) = 0.;
            A3(lfac, c, i, nz) = 0.;
          }
  }
#undef CI
#undef A3
#undef A2
  free(omega);
  free(lfac);
  free(wp);
  free(fz);
  free(www);
  free(mx);
  free(mn);
  free(tq);
  free(rhofac);
  free(irhoadz);
  free(iwmax);
}

/* ---- SGS scalar diffusion: diffuse_scalar.F90:8-72, diffuse_scalar2D.F90:5-141 ------- */
static void diffuse_scalar(fld *fp, double *fdiff, double *flux) {
  int nzm = S.nzm, nx = S.nx;
  double rdx2 = 1. / (S.dx * S.dx);
  size_t ni = nx + 1;
  double *flx = (double *)calloc(ni * (nzm + 1), sizeof(double));
  double *dfdt = (double *)calloc(ni * (nzm + 1), sizeof(double));
  double *f0 = (double *)calloc(ni * (nzm + 1), sizeof(double));
#define FX(i, k) flx[(size_t)(k) * ni + (i)]
#define DF(i, k) dfdt[(size_t)(k) * ni + (i)]
  for (int c = 0; c < S.ncrms; c++) {
    for (int k = 1; k <= nzm; k++)
      for (int i = 1; i <= nx; i++) {
        DF(i, k) = 0.;
        f0[(size_t)k * ni + i] = FLD(*fp, c, i, k);
      }
    for (int k = 1; k <= nzm; k++) {
      double rdx5 = 0.5 * rdx2 * PR(S.grdf_x, c, k);
      for (int i = 0; i <= nx; i++) {
        int ic = i + 1;
        double tkx = rdx5 * (FLD(S.tkh, c, i, k) + FLD(S.tkh, c, ic, k));
        FX(i, k) = -tkx * (FLD(*fp, c, ic, k) - FLD(*fp, c, i, k));
      }
      for (int i = 1; i <= nx; i++) DF(i, k) = DF(i, k) - (FX(i, k) - FX(i - 1, k));
    }
    for (int k = 1; k <= nzm; k++) flux[k - 1] = 0.;
    for (int k = 1; k <= nzm; k++) {
      if (k <= nzm - 1) {
        int kc = k + 1;
        double rhoi = PR(S.rhow, c, kc) / PR(S.adzw, c, kc);
        double rdz2 = 1. / (S.dz[c] * S.dz[c]);
        double rdz5 = 0.5 * rdz2 * PR(S.grdf_z, c, k);
        for (int i = 1; i <= nx; i++) {
          double tkz = rdz5 * (FLD(S.tkh, c, i, k) + FLD(S.tkh, c, i, kc));
          FX(i, k) = -tkz * (FLD(*fp, c, i, kc) - FLD(*fp, c, i, k)) * rhoi;
          flux[kc - 1] += FX(i, k);
        }
      } else {
        double tmp = 1. / PR(S.adzw, c, S.nz);
        double rdz = 1. / S.dz[c];
        for (int i = 1; i <= nx; i++) {
          FX(i, 0) = 0.0 * rdz * PR(S.rhow, c, 1);
          FX(i, nzm) = 0.0 * rdz * tmp * PR(S.rhow, c, S.nz);
          flux[0] += FX(i, 0);
        }
      }
    }
    for (int k = 1; k <= nzm; k++) {
      double rhoi = 1. / (PR(S.adz, c, k) * PR(S.rho, c, k));
      for (int i = 1; i <= nx; i++) {
        DF(i, k) = S.dtn * (DF(i, k) - (FX(i, k) - FX(i, k - 1)) * rhoi);
        FLD(*fp, c, i, k) = FLD(*fp, c, i, k) + DF(i, k);
      }
    }
    for (int k = 1; k <= nzm; k++) {
      fdiff[k - 1] = 0.;
      for (int i = 1; i <= nx; i++) fdiff[k - 1] += FLD(*fp, c, i, k) - f0[(size_t)k * ni + i];
    }
    fdiff += S.nz;
    flux += S.nz;
  }
#undef FX
#undef DF
  free(flx);
  free(dfdt);
  free(f0);
}

/* ---- microphysics: precip_init.F90:8-97, cloud.F90:7-114, precip_proc.F90:7-151,
 *      microphysics.F90:331-357 ------------------------------------------------------------ */
static int precip_init(void) {
  S.gam3 = 3.f;
  S.gamr1 = (float)(3. + B_RAIN);
  S.gamr2 = (float)((5. + B_RAIN) / 2.);
  S.gamr3 = (float)(4. + B_RAIN);
  S.gams1 = (float)(3. + B_SNOW);
  S.gams2 = (float)((5. + B_SNOW) / 2.);
  S.gams3 = (float)(4. + B_SNOW);
  S.gamg1 = (float)(3. + B_GRAU);
  S.gamg2 = (float)((5. + B_GRAU) / 2.);
  S.gamg3 = (float)(4. + B_GRAU);
  S.gam3 = gammafff(S.gam3);
  S.gamr1 = gammafff(S.gamr1);
  S.gamr2 = gammafff(S.gamr2);
  S.gamr3 = gammafff(S.gamr3);
  S.gams1 = gammafff(S.gams1);
  S.gams2 = gammafff(S.gams2);
  S.gams3 = gammafff(S.gams3);
  S.gamg1 = gammafff(S.gamg1);
  S.gamg2 = gammafff(S.gamg2);
  S.gamg3 = gammafff(S.gamg3);
  if (lrintf(S.gam3) != 2) return CRM_ERR_UNSUPPORTED;
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++) {
      double rho = PR(S.rho, c, k), t0 = PR(S.tabs0, c, k), pres = PR(S.pres, c, k);
      double pratio = sqrt(R4(1.29) / rho);
      double rrr1 = 393. / (t0 + 120.) * pow(t0 / 273., 1.5);
      double rrr2 = pow(t0 / 273., R4(1.94)) * (1000. / pres);
      double estw = 100. * esatw(t0);
      double esti = 100. * esati(t0);
      double coef1 = 0.25 * PI_R4 * NZEROS * A_SNOW * S.gams1 * pratio / pow(PI_R4 * RHOS * NZEROS / rho, (3 + B_SNOW) / 4.);
      double coef2 = exp(R4(0.025) * (t0 - R4(273.15)));
      PR(S.accrsi, c, k) = coef1 * coef2 * ESICOEF;
      PR(S.accrsc, c, k) = coef1 * ESCCOEF;
      PR(S.coefice, c, k) = coef2;
      coef1 = (LSUB / (t0 * RV) - 1.) * LSUB / (THERCO * rrr1 * t0);
      coef2 = RV * t0 / (DIFFELQ * rrr2 * esti);
      PR(S.evaps1, c, k) = R4(0.65) * 4. * NZEROS / sqrt(PI_R4 * RHOS * NZEROS) / (coef1 + coef2) / sqrt(rho);
      PR(S.evaps2, c, k) = R4(0.49) * 4. * NZEROS * S.gams2 * sqrt(A_SNOW / (MUELQ * rrr1)) /
                           pow(PI_R4 * RHOS * NZEROS, (5 + B_SNOW) / 8.) / (coef1 + coef2) * pow(rho, (1 + B_SNOW) / 8.) *
                           sqrt(pratio);
      coef1 = 0.25 * PI_R4 * NZEROG * A_GRAU * S.gamg1 * pratio / pow(PI_R4 * RHOG * NZEROG / rho, (3 + B_GRAU) / 4.);
      coef2 = exp(R4(0.025) * (t0 - R4(273.15)));
      PR(S.accrgi, c, k) = coef1 * coef2 * EGICOEF;
      PR(S.accrgc, c, k) = coef1 * EGCCOEF;
      coef1 = (LSUB / (t0 * RV) - 1.) * LSUB / (THERCO * rrr1 * t0);
      coef2 = RV * t0 / (DIFFELQ * rrr2 * esti);
      PR(S.evapg1, c, k) = R4(0.65) * 4. * NZEROG / sqrt(PI_R4 * RHOG * NZEROG) / (coef1 + coef2) / sqrt(rho);
      PR(S.evapg2, c, k) = R4(0.49) * 4. * NZEROG * S.gamg2 * sqrt(A_GRAU / (MUELQ * rrr1)) /
                           pow(PI_R4 * RHOG * NZEROG, (5 + B_GRAU) / 8.) / (coef1 + coef2) * pow(rho, (1 + B_GRAU) / 8.) *
                           sqrt(pratio);
      PR(S.accrrc, c, k) = 0.25 * PI_R4 * NZEROR * A_RAIN * S.gamr1 * pratio / pow(PI_R4 * RHOR * NZEROR / rho, (3 + B_RAIN) / 4.) *
                           ERCCOEF;
      coef1 = (LCOND / (t0 * RV) - 1.) * LCOND / (THERCO * rrr1 * t0);
      coef2 = RV * t0 / (DIFFELQ * rrr2 * estw);
      PR(S.evapr1, c, k) = R4(0.78) * 2. * PI_R4 * NZEROR / sqrt(PI_R4 * RHOR * NZEROR) / (coef1 + coef2) / sqrt(rho);
      PR(S.evapr2, c, k) = R4(0.31) * 2. * PI_R4 * NZEROR * S.gamr2 * R4(0.89) * sqrt(A_RAIN / (MUELQ * rrr1)) /
                           pow(PI_R4 * RHOR * NZEROR, (5 + B_RAIN) / 8.) / (coef1 + coef2) * pow(rho, (1 + B_RAIN) / 8.) *
                           sqrt(pratio);
    }
  return CRM_OK;
}

static void micro_diagnose(void) {
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++)
      for (int i = 1; i <= S.nx; i++) {
        double qn = FLD(S.qn, c, i, k), tabs = FLD(S.tabs, c, i, k), qp = FLD(S.qp, c, i, k);
        FLD(S.qv, c, i, k) = FLD(S.qt, c, i, k) - qn;
        double omn = MAX(0.0, MIN(1.0, (tabs - TBGMIN) * S.a_bg));
        FLD(S.qcl, c, i, k) = qn * omn;
        FLD(S.qci, c, i, k) = qn * (1. - omn);
        double omp = MAX(0.0, MIN(1.0, (tabs - TPRMIN) * S.a_pr));
        FLD(S.qpl, c, i, k) = qp * omp;
        FLD(S.qpi, c, i, k) = qp * (1. - omp);
      }
}

static void cloud(void) {
  double an = 1. / (TBGMAX - TBGMIN), bn = TBGMIN * an;
  double ap = 1. / (TPRMAX - TPRMIN), bp = TPRMIN * ap;
  double fac1 = FAC_COND + (1 + bp) * FAC_FUS, fac2 = FAC_FUS * ap;
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++) {
      double pres = PR(S.pres, c, k);
      for (int i = 1; i <= S.nx; i++) {
        double q = MAX(0.0, FLD(S.qt, c, i, k));
        FLD(S.qt, c, i, k) = q;
        double qp = FLD(S.qp, c, i, k);
        double tabs = FLD(S.t, c, i, k) - PR(S.gamaz, c, k);
        double tabs1 = (tabs + fac1 * qp) / (1. + fac2 * qp);
        double qsatt, om;
        if (tabs1 >= TBGMAX) {
          tabs1 = tabs + FAC_COND * qp;
          qsatt = qsatw(tabs1, pres);
        } else if (tabs1 <= TBGMIN) {
          tabs1 = tabs + FAC_SUB * qp;
          qsatt = qsati(tabs1, pres);
        } else {
          om = an * tabs1 - bn;
          qsatt = om * qsatw(tabs1, pres) + (1. - om) * qsati(tabs1, pres);
        }
        double qn;
        if (q > qsatt) {
          int niter = 0;
          double dtabs = 100., dqsat = 0.;
          while (fabs(dtabs) > R4(0.01) && niter < 10) {
            double lstarn, dlstarn, lstarp, dlstarp;
            if (tabs1 >= TBGMAX) {
              lstarn = FAC_COND;
              dlstarn = 0.;
              qsatt = qsatw(tabs1, pres);
              dqsat = dtqsatw(tabs1, pres);
            } else if (tabs1 <= TBGMIN) {
              lstarn = FAC_SUB;
              dlstarn = 0.;
              qsatt = qsati(tabs1, pres);
              dqsat = dtqsati(tabs1, pres);
            } else {
              om = an * tabs1 - bn;
              lstarn = FAC_COND + (1. - om) * FAC_FUS;
              dlstarn = an * FAC_FUS;
              qsatt = om * qsatw(tabs1, pres) + (1. - om) * qsati(tabs1, pres);
              dqsat = om * dtqsatw(tabs1, pres) + (1. - om) * dtqsati(tabs1, pres);
            }
            if (tabs1 >= TPRMAX) {
              lstarp = FAC_COND;
              dlstarp = 0.;
            } else if (tabs1 <= TPRMIN) {
              lstarp = FAC_SUB;
              dlstarp = 0.;
            } else {
              double omp = ap * tabs1 - bp;
              lstarp = FAC_COND + (1. - omp) * FAC_FUS;
              dlstarp = ap * FAC_FUS;
            }
            double fff = tabs - tabs1 + lstarn * (q - qsatt) + lstarp * qp;
            double dfff = dlstarn * (q - qsatt) + dlstarp * qp - lstarn * dqsat - 1.;
            dtabs = -fff / dfff;
            niter++;
            tabs1 = tabs1 + dtabs;
          }
          qsatt = qsatt + dqsat * dtabs;
          qn = MAX(0.0, q - qsatt);
        } else {
          qn = 0.;
        }
        FLD(S.qn, c, i, k) = qn;
        FLD(S.tabs, c, i, k) = tabs1;
        FLD(S.qp, c, i, k) = MAX(0.0, qp);
      }
    }
}

static void precip_proc(void) {
  double powr1 = (3 + B_RAIN) / 4., powr2 = (5 + B_RAIN) / 8.;
  double pows1 = (3 + B_SNOW) / 4., pows2 = (5 + B_SNOW) / 8.;
  double powg1 = (3 + B_GRAU) / 4., powg2 = (5 + B_GRAU) / 8.;
  double dtn = S.dtn;
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++) {
      PR(S.qpsrc, c, k) = 0.;
      PR(S.qpevp, c, k) = 0.;
    }
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++)
      for (int i = 1; i <= S.nx; i++) {
        double q = FLD(S.qt, c, i, k), qp = FLD(S.qp, c, i, k), qn = FLD(S.qn, c, i, k);
        double tabs = FLD(S.tabs, c, i, k);
        if (qn + qp > 0.) {
          double omn = MAX(0.0, MIN(1.0, (tabs - TBGMIN) * S.a_bg));
          double omp = MAX(0.0, MIN(1.0, (tabs - TPRMIN) * S.a_pr));
          double omg = MAX(0.0, MIN(1.0, (tabs - TGRMIN) * S.a_gr));
          if (qn > 0.) {
            double qcc = qn * omn, qii = qn * (1. - omn);
            double autor = (qcc > QCW0) ? ALPHAELQ : 0.;
            double autos = (qii > QCI0) ? BETAELQ * PR(S.coefice, c, k) : 0.;
            double accrr = 0.;
            if (omp > R4(0.001)) {
              double qrr = qp * omp;
              accrr = PR(S.accrrc, c, k) * pow(qrr, powr1);
            }
            double accrcs = 0., accris = 0.;
            if (omp < R4(0.999) && omg < R4(0.999)) {
              double qss = qp * (1. - omp) * (1. - omg);
              double tmp = pow(qss, pows1);
              accrcs = PR(S.accrsc, c, k) * tmp;
              accris = PR(S.accrsi, c, k) * tmp;
            }
            double accrcg = 0., accrig = 0.;
            if (omp < R4(0.999) && omg > R4(0.001)) {
              double qgg = qp * (1. - omp) * omg;
              double tmp = pow(qgg, powg1);
              accrcg = PR(S.accrgc, c, k) * tmp;
              accrig = PR(S.accrgi, c, k) * tmp;
            }
            qcc = (qcc + dtn * autor * QCW0) / (1. + dtn * (accrr + accrcs + accrcg + autor));
            qii = (qii + dtn * autos * QCI0) / (1. + dtn * (accris + accrig + autos));
            double dq = dtn * (accrr * qcc + autor * (qcc - QCW0) + (accris + accrig) * qii + (accrcs + accrcg) * qcc +
                               autos * (qii - QCI0));
            dq = MIN(dq, qn);
            qp = qp + dq;
            q = q - dq;
            qn = qn - dq;
            PR(S.qpsrc, c, k) += dq;
          } else if (qp > QP_THRESHOLD && qn == 0.) {
            double qsatt = 0.;
            double pres = PR(S.pres, c, k);
            if (omn > R4(0.001)) qsatt = qsatt + omn * qsatw(tabs, pres);
            if (omn < R4(0.999)) qsatt = qsatt + (1. - omn) * qsati(tabs, pres);
            double dq = 0.;
            if (omp > R4(0.001)) {
              double qrr = qp * omp;
              dq = dq + PR(S.evapr1, c, k) * sqrt(qrr) + PR(S.evapr2, c, k) * pow(qrr, powr2);
            }
            if (omp < R4(0.999) && omg < R4(0.999)) {
              double qss = qp * (1. - omp) * (1. - omg);
              dq = dq + PR(S.evaps1, c, k) * sqrt(qss) + PR(S.evaps2, c, k) * pow(qss, pows2);
            }
            if (omp < R4(0.999) && omg > R4(0.001)) {
              double qgg = qp * (1. - omp) * omg;
              dq = dq + PR(S.evapg1, c, k) * sqrt(qgg) + PR(S.evapg2, c, k) * pow(qgg, powg2);
            }
            dq = dq * dtn * (q / qsatt - 1.);
            dq = MAX(-0.5 * qp, dq);
            qp = qp + dq;
            q = q - dq;
            PR(S.qpevp, c, k) += dq;
          } else {
            q = q + qp;
            PR(S.qpevp, c, k) -= qp;
            qp = 0.;
          }
        }
        double dq = qp;
        qp = MAX(0.0, qp);
        q = q + (dq - qp);
        FLD(S.qt, c, i, k) = q;
        FLD(S.qp, c, i, k) = qp;
        FLD(S.qn, c, i, k) = qn;
      }
}

/* diagnose.F90:8-177 (the prognostic-relevant part: tabs and the level means) */
static void diagnose(void) {
  double coef = 1. / (double)(S.nx);
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= S.nzm; k++) {
      double u0 = 0, t0 = 0, tabs0 = 0, q0 = 0, qn0 = 0, qp0 = 0, p0 = 0;
      for (int i = 1; i <= S.nx; i++) {
        double tabs = FLD(S.t, c, i, k) - PR(S.gamaz, c, k) + FAC_COND * (FLD(S.qcl, c, i, k) + FLD(S.qpl, c, i, k)) +
                      FAC_SUB * (FLD(S.qci, c, i, k) + FLD(S.qpi, c, i, k));
        FLD(S.tabs, c, i, k) = tabs;
        u0 += FLD(S.u, c, i, k);
        p0 += FLD(S.p, c, i, k);
        t0 += FLD(S.t, c, i, k);
        tabs0 += tabs;
        q0 += FLD(S.qv, c, i, k) + FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k);
        qn0 += FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k);
        qp0 += FLD(S.qpl, c, i, k) + FLD(S.qpi, c, i, k);
      }
      PR(S.u0, c, k) = u0 * coef;
      PR(S.v0, c, k) = 0.0;
      PR(S.t0, c, k) = t0 * coef;
      PR(S.tabs0, c, k) = tabs0 * coef;
      PR(S.q0, c, k) = q0 * coef;
      PR(S.qn0, c, k) = qn0 * coef;
      PR(S.qp0, c, k) = qp0 * coef;
      PR(S.p0, c, k) = p0 * coef;
      PR(S.qv0, c, k) = PR(S.q0, c, k) - PR(S.qn0, c, k);
    }
}

/* Mean-state acceleration, accelerate_crm.F90 (Jones, Bretherton and Pritchard 2015).
 * Settings of crm_accel_init (accelerate_crm.F90:30-38,45-75): use_crm_accel, crm_accel_factor
 * (1 + factor = 'a'), crm_accel_uv. */
static struct {
  int use;
  double factor;
  int uv;
} ACC = {0, 0.0, 0};

void oracle_accel_init(int use_crm_accel, double crm_accel_factor, int crm_accel_uv) {
  ACC.use = use_crm_accel;
  ACC.factor = crm_accel_factor;
  ACC.uv = crm_accel_uv;
}

/* accelerate_crm.F90:109-366 for one call inside the subcycle loop (after micro_proc, before
 * diagnose): the change of the horizontal means of t and qt (and u, v) since the last
 * diagnose, amplified by crm_accel_factor; qt clipped at zero level by level with the
 * negative excess removed proportionally from the positive values; qv, qcl, qci re-split.
 * Raises *ceaseflag (and changes nothing) when the mean t change exceeds 5 K anywhere. */
static void accelerate_crm(int *ceaseflag) {
  const int ncrms = S.ncrms, nx = S.nx, nzm = S.nzm;
  const double coef = 1.0 / (double)(nx * 1);
  const double ttend_threshold = 5., tmin = 50.;
  const size_t n = (size_t)ncrms * (nzm + 1);
  double *tt = (double *)calloc(n, sizeof(double)), *qq = (double *)calloc(n, sizeof(double));
  double *uu = (double *)calloc(n, sizeof(double));
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      double tb = 0, qtb = 0, ub = 0;
      for (int i = 1; i <= nx; i++) {
        tb += FLD(S.t, c, i, k) * coef;
        qtb += (FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k) + FLD(S.qv, c, i, k)) * coef;
        if (ACC.uv) ub += FLD(S.u, c, i, k) * coef;
      }
      size_t o = (size_t)c * (nzm + 1) + k;
      tt[o] = tb - PR(S.t0, c, k);
      qq[o] = qtb - PR(S.q0, c, k);
      if (ACC.uv) uu[o] = ub - PR(S.u0, c, k);
      if (fabs(tt[o]) > ttend_threshold) *ceaseflag = 1;
    }
  if (*ceaseflag) goto done;
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      size_t o = (size_t)c * (nzm + 1) + k;
      double qpoz = 0., qneg = 0.;
      for (int i = 1; i <= nx; i++) {
        FLD(S.t, c, i, k) = MAX(tmin, FLD(S.t, c, i, k) + ACC.factor * tt[o]);
        if (ACC.uv) FLD(S.u, c, i, k) = FLD(S.u, c, i, k) + ACC.factor * uu[o];
        FLD(S.qt, c, i, k) = FLD(S.qt, c, i, k) + ACC.factor * qq[o];
      }
      for (int i = 1; i <= nx; i++) {
        if (FLD(S.qt, c, i, k) < 0.)
          qneg = qneg + FLD(S.qt, c, i, k);
        else
          qpoz = qpoz + FLD(S.qt, c, i, k);
      }
      for (int i = 1; i <= nx; i++) {
        if (qpoz + qneg <= 0.) {
          FLD(S.qt, c, i, k) = 0.;
          FLD(S.qv, c, i, k) = 0.;
          FLD(S.qcl, c, i, k) = 0.;
          FLD(S.qci, c, i, k) = 0.;
        } else {
          double factor = 1. + qneg / qpoz;
          FLD(S.qt, c, i, k) = MAX(0., FLD(S.qt, c, i, k) * factor);
          if (FLD(S.qt, c, i, k) <= 0.) {
            FLD(S.qv, c, i, k) = 0.;
            FLD(S.qcl, c, i, k) = 0.;
            FLD(S.qci, c, i, k) = 0.;
          } else {
            double qt_res = FLD(S.qt, c, i, k) - FLD(S.qcl, c, i, k) - FLD(S.qci, c, i, k);
            FLD(S.qv, c, i, k) = MAX(0., qt_res);
            if (qt_res < 0.) {
              factor = 1. + qt_res / (FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k));
              FLD(S.qcl, c, i, k) = FLD(S.qcl, c, i, k) * factor;
              FLD(S.qci, c, i, k) = FLD(S.qci, c, i, k) * factor;
            }
          }
        }
      }
    }
done:
  free(tt);
  free(qq);
  free(uu);
}

/* crmsurface.F90:104-131 */
static double z0_est(double z, double bflx, double wnd, double ustar) {
  const double vonk = R4(0.4), am = R4(4.8), bm = R4(19.3), eps = R4(1.e-10);
  double c1 = (double)(3.14159f / 2.f - 3.f * logf(2.f));
  double rlmo = -bflx * vonk / (ustar * ustar * ustar + eps);
  double zeta = MIN(1.0, z * rlmo), psi1;
  if (zeta >= 0.) {
    psi1 = -am * zeta;
  } else {
    double x = sqrt(sqrt(1.0 - bm * zeta));
    psi1 = 2. * log(1.0 + x) + log(1.0 + x * x) - 2. * atan(x) + c1;
  }
  double lnz = MAX(0.0, vonk * wnd / (ustar + eps) + psi1);
  return z * exp(-lnz);
}

/* setperturb.F90:7-86 (setperturb_sgs is a no-op under Smagorinsky) */
static void setperturb(int c, int iseed) {
  double factor_xy = 1. / (double)(S.nx);
  mt_seed(iseed);
  mti = 624;
  for (int k = 1; k <= 5; k++) {
    double scaling = (double)(5 + 1 - k) / 5.0;
    double t02 = 0.0;
    for (int i = 1; i <= S.nx; i++) {
      double r = mt_rand();
      r = 1. - 2. * r;
      FLD(S.t, c, i, k) = FLD(S.t, c, i, k) + 1 * r * scaling;
      t02 = t02 + FLD(S.t, c, i, k) * factor_xy;
    }
    for (int i = 1; i <= S.nx; i++) FLD(S.t, c, i, k) = FLD(S.t, c, i, k) * PR(S.t0, c, k) / t02;
  }
}

/* ---- the driver: crm_module.F90:45-1718 --------------------------------------------------- */
#define IN2(p, c, k) ((p)[(size_t)(c) + (size_t)ld * (size_t)(k)])
#define G4(p, c, i, k) ((p)[(size_t)(c) + (size_t)ld * ((size_t)(i) + (size_t)nx * (size_t)(k))])

/* ---- per-substep statistics and end-of-call outputs ------------------------------------
 * The statistics of crm_module.F90:896-1073 (cloud fractions and tops, water paths, mass
 * fluxes, radiation-column sums, interface mass fluxes) and the outputs of
 * crm_module.F90:1078-1659.  They are separate functions so that oracle_stats_probe can run
 * them on a constructed state. */
typedef struct {
  int ncrms, ld, plev, nx, nzm, nz, nxr, ptop; /* ptop 1-based as in the reference */
  double factor_xy, icrm_run_time;
  const crm_input_view *in;
  crm_rad_view *rad;
  crm_output_view *out;
  double *mui, *mdi; /* interface mass fluxes [c][plev + 2] */
} stats_args;

/* A planted mistake in the statistics, selected by oracle_mutate (0: none).  The tests use it
 * to show that their closed forms notice each kind of slip; no other code sets it. */
static int MUT = 0;
enum {
  MUT_NONE = 0,
  MUT_EUDU,      /* entrainment / detrainment branches swapped                   */
  MUT_SPLIT,     /* 700 / 400 hPa bounds of the low / high cloud layers swapped  */
  MUT_CLD_XY,    /* cloud fraction averaged over columns only, not substeps      */
  MUT_MC_WMIN,   /* cloudy updraft / downdraft threshold wmin instead of 2 wmin  */
  MUT_JTMX,      /* jt / mx taken as the last / first flux level                 */
  MUT_WP_LEVEL,  /* water paths scaled by the pdel of the level below           */
  MUT_CLDTOP,    /* every level below the cloud top counted as a top             */
  MUT_TKEZ,      /* resolved kinetic energy without the factor 1/2               */
  MUT_FLUX_DT,   /* resolved scalar flux converted with dt instead of dtn        */
  MUT_MU_LEVEL,  /* updraft cloud test at the level above the interface          */
  MUT_RAD_QV,    /* radiation vapour sum without the clip at zero                */
  MUT_PREC_SPLIT /* convective / stratiform split at 10 mm/h instead of 10 mm/day */
};
void oracle_mutate(int m) { MUT = m; }

static void stats_step(const stats_args *a) {
  const int ncrms = a->ncrms, ld = a->ld, plev = a->plev, nx = a->nx, nzm = a->nzm, nz = a->nz, nxr = a->nxr;
  const crm_input_view *in = a->in;
  crm_rad_view *rad = a->rad;
  crm_output_view *out = a->out;
  double *mui = a->mui, *mdi = a->mdi;
  const double wmin = 2.;
  const double cwp_threshold = R4(0.001);
  const double p_low = MUT == MUT_SPLIT ? 400. : 700., p_high = MUT == MUT_SPLIT ? 700. : 400.;
  const double wcut = MUT == MUT_MC_WMIN ? wmin : 2 * wmin;
  (void)in;

  for (int c = 0; c < ncrms; c++)
    for (int i = 1; i <= nx; i++) {
      double cwp = 0, cwph = 0, cwpm = 0, cwpl = 0, clt = 0, cmt = 0, cht = 0, ctt = 0;
      int flag_top = 1;
      for (int k = 1; k <= nzm; k++) {
        int l = plev - k; /* 0-based */
        int kk = nz - k;
        double tmp1 = PR(S.rho, c, kk) * PR(S.adz, c, kk) * S.dz[c] * (FLD(S.qcl, c, i, kk) + FLD(S.qci, c, i, kk));
        cwp += tmp1;
        ctt = MAX(1.0, ctt);
        if (cwp > cwp_threshold && flag_top) {
          IN2(out->cldtop, c, l) += 1;
          flag_top = MUT == MUT_CLDTOP;
        }
        if (PR(S.pres, c, kk) >= p_low) {
          cwpl += tmp1;
          clt = MAX(1.0, clt);
        } else if (PR(S.pres, c, kk) < p_high) {
          cwph += tmp1;
          cht = MAX(1.0, cht);
        } else {
          cwpm += tmp1;
          cmt = MAX(1.0, cmt);
        }
        tmp1 = PR(S.rho, c, k) * PR(S.adz, c, k) * S.dz[c];
        double ws = FLD(S.w, c, i, k + 1) + FLD(S.w, c, i, k);
        double rw = PR(S.rho, c, k) * 0.5 * ws;
        if (tmp1 * (FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k)) > cwp_threshold) {
          IN2(out->cld, c, l) += 1.0;
          if (ws > wcut) {
            IN2(out->mcup, c, l) += rw * 1.0;
            IN2(out->mcuup, c, l) += rw * (1.0 - 1.0);
          }
          if (ws < -wcut) {
            IN2(out->mcdn, c, l) += rw * 1.0;
            IN2(out->mcudn, c, l) += rw * (1. - 1.0);
          }
        } else {
          if (ws > 2 * wmin) IN2(out->mcuup, c, l) += rw;
          if (ws < -2 * wmin) IN2(out->mcudn, c, l) += rw;
        }
        IN2(out->gliqwp, c, l) += FLD(S.qcl, c, i, k);
        IN2(out->gicewp, c, l) += FLD(S.qci, c, i, k);
      }
      if (cwp > cwp_threshold) out->cltot[c] += ctt;
      if (cwph > cwp_threshold) out->clhgh[c] += cht;
      if (cwpm > cwp_threshold) out->clmed[c] += cmt;
      if (cwpl > cwp_threshold) out->cllow[c] += clt;
    }
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++)
      for (int i = 1; i <= nx; i++) {
        int ir = (i - 1) / (nx / nxr);
        size_t o = (size_t)c + (size_t)ld * ((size_t)ir + (size_t)nxr * (k - 1));
        rad->temperature[o] += FLD(S.tabs, c, i, k);
        rad->qv[o] += MUT == MUT_RAD_QV ? FLD(S.qv, c, i, k) : MAX(0.0, FLD(S.qv, c, i, k));
        rad->qc[o] += FLD(S.qcl, c, i, k);
        rad->qi[o] += FLD(S.qci, c, i, k);
        if (FLD(S.qcl, c, i, k) + FLD(S.qci, c, i, k) > 0) rad->cld[o] += 1.0;
      }
  for (int c = 0; c < ncrms; c++)
    for (int i = 1; i <= nx; i++)
      for (int k = 1; k <= nzm + 1; k++) {
        int l = plev + 1 - k; /* 0-based index of plev+1-k+1 */
        double wk = FLD(S.w, c, i, k);
        if (wk > 0.) {
          int kx = MUT == MUT_MU_LEVEL ? MIN(k, nzm) : MAX(1, k - 1);
          double qsat = qsatw(FLD(S.tabs, c, i, kx), PR(S.pres, c, kx));
          if (FLD(S.qcl, c, i, kx) + FLD(S.qci, c, i, kx) > MIN(R4(1.e-5), 0.01 * qsat))
            mui[(size_t)c * (plev + 2) + l] += PR(S.rhow, c, k) * wk;
        } else if (wk < 0.) {
          int kx = MIN(k + 1, nzm);
          double qsat = qsatw(FLD(S.tabs, c, i, kx), PR(S.pres, c, kx));
          if (FLD(S.qcl, c, i, kx) + FLD(S.qci, c, i, kx) > MIN(R4(1.e-5), 0.01 * qsat))
            mdi[(size_t)c * (plev + 2) + l] += PR(S.rhow, c, k) * wk;
          else if (FLD(S.qpl, c, i, kx) + FLD(S.qpi, c, i, kx) > R4(1.0e-4))
            mdi[(size_t)c * (plev + 2) + l] += PR(S.rhow, c, k) * wk;
        }
      }
}

static void stats_final(const stats_args *a, crm_state_view *st) {
  const int ncrms = a->ncrms, ld = a->ld, plev = a->plev, nx = a->nx, nzm = a->nzm, nxr = a->nxr, ptop = a->ptop;
  const double factor_xy = a->factor_xy, icrm_run_time = a->icrm_run_time;
  const crm_input_view *in = a->in;
  crm_rad_view *rad = a->rad;
  crm_output_view *out = a->out;
  double *mui = a->mui, *mdi = a->mdi;
  {
    double factor_xyt = factor_xy / (double)S.nstop;
    double tmp1 = ((double)nxr / (double)nx) * 1.0 / (double)S.nstop;
    for (int c = 0; c < ncrms; c++)
      for (int k = 0; k < nzm; k++)
        for (int ir = 0; ir < nxr; ir++) {
          size_t o = (size_t)c + (size_t)ld * ((size_t)ir + (size_t)nxr * k);
          rad->temperature[o] *= tmp1;
          rad->qv[o] *= tmp1;
          rad->qc[o] *= tmp1;
          rad->qi[o] *= tmp1;
          rad->cld[o] *= tmp1;
        }
    double *tln = (double *)calloc(plev, sizeof(double)), *qln = (double *)calloc(plev, sizeof(double));
    double *qccln = (double *)calloc(plev, sizeof(double)), *qiiln = (double *)calloc(plev, sizeof(double));
    for (int c = 0; c < ncrms; c++) {
      for (int k = 0; k < plev; k++) {
        if (k < ptop - 1) {
          tln[k] = IN2(in->tl, c, k);
          qln[k] = IN2(in->ql, c, k);
          qccln[k] = IN2(in->qccl, c, k);
          qiiln[k] = IN2(in->qiil, c, k);
        } else {
          tln[k] = qln[k] = qccln[k] = qiiln[k] = 0.;
        }
      }
      double cp_ = 0., cps = 0.;
      for (int k = 1; k <= nzm; k++)
        for (int i = 1; i <= nx; i++) {
          int l = plev - k;
          double pd = IN2(in->pdel, c, l);
          cp_ += (FLD(S.qpl, c, i, k) + FLD(S.qpi, c, i, k)) * pd;
          cps += FLD(S.qpi, c, i, k) * pd;
          tln[l] += FLD(S.tabs, c, i, k);
          qln[l] += FLD(S.qv, c, i, k);
          qccln[l] += FLD(S.qcl, c, i, k);
          qiiln[l] += FLD(S.qci, c, i, k);
        }
      for (int k = ptop - 1; k < plev; k++) {
        tln[k] *= factor_xy;
        qln[k] *= factor_xy;
        qccln[k] *= factor_xy;
        qiiln[k] *= factor_xy;
      }
      for (int k = 0; k < plev; k++) {
        IN2(out->sltend, c, k) = CP * (tln[k] - IN2(in->tl, c, k)) * icrm_run_time;
        IN2(out->qltend, c, k) = (qln[k] - IN2(in->ql, c, k)) * icrm_run_time;
        IN2(out->qcltend, c, k) = (qccln[k] - IN2(in->qccl, c, k)) * icrm_run_time;
        IN2(out->qiltend, c, k) = (qiiln[k] - IN2(in->qiil, c, k)) * icrm_run_time;
      }
      out->prectend[c] = (cp_ - out->prectend[c]) / GGR * factor_xy * icrm_run_time;
      out->precstend[c] = (cps - out->precstend[c]) / GGR * factor_xy * icrm_run_time;
      for (int k = ptop - 1; k <= ptop; k++) {
        IN2(out->sltend, c, k) = 0.;
        IN2(out->qltend, c, k) = 0.;
        IN2(out->qcltend, c, k) = 0.;
        IN2(out->qiltend, c, k) = 0.;
      }
    }
    free(tln);
    free(qln);
    free(qccln);
    free(qiiln);
    for (int c = 0; st && c < ncrms; c++)
      for (int k = 1; k <= nzm; k++)
        for (int i = 1; i <= nx; i++) {
          G4(st->u_wind, c, i - 1, k - 1) = FLD(S.u, c, i, k);
          G4(st->v_wind, c, i - 1, k - 1) = 0.0;
          G4(st->w_wind, c, i - 1, k - 1) = FLD(S.w, c, i, k);
          G4(st->temperature, c, i - 1, k - 1) = FLD(S.tabs, c, i, k);
          G4(st->qt, c, i - 1, k - 1) = FLD(S.qt, c, i, k);
          G4(st->qp, c, i - 1, k - 1) = FLD(S.qp, c, i, k);
          G4(st->qn, c, i - 1, k - 1) = FLD(S.qn, c, i, k);
          G4(out->tk, c, i - 1, k - 1) = FLD(S.tk, c, i, k);
          G4(out->tkh, c, i - 1, k - 1) = FLD(S.tkh, c, i, k);
          G4(out->qcl, c, i - 1, k - 1) = FLD(S.qcl, c, i, k);
          G4(out->qci, c, i - 1, k - 1) = FLD(S.qci, c, i, k);
          G4(out->qpl, c, i - 1, k - 1) = FLD(S.qpl, c, i, k);
          G4(out->qpi, c, i - 1, k - 1) = FLD(S.qpi, c, i, k);
        }
    for (int c = 0; c < ncrms; c++) {
      out->z0m[c] = S.z0[c];
      out->taux[c] = S.taux0[c] / (double)S.nstop;
      out->tauy[c] = S.tauy0[c] / (double)S.nstop;
    }
    for (int c = 0; c < ncrms; c++)
      for (int k = 1; k <= nzm; k++)
        for (int i = 1; i <= nx; i++) {
          int l = plev - k;
          IN2(out->qc_mean, c, l) += FLD(S.qcl, c, i, k);
          IN2(out->qi_mean, c, l) += FLD(S.qci, c, i, k);
          IN2(out->qr_mean, c, l) += FLD(S.qpl, c, i, k);
          double omg = MAX(0.0, MIN(1.0, (FLD(S.tabs, c, i, k) - TGRMIN) * S.a_gr));
          IN2(out->qg_mean, c, l) += FLD(S.qpi, c, i, k) * omg;
          IN2(out->qs_mean, c, l) += FLD(S.qpi, c, i, k) * (1. - omg);
        }
    for (int c = 0; c < ncrms; c++)
      for (int k = 0; k < plev; k++) {
        IN2(out->cld, c, k) = MIN(1.0, IN2(out->cld, c, k) * (MUT == MUT_CLD_XY ? factor_xy : factor_xyt));
        IN2(out->cldtop, c, k) = MIN(1.0, IN2(out->cldtop, c, k) * factor_xyt);
        const int kp = MUT == MUT_WP_LEVEL ? MIN(k + 1, plev - 1) : k;
        IN2(out->gicewp, c, k) = IN2(out->gicewp, c, k) * IN2(in->pdel, c, kp) * 1000. / GGR * factor_xyt;
        IN2(out->gliqwp, c, k) = IN2(out->gliqwp, c, k) * IN2(in->pdel, c, kp) * 1000. / GGR * factor_xyt;
        IN2(out->mcup, c, k) *= factor_xyt;
        IN2(out->mcdn, c, k) *= factor_xyt;
        IN2(out->mcuup, c, k) *= factor_xyt;
        IN2(out->mcudn, c, k) *= factor_xyt;
        IN2(out->mctot, c, k) = IN2(out->mcup, c, k) + IN2(out->mcdn, c, k) + IN2(out->mcuup, c, k) + IN2(out->mcudn, c, k);
        IN2(out->qc_mean, c, k) *= factor_xy;
        IN2(out->qi_mean, c, k) *= factor_xy;
        IN2(out->qs_mean, c, k) *= factor_xy;
        IN2(out->qg_mean, c, k) *= factor_xy;
        IN2(out->qr_mean, c, k) *= factor_xy;
      }
    for (int c = 0; c < ncrms; c++) {
      double precc = 0, precl = 0, precsc = 0, precsl = 0;
      for (int i = 1; i <= nx; i++) {
        double ps = FLD(S.precsfc, c, i, 1) * S.dz[c] / S.dt / (double)S.nstop;
        double pss = FLD(S.precssfc, c, i, 1) * S.dz[c] / S.dt / (double)S.nstop;
        if (ps > 10. / (MUT == MUT_PREC_SPLIT ? 3600. : 86400.)) {
          precc += ps;
          precsc += pss;
        } else {
          precl += ps;
          precsl += pss;
        }
        out->prec_crm[(size_t)c + (size_t)ld * (i - 1)] = ps / 1000.;
      }
      out->precc[c] = precc * factor_xy / 1000.;
      out->precl[c] = precl * factor_xy / 1000.;
      out->precsc[c] = precsc * factor_xy / 1000.;
      out->precsl[c] = precsl * factor_xy / 1000.;
      out->cltot[c] *= factor_xyt;
      out->clhgh[c] *= factor_xyt;
      out->clmed[c] *= factor_xyt;
      out->cllow[c] *= factor_xyt;
      out->jt_crm[c] = plev * 1.0;
      out->mx_crm[c] = 1.0;
    }
    for (int c = 0; c < ncrms; c++)
      for (int k = 0; k < plev; k++) {
        double *mu = &mui[(size_t)c * (plev + 2)], *md = &mdi[(size_t)c * (plev + 2)];
        double pd = IN2(in->pdel, c, k);
        double mu_ = 0.5 * (mu[k] + mu[k + 1]) * GGR / 100.;
        double md_ = 0.5 * (md[k] + md[k + 1]) * GGR / 100.;
        IN2(out->mu_crm, c, k) = mu_;
        IN2(out->md_crm, c, k) = md_;
        IN2(out->eu_crm, c, k) = 0.;
        if (MUT == MUT_EUDU ? mu[k] - mu[k + 1] < 0 : mu[k] - mu[k + 1] > 0)
          IN2(out->eu_crm, c, k) = (mu[k] - mu[k + 1]) * GGR / pd;
        else
          IN2(out->du_crm, c, k) = -1.0 * (mu[k] - mu[k + 1]) * GGR / pd;
        if (md[k + 1] - md[k] < 0) IN2(out->ed_crm, c, k) = (md[k] - md[k + 1]) * GGR / pd;
        if (fabs(mu_) > R4(1.0e-15) || fabs(md_) > R4(1.0e-15)) {
          double kk = k + 1;
          if (MUT == MUT_JTMX) {
            out->jt_crm[c] = MAX(kk, out->jt_crm[c]);
            out->mx_crm[c] = MIN(kk, out->mx_crm[c]);
          } else {
            out->jt_crm[c] = MIN(kk, out->jt_crm[c]);
            out->mx_crm[c] = MAX(kk, out->mx_crm[c]);
          }
        }
      }
    for (int c = 0; c < ncrms; c++)
      for (int k = 1; k <= nzm; k++) {
        double u2z = 0, w2z = 0, tke = 0, tkz = 0;
        for (int i = 1; i <= nx; i++) {
          double du = FLD(S.u, c, i, k) - PR(S.u0, c, k);
          u2z += du * du;
          w2z += 0.5 * (FLD(S.w, c, i, k + 1) * FLD(S.w, c, i, k + 1) + FLD(S.w, c, i, k) * FLD(S.w, c, i, k));
          tke += FLD(S.tke, c, i, k);
          tkz += FLD(S.tk, c, i, k);
        }
        double t1 = S.dz[c] / PR(S.rhow, c, k);
        double t2 = t1 / (MUT == MUT_FLUX_DT ? S.dt : S.dtn);
        int l = plev - k;
        double mkwsb1 = PR(S.mkwsb[0], c, k) * t1 * PR(S.rhow, c, k) * factor_xy / S.nstop;
        double mkwsb2 = PR(S.mkwsb[1], c, k) * t1 * PR(S.rhow, c, k) * factor_xy / S.nstop;
        double mkwle1 = PR(S.mkwle[0], c, k) * t2 * PR(S.rhow, c, k) * factor_xy / S.nstop;
        double mkwle2 = PR(S.mkwle[1], c, k) * t2 * PR(S.rhow, c, k) * factor_xy / S.nstop;
        double mkadv1 = PR(S.mkadv[0], c, k) * factor_xy * icrm_run_time;
        double mkadv2 = PR(S.mkadv[1], c, k) * factor_xy * icrm_run_time;
        double mkdiff1 = PR(S.mkdiff[0], c, k) * factor_xy * icrm_run_time;
        double mkdiff2 = PR(S.mkdiff[1], c, k) * factor_xy * icrm_run_time;
        double qpsrc = PR(S.qpsrc, c, k) * factor_xy * icrm_run_time;
        double qpevp = PR(S.qpevp, c, k) * factor_xy * icrm_run_time;
        double qpfall = PR(S.qpfall, c, k) * factor_xy * icrm_run_time;
        double precflux = PR(S.precflux, c, k) * factor_xy * S.dz[c] / S.dt / S.nstop;
        IN2(out->flux_u, c, l) = (PR(S.uwle, c, k) + PR(S.uwsb, c, k)) * t1 * factor_xy / S.nstop;
        IN2(out->flux_v, c, l) = 0.0 * t1 * factor_xy / S.nstop;
        IN2(out->flux_qt, c, l) = mkwle1 + mkwsb1;
        IN2(out->fluxsgs_qt, c, l) = mkwsb1;
        IN2(out->flux_qp, c, l) = mkwle2 + mkwsb2;
        IN2(out->qt_trans, c, l) = mkadv1 + mkdiff1;
        IN2(out->qp_trans, c, l) = mkadv2 + mkdiff2;
        IN2(out->tkesgsz, c, l) = PR(S.rho, c, k) * tke * factor_xy;
        IN2(out->tkez, c, l) = PR(S.rho, c, k) * (MUT == MUT_TKEZ ? 1.0 : 0.5) * (u2z + 0.0 + w2z) * factor_xy + IN2(out->tkesgsz, c, l);
        IN2(out->tkz, c, l) = tkz * factor_xy;
        IN2(out->precflux, c, l) = precflux / 1000.;
        IN2(out->qp_fall, c, l) = qpfall;
        IN2(out->qp_evp, c, l) = qpevp;
        IN2(out->qp_src, c, l) = qpsrc;
        IN2(out->qt_ls, c, l) = PR(S.qtend, c, k);
        IN2(out->t_ls, c, l) = PR(S.ttend, c, k);
      }
    for (int c = 0; c < ncrms; c++) out->timing_factor[c] = out->timing_factor[c] / S.nstop;
  }
}

int oracle_crm(const crm_dims *d, const crm_column_info *col, double dt_gl, const crm_input_view *in,
               crm_state_view *st, crm_rad_view *rad, crm_output_view *out, int phase_stop) {
  int ncrms = d->ncrms, ld = d->ld, plev = d->plev, nx = d->nx, nzm = d->nz;
  if (d->ny != 1) return CRM_ERR_UNSUPPORTED;
  if (ncrms < 1 || ld < ncrms || nx < 2 || nzm < 3 || plev < nzm + 1) return CRM_ERR_ARG;
  if (nx % d->nx_rad != 0 || d->ny_rad != 1) return CRM_ERR_NXRAD;
  if (60. < 2 * d->dt) return CRM_ERR_DAMPING;
  state_free();
  state_alloc(ncrms, nx, nzm);
  oracle_nprec_max = 1;
  int nz = nzm + 1;
  S.plev = plev;
  S.dx = d->dx;
  S.dy = d->dx;
  S.dt = d->dt;
  S.na = 0;
  S.nb = 1;
  S.nc = 2;
  S.dt3[0] = S.dt3[1] = S.dt3[2] = 0.0;
  S.nstep = 0;
  const double umax = 0.5 * d->dx / d->dt;
  int nxr = d->nx_rad, ptop = plev - nzm + 1; /* 1-based as in the reference */
  double idt_gl = 1. / dt_gl, factor_xy = 1. / (double)(nx);
  int status = CRM_OK;

  for (int c = 0; c < ncrms; c++)
    for (int k = 0; k < nzm; k++)
      for (int ir = 0; ir < nxr; ir++) {
        size_t o = (size_t)c + (size_t)ld * ((size_t)ir + (size_t)nxr * k);
        rad->temperature[o] = rad->qv[o] = rad->qc[o] = rad->qi[o] = rad->cld[o] = 0.0;
      }

  /* reference profiles, crm_module.F90:320-375 */
  for (int c = 0; c < ncrms; c++) {
    for (int k = 1; k <= nzm; k++) {
      PR(S.z, c, k) = IN2(in->zmid, c, plev - k) - IN2(in->zint, c, plev);
      PR(S.zi, c, k) = IN2(in->zint, c, plev - k + 1) - IN2(in->zint, c, plev);
      PR(S.pres, c, k) = IN2(in->pmid, c, plev - k) / 100.;
      PR(S.presi, c, k) = IN2(in->pint, c, plev - k + 1) / 100.;
      PR(S.prespot, c, k) = pow(1000. / PR(S.pres, c, k), RGAS / CP);
      PR(S.bet, c, k) = GGR / IN2(in->tl, c, plev - k);
      PR(S.gamaz, c, k) = GGR / CP * PR(S.z, c, k);
    }
    PR(S.zi, c, nz) = IN2(in->zint, c, plev - nz + 1) - IN2(in->zint, c, plev);
    PR(S.presi, c, nz) = IN2(in->pint, c, plev - nz + 1) / 100.;
    S.dz[c] = 0.5 * (PR(S.z, c, 1) + PR(S.z, c, 2));
    double dz = S.dz[c];
    for (int k = 2; k <= nzm; k++) PR(S.adzw, c, k) = (PR(S.z, c, k) - PR(S.z, c, k - 1)) / dz;
    PR(S.adzw, c, 1) = 1.;
    PR(S.adzw, c, nz) = PR(S.adzw, c, nzm);
    for (int k = 1; k <= nzm; k++) PR(S.adz, c, k) = (PR(S.zi, c, k + 1) - PR(S.zi, c, k)) / dz;
    for (int k = 1; k <= nzm; k++) PR(S.rho, c, k) = IN2(in->pdel, c, plev - k) / GGR / (PR(S.adz, c, k) * dz);
    for (int k = 2; k <= nzm; k++)
      PR(S.rhow, c, k) = (IN2(in->pmid, c, plev - k + 1) - IN2(in->pmid, c, plev - k)) / GGR / (PR(S.adzw, c, k) * dz);
    PR(S.rhow, c, 1) = 2. * PR(S.rhow, c, 2) - PR(S.rhow, c, 3);
    PR(S.rhow, c, nz) = 2. * PR(S.rhow, c, nzm) - PR(S.rhow, c, nzm - 1);
    /* sgs_init (sgs.F90:219-228, LES = .false.) */
    for (int k = 1; k <= nzm; k++) {
      double a = PR(S.adz, c, k) * dz;
      PR(S.grdf_x, c, k) = MIN(16.0, S.dx * S.dx / (a * a));
      PR(S.grdf_z, c, k) = 1.;
    }
  }

  /* initial fields, crm_module.F90:379-452 */
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++)
      for (int i = 1; i <= nx; i++) {
        FLD(S.u, c, i, k) = G4(st->u_wind, c, i - 1, k - 1);
        FLD(S.w, c, i, k) = G4(st->w_wind, c, i - 1, k - 1);
        FLD(S.tabs, c, i, k) = G4(st->temperature, c, i - 1, k - 1);
      }
  {
    /* instance 0 decides (u array halos are still zero here, as in the reference) */
    double u11 = FLD(S.u, 0, 1, 1), u21 = FLD(S.u, 0, 2, 1), u32 = FLD(S.u, 0, 3, 2), u42 = FLD(S.u, 0, 4, 2);
    if (u11 == u21 && u32 == u42)
      for (int c = 0; c < ncrms; c++)
        for (int k = 1; k <= nzm; k++)
          for (int i = 1; i <= nx; i++) FLD(S.u, c, i, k) = MIN(umax, MAX(-umax, FLD(S.u, c, i, k)));
  }
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++)
      for (int i = 1; i <= nx; i++) {
        FLD(S.qt, c, i, k) = G4(st->qt, c, i - 1, k - 1);
        FLD(S.qp, c, i, k) = G4(st->qp, c, i - 1, k - 1);
        FLD(S.qn, c, i, k) = G4(st->qn, c, i - 1, k - 1);
      }
  /* micro_init (microphysics.F90:160-245) */
  S.a_bg = 1. / (TBGMAX - TBGMIN);
  S.a_pr = 1. / (TPRMAX - TPRMIN);
  S.a_gr = 1. / (TGRMAX - TGRMIN);
  micro_diagnose();

  /* domain means and large-scale tendencies, crm_module.F90:482-594 */
  double *colprec = (double *)calloc(ncrms, sizeof(double));
  double *colprecs = (double *)calloc(ncrms, sizeof(double));
  double *t00 = (double *)calloc((size_t)ncrms * nz, sizeof(double));
  for (int c = 0; c < ncrms; c++)
    for (int k = 1; k <= nzm; k++) {
      double u0 = 0, t0 = 0, t00k = 0, tabs0 = 0, q0 = 0, qv0 = 0, qn0 = 0, qp0 = 0;
      for (int i = 1; i <= nx; i++) {
        double qcl = FLD(S.qcl, c, i, k), qci = FLD(S.qci, c, i, k), qpl = FLD(S.qpl, c, i, k), qpi = FLD(S.qpi, c, i, k);
        double t = FLD(S.tabs, c, i, k) + PR(S.gamaz, c, k) - FAC_COND * qcl - FAC_SUB * qci - FAC_COND * qpl - FAC_SUB * qpi;
        FLD(S.t, c, i, k) = t;
        double pd = IN2(in->pdel, c, plev - k);
        colprec[c] += (qpl + qpi) * pd;
        colprecs[c] += qpi * pd;
        u0 += FLD(S.u, c, i, k);
        t0 += t;
        t00k += t + FAC_COND * qpl + FAC_SUB * qpi;
        tabs0 += FLD(S.tabs, c, i, k);
        q0 += FLD(S.qv, c, i, k) + qcl + qci;
        qv0 += FLD(S.qv, c, i, k);
        qn0 += qcl + qci;
        qp0 += qpl + qpi;
      }
      PR(S.u0, c, k) = u0 * factor_xy;
      PR(S.v0, c, k) = 0.0;
      PR(S.t0, c, k) = t0 * factor_xy;
      t00[(size_t)c * nz + k - 1] = t00k * factor_xy;
      PR(S.tabs0, c, k) = tabs0 * factor_xy;
      PR(S.q0, c, k) = q0 * factor_xy;
      PR(S.qv0, c, k) = qv0 * factor_xy;
      PR(S.qn0, c, k) = qn0 * factor_xy;
      PR(S.qp0, c, k) = qp0 * factor_xy;
      PR(S.tke0, c, k) = 0.0;
      int l = plev - k; /* 0-based GCM level */
      double uln = MIN(umax, MAX(-umax, IN2(in->ul, c, l)));
      PR(S.ttend, c, k) = (IN2(in->tl, c, l) + PR(S.gamaz, c, k) - FAC_COND * (IN2(in->qccl, c, l) + IN2(in->qiil, c, l)) -
                           FAC_FUS * IN2(in->qiil, c, l) - t00[(size_t)c * nz + k - 1]) * idt_gl;
      PR(S.qtend, c, k) = (IN2(in->ql, c, l) + IN2(in->qccl, c, l) + IN2(in->qiil, c, l) - PR(S.q0, c, k)) * idt_gl;
      PR(S.utend, c, k) = (uln - PR(S.u0, c, k)) * idt_gl;
      PR(S.vtend, c, k) = 0.0;
      PR(S.ug0, c, k) = uln;
      PR(S.tg0, c, k) = IN2(in->tl, c, l) + PR(S.gamaz, c, k) - FAC_COND * IN2(in->qccl, c, l) - FAC_SUB * IN2(in->qiil, c, l);
      PR(S.qg0, c, k) = IN2(in->ql, c, l) + IN2(in->qccl, c, l) + IN2(in->qiil, c, l);
    }
  for (int c = 0; c < ncrms; c++) {
    double ustar = sqrt(in->tau00[c] / PR(S.rho, c, 1));
    double z0 = z0_est(PR(S.z, c, 1), in->bflxls[c], in->wndls[c], ustar);
    S.z0[c] = MAX(R4(0.00001), MIN(1.0, z0));
    out->timing_factor[c] = 0.;
    out->prectend[c] = colprec[c];
    out->precstend[c] = colprecs[c];
    out->jt_crm[c] = 0.;
    out->mx_crm[c] = 0.;
  }
  double *plev_outs[] = {out->cld, out->cldtop, out->gicewp, out->gliqwp, out->mctot, out->mcup, out->mcdn,
                         out->mcuup, out->mcudn, out->qc_mean, out->qi_mean, out->qs_mean, out->qg_mean,
                         out->qr_mean, out->mu_crm, out->md_crm, out->eu_crm, out->du_crm, out->ed_crm,
                         out->flux_qt, out->flux_u, out->flux_v, out->fluxsgs_qt, out->tkez, out->tkesgsz,
                         out->tkz, out->flux_qp, out->precflux, out->qt_trans, out->qp_trans, out->qp_fall,
                         out->qp_evp, out->qp_src, out->qt_ls, out->t_ls};
  for (size_t n = 0; n < sizeof(plev_outs) / sizeof(plev_outs[0]); n++)
    for (int k = 0; k < plev; k++)
      for (int c = 0; c < ncrms; c++) IN2(plev_outs[n], c, k) = 0.;
  double *mui = (double *)calloc((size_t)ncrms * (plev + 2), sizeof(double));
  double *mdi = (double *)calloc((size_t)ncrms * (plev + 2), sizeof(double));
  const stats_args sa = {ncrms, ld, plev, nx, nzm, nz, nxr, ptop, factor_xy, 1. / dt_gl, in, rad, out, mui, mdi};

  if (precip_init() != CRM_OK) {
    status = CRM_ERR_UNSUPPORTED;
    goto cleanup;
  }
  for (int c = 0; c < ncrms; c++)
    if (col->igstep <= 1) setperturb(c, col->gcol[c] * 1000);

  S.nstop = (int)(dt_gl / S.dt);
  S.dt = dt_gl / S.nstop;
  /* crm_accel_nstop (accelerate_crm.F90:78-106), crm_module.F90:718-720 */
  int accel_cease = 0;
  if (ACC.use) {
    if (S.nstop % (int)(1 + ACC.factor) != 0) {
      status = CRM_ERR_ACCEL;
      goto cleanup;
    }
    S.nstop = (int)(S.nstop / (1 + ACC.factor));
  }

  /* main loop, crm_module.F90:727-1075 */
  S.nstep = 0;
  while (S.nstep < S.nstop) {
    S.nstep++;
    for (int c = 0; c < ncrms; c++) out->timing_factor[c] += 1;
    if (kurant() != CRM_OK) {
      status = CRM_ERR_CFL;
      goto cleanup;
    }
    int last = (S.nstep == S.nstop);
#define STOP_AT(id)                         \
  if (last && phase_stop == (id)) {         \
    status = CRM_OK;                        \
    goto cleanup_keep;                      \
  }
    for (int icyc = 1; icyc <= S.ncycle; icyc++) {
      S.icycle = icyc;
      S.dtn = S.dt / S.ncycle;
      S.dt3[S.na] = S.dtn;
      S.dtfactor = S.dtn / S.dt;
      abcoefs();
      zero();
      STOP_AT(1);
      buoyancy();
      STOP_AT(2);
      forcing();
      STOP_AT(3);
      for (int c = 0; c < ncrms; c++)
        for (int k = 1; k <= nzm; k++)
          for (int i = 1; i <= nx; i++) {
            int ir = (i - 1) / (nx / nxr);
            FLD(S.t, c, i, k) = FLD(S.t, c, i, k) + rad->qrad[(size_t)c + (size_t)ld * ((size_t)ir + (size_t)nxr * (k - 1))] * S.dtn;
          }
      STOP_AT(4);
      damping();
      STOP_AT(5);
      ice_fall();
      STOP_AT(6);
      boundaries(3);
      boundaries(0);
      STOP_AT(7);
      tke_full();
      STOP_AT(8);
      boundaries(4);
      STOP_AT(9);
      advect_mom();
      STOP_AT(10);
      diffuse_mom();
      STOP_AT(11);
      pressure();
      STOP_AT(12);
      adams();
      STOP_AT(13);
      boundaries(2);
      STOP_AT(14);
      {
        double *dummy = (double *)calloc((size_t)ncrms * nz, sizeof(double));
        double *dummy2 = (double *)calloc((size_t)ncrms * nz, sizeof(double));
        advect_scalar(&S.t, dummy, dummy2);
        free(dummy);
        free(dummy2);
      }
      advect_scalar(&S.qt, S.mkadv[0], S.mkwle[0]);
      advect_scalar(&S.qp, S.mkadv[1], S.mkwle[1]);
      precip_fall();
      STOP_AT(15);
      uvw();
      STOP_AT(16);
      boundaries(3);
      STOP_AT(17);
      {
        double *dummy = (double *)calloc((size_t)ncrms * nz, sizeof(double));
        double *dummy2 = (double *)calloc((size_t)ncrms * nz, sizeof(double));
        diffuse_scalar(&S.t, dummy, dummy2);
        free(dummy);
        free(dummy2);
      }
      diffuse_scalar(&S.qt, S.mkdiff[0], S.mkwsb[0]);
      diffuse_scalar(&S.qp, S.mkdiff[1], S.mkwsb[1]);
      STOP_AT(18);
      if (S.icycle == 1 && precip_init() != CRM_OK) {
        status = CRM_ERR_UNSUPPORTED;
        goto cleanup;
      }
      cloud();
      precip_proc();
      micro_diagnose();
      STOP_AT(19);
      /* crm_module.F90:872-876 */
      if (ACC.use && !accel_cease) {
        accelerate_crm(&accel_cease);
        if (accel_cease) S.nstop = (int)(S.nstop + (S.nstop - S.nstep + 1) * ACC.factor);
      }
      diagnose();
      STOP_AT(20);
      int nn = S.na;
      S.na = S.nc;
      S.nc = S.nb;
      S.nb = nn;
    }
#undef STOP_AT

    /* per-step statistics, crm_module.F90:896-1073 */
    stats_step(&sa);
  }

  /* end-of-call outputs, crm_module.F90:1078-1659 */
  stats_final(&sa, st);
cleanup_keep:
  free(colprec);
  free(colprecs);
  free(t00);
  free(mui);
  free(mdi);
  return status;
cleanup:
  free(colprec);
  free(colprecs);
  free(t00);
  free(mui);
  free(mdi);
  return status;
}

/* Copy an internal field (interior points) in the layout of ref_get: [c][k][i] grid fields,
 * [c][k] profiles.  Returns the number of values or -1. */
int oracle_get(const char *name, double *buf) {
  if (!S.live) return -1;
  size_t m = 0;
  int nx = S.nx;
#define GRID(f, nk)                                                       \
  do {                                                                    \
    for (int c = 0; c < S.ncrms; c++)                                     \
      for (int k = 1; k <= (nk); k++)                                     \
        for (int i = 1; i <= nx; i++) buf[m++] = FLD(f, c, i, k);         \
  } while (0)
#define PROF(a, nk)                                                       \
  do {                                                                    \
    for (int c = 0; c < S.ncrms; c++)                                     \
      for (int k = 1; k <= (nk); k++) buf[m++] = PR(a, c, k);             \
  } while (0)
  if (!strcmp(name, "nprec_max")) buf[m++] = oracle_nprec_max;
  else if (!strcmp(name, "u")) GRID(S.u, S.nzm);
  else if (!strcmp(name, "w")) GRID(S.w, S.nz);
  else if (!strcmp(name, "t")) GRID(S.t, S.nzm);
  else if (!strcmp(name, "tabs")) GRID(S.tabs, S.nzm);
  else if (!strcmp(name, "qv")) GRID(S.qv, S.nzm);
  else if (!strcmp(name, "qcl")) GRID(S.qcl, S.nzm);
  else if (!strcmp(name, "qci")) GRID(S.qci, S.nzm);
  else if (!strcmp(name, "qpl")) GRID(S.qpl, S.nzm);
  else if (!strcmp(name, "qpi")) GRID(S.qpi, S.nzm);
  else if (!strcmp(name, "p")) GRID(S.p, S.nzm);
  else if (!strcmp(name, "qt")) GRID(S.qt, S.nzm);
  else if (!strcmp(name, "qp")) GRID(S.qp, S.nzm);
  else if (!strcmp(name, "qn")) GRID(S.qn, S.nzm);
  else if (!strcmp(name, "tke")) GRID(S.tke, S.nzm);
  else if (!strcmp(name, "tk")) GRID(S.tk, S.nzm);
  else if (!strcmp(name, "tkh")) GRID(S.tkh, S.nzm);
  else if (!strcmp(name, "misc")) GRID(S.misc, S.nz);
  else if (!strcmp(name, "dudt_na")) GRID(S.dudt[S.na], S.nzm);
  else if (!strcmp(name, "dudt_nb")) GRID(S.dudt[S.nb], S.nzm);
  else if (!strcmp(name, "dudt_nc")) GRID(S.dudt[S.nc], S.nzm);
  else if (!strcmp(name, "dwdt_na")) GRID(S.dwdt[S.na], S.nz);
  else if (!strcmp(name, "dwdt_nb")) GRID(S.dwdt[S.nb], S.nz);
  else if (!strcmp(name, "dwdt_nc")) GRID(S.dwdt[S.nc], S.nz);
  else if (!strcmp(name, "precsfc")) GRID(S.precsfc, 1);
  else if (!strcmp(name, "precssfc")) GRID(S.precssfc, 1);
  else if (!strcmp(name, "t0")) PROF(S.t0, S.nzm);
  else if (!strcmp(name, "q0")) PROF(S.q0, S.nzm);
  else if (!strcmp(name, "qv0")) PROF(S.qv0, S.nzm);
  else if (!strcmp(name, "tabs0")) PROF(S.tabs0, S.nzm);
  else if (!strcmp(name, "u0")) PROF(S.u0, S.nzm);
  else if (!strcmp(name, "qn0")) PROF(S.qn0, S.nzm);
  else if (!strcmp(name, "qp0")) PROF(S.qp0, S.nzm);
  else if (!strcmp(name, "p0")) PROF(S.p0, S.nzm);
  else if (!strcmp(name, "rho")) PROF(S.rho, S.nzm);
  else if (!strcmp(name, "rhow")) PROF(S.rhow, S.nz);
  else if (!strcmp(name, "adz")) PROF(S.adz, S.nzm);
  else if (!strcmp(name, "adzw")) PROF(S.adzw, S.nz);
  else if (!strcmp(name, "pres")) PROF(S.pres, S.nzm);
  else if (!strcmp(name, "presi")) PROF(S.presi, S.nz);
  else if (!strcmp(name, "z")) PROF(S.z, S.nz);
  else if (!strcmp(name, "zi")) PROF(S.zi, S.nz);
  else if (!strcmp(name, "bet")) PROF(S.bet, S.nzm);
  else if (!strcmp(name, "gamaz")) PROF(S.gamaz, S.nzm);
  else if (!strcmp(name, "ttend")) PROF(S.ttend, S.nzm);
  else if (!strcmp(name, "qtend")) PROF(S.qtend, S.nzm);
  else if (!strcmp(name, "utend")) PROF(S.utend, S.nzm);
  else if (!strcmp(name, "uwle")) PROF(S.uwle, S.nz);
  else if (!strcmp(name, "uwsb")) PROF(S.uwsb, S.nz);
  else if (!strcmp(name, "mkwle1")) PROF(S.mkwle[0], S.nz);
  else if (!strcmp(name, "mkwsb1")) PROF(S.mkwsb[0], S.nz);
  else if (!strcmp(name, "mkadv1")) PROF(S.mkadv[0], S.nz);
  else if (!strcmp(name, "mkdiff1")) PROF(S.mkdiff[0], S.nz);
  else if (!strcmp(name, "qpsrc")) PROF(S.qpsrc, S.nz);
  else if (!strcmp(name, "qpevp")) PROF(S.qpevp, S.nz);
  else if (!strcmp(name, "qpfall")) PROF(S.qpfall, S.nz);
  else if (!strcmp(name, "precflux")) PROF(S.precflux, S.nz);
  else if (!strcmp(name, "dz")) {
    for (int c = 0; c < S.ncrms; c++) buf[m++] = S.dz[c];
  } else
    return -1;
#undef GRID
#undef PROF
  return (int)m;
}

void oracle_release(void) { state_free(); }

/* Overwrite an internal field of the live state, in the layout of oracle_get: grid fields
 * [c][k][i] (interior points), profiles [c][k].  Covers the fields the statistics read.
 * Returns the number of values read from buf, or -1 for an unknown name or no live state. */
int oracle_set(const char *name, const double *buf) {
  if (!S.live) return -1;
  size_t m = 0;
  int nx = S.nx;
  fld *f = NULL;
  double *p = NULL;
  int nk = S.nzm;
  if (!strcmp(name, "u")) f = &S.u;
  else if (!strcmp(name, "w")) { f = &S.w; nk = S.nz; }
  else if (!strcmp(name, "tabs")) f = &S.tabs;
  else if (!strcmp(name, "qv")) f = &S.qv;
  else if (!strcmp(name, "qcl")) f = &S.qcl;
  else if (!strcmp(name, "qci")) f = &S.qci;
  else if (!strcmp(name, "qpl")) f = &S.qpl;
  else if (!strcmp(name, "qpi")) f = &S.qpi;
  else if (!strcmp(name, "tke")) f = &S.tke;
  else if (!strcmp(name, "tk")) f = &S.tk;
  else if (!strcmp(name, "precsfc")) { f = &S.precsfc; nk = 1; }
  else if (!strcmp(name, "precssfc")) { f = &S.precssfc; nk = 1; }
  else if (!strcmp(name, "u0")) p = S.u0;
  else if (!strcmp(name, "uwle")) { p = S.uwle; nk = S.nz; }
  else if (!strcmp(name, "uwsb")) { p = S.uwsb; nk = S.nz; }
  else if (!strcmp(name, "mkwle1")) { p = S.mkwle[0]; nk = S.nz; }
  else if (!strcmp(name, "mkwle2")) { p = S.mkwle[1]; nk = S.nz; }
  else if (!strcmp(name, "mkwsb1")) { p = S.mkwsb[0]; nk = S.nz; }
  else if (!strcmp(name, "mkwsb2")) { p = S.mkwsb[1]; nk = S.nz; }
  else if (!strcmp(name, "mkadv1")) { p = S.mkadv[0]; nk = S.nz; }
  else if (!strcmp(name, "mkadv2")) { p = S.mkadv[1]; nk = S.nz; }
  else if (!strcmp(name, "mkdiff1")) { p = S.mkdiff[0]; nk = S.nz; }
  else if (!strcmp(name, "mkdiff2")) { p = S.mkdiff[1]; nk = S.nz; }
  else if (!strcmp(name, "qpsrc")) { p = S.qpsrc; nk = S.nz; }
  else if (!strcmp(name, "qpevp")) { p = S.qpevp; nk = S.nz; }
  else if (!strcmp(name, "qpfall")) { p = S.qpfall; nk = S.nz; }
  else if (!strcmp(name, "precflux")) { p = S.precflux; nk = S.nz; }
  else
    return -1;
  for (int c = 0; c < S.ncrms; c++)
    for (int k = 1; k <= nk; k++) {
      if (f)
        for (int i = 1; i <= nx; i++) FLD(*f, c, i, k) = buf[m++];
      else
        PR(p, c, k) = buf[m++];
    }
  return (int)m;
}

/* Run the per-substep statistics (crm_module.F90:896-1073) nsteps times on the live state,
 * as a previous oracle_crm call set it up (profiles from `in`) and oracle_set adjusted it,
 * then the end-of-call outputs (crm_module.F90:1078-1659) into rad and out; the state views
 * are not written.  The substep length is dt_gl / nsteps and dtn is the length of the last
 * subcycle.  As in a call, the plev-level outputs and the radiation sums start from zero, and
 * cltot..cllow and prectend/precstend start from the values the caller passes. */
int oracle_stats_probe(const crm_dims *d, double dt_gl, double dtn, int nsteps, const crm_input_view *in,
                       crm_rad_view *rad, crm_output_view *out) {
  if (!S.live || !d || d->ncrms != S.ncrms || d->nx != S.nx || d->nz != S.nzm || nsteps < 1 ||
      d->ny_rad != 1 || d->nx_rad < 1 || d->nx % d->nx_rad)
    return CRM_ERR_ARG;
  const int ncrms = d->ncrms, ld = d->ld, plev = d->plev, nzm = d->nz, nxr = d->nx_rad;
  for (int c = 0; c < ncrms; c++)
    for (int k = 0; k < nzm; k++)
      for (int ir = 0; ir < nxr; ir++) {
        size_t o = (size_t)c + (size_t)ld * ((size_t)ir + (size_t)nxr * k);
        rad->temperature[o] = rad->qv[o] = rad->qc[o] = rad->qi[o] = rad->cld[o] = 0.0;
      }
  double *plev_outs[] = {out->cld, out->cldtop, out->gicewp, out->gliqwp, out->mctot, out->mcup, out->mcdn,
                         out->mcuup, out->mcudn, out->qc_mean, out->qi_mean, out->qs_mean, out->qg_mean,
                         out->qr_mean, out->mu_crm, out->md_crm, out->eu_crm, out->du_crm, out->ed_crm,
                         out->flux_qt, out->flux_u, out->flux_v, out->fluxsgs_qt, out->tkez, out->tkesgsz,
                         out->tkz, out->flux_qp, out->precflux, out->qt_trans, out->qp_trans, out->qp_fall,
                         out->qp_evp, out->qp_src, out->qt_ls, out->t_ls};
  for (size_t n = 0; n < sizeof(plev_outs) / sizeof(plev_outs[0]); n++)
    for (int k = 0; k < plev; k++)
      for (int c = 0; c < ncrms; c++) IN2(plev_outs[n], c, k) = 0.;
  double *mui = (double *)calloc((size_t)ncrms * (plev + 2), sizeof(double));
  double *mdi = (double *)calloc((size_t)ncrms * (plev + 2), sizeof(double));
  S.nstop = nsteps;
  S.dt = dt_gl / nsteps;
  S.dtn = dtn;
  const stats_args sa = {ncrms, ld, plev, d->nx, nzm, nzm + 1, nxr, plev - nzm + 1, 1. / (double)d->nx,
                         1. / dt_gl, in, rad, out, mui, mdi};
  for (int n = 0; n < nsteps; n++) stats_step(&sa);
  stats_final(&sa, NULL);
  free(mui);
  free(mdi);
  return CRM_OK;
}
