// crm_common.hpp -- constants, device state layout and wave-segment helpers of the HIP CRM.
//
// Layout in HBM (instance-major structure of arrays, no halos):
//   scalar fields   F[(c*nzm + k)*NX + i]   k = 0..nzm-1 (CRM level k+1 of the reference)
//   w-level fields  W[(c*nz  + k)*NX + i]   k = 0..nz-1  (w level k+1; k = nzm is the lid)
//   profiles        P[c*nz + k]
// One wave64 holds 64/NX instances; lane i of a segment owns column i of its instance, so
// the reference's periodic x-halos (crm/periodic.F90, crm/bound_exchange.F90) become lane
// rotations inside the segment and every domain mean is a butterfly over NX lanes (fixed
// order, so results are bit-reproducible run to run).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crm {

// Single-precision value of a default-kind (REAL(4)) literal of the reference, widened.
#define R4(x) ((double)(x##f))

// crm/params.F90:49-69 with cime/src/share/util/shr_const_mod.F90:22-55 under -DCRM
constexpr double SHR_RGAS = 6.02214e26 * 1.38065e-23;
constexpr double CP = 1.00464e3;
constexpr double GGR = 9.80616;
constexpr double LCOND = 2.501e6;
constexpr double LFUS = 3.337e5;
constexpr double LSUB = LCOND + LFUS;
constexpr double RGAS = SHR_RGAS / 28.966;
constexpr double RV = SHR_RGAS / 18.016;
constexpr double FAC_COND = LCOND / CP;
constexpr double FAC_FUS = LFUS / CP;
constexpr double FAC_SUB = LSUB / CP;
#define PI_R4 R4(3.141592653589793)
#define EPSV R4(0.61)
#define DIFFELQ R4(2.21e-05)
#define THERCO R4(2.40e-02)
#define MUELQ R4(1.717e-05)
// crm/MICRO_SAM1MOM/micro_params.F90:12-61
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
#define ESICOEF R4(0.1)
#define EGICOEF R4(0.1)
#define NZEROR 8.e6
#define NZEROS 3.e6
#define NZEROG 4.e6
#define QP_THRESHOLD R4(1.e-8)

constexpr int NCOEF = 12;  // precip_init coefficients per level
// Largest number of precipitation sub-steps of one substep (the reference has no limit; a
// count this large only arises from a non-physical state, which is reported as an error).
constexpr int NPREC_MAX = 512;
constexpr unsigned long long ERR_FLAG_PRECIP = 1, ERR_FLAG_NAN = 2;
enum {
  C_ACCRSC = 0, C_ACCRSI, C_ACCRRC, C_COEFICE, C_ACCRGC, C_ACCRGI,
  C_EVAPS1, C_EVAPS2, C_EVAPR1, C_EVAPR2, C_EVAPG1, C_EVAPG2
};

// Profile arrays, each [ncrms][nz].
enum Prof {
  P_Z = 0, P_ZI, P_PRES, P_PRESI, P_ADZ, P_ADZW, P_RHO, P_RHOW, P_BET, P_GAMAZ,
  P_GRDFX, P_TAU,                                          // reference state (per call)
  P_T0, P_Q0, P_QV0, P_TABS0, P_U0, P_QN0, P_QP0,          // domain means (diagnose)
  P_TTEND, P_QTEND, P_UTEND,                               // large-scale tendencies
  P_UWLE, P_UWSB, P_MKWLE1, P_MKWLE2, P_MKWSB1, P_MKWSB2,  // last-substep fluxes
  P_MKADV1, P_MKADV2, P_MKDIFF1, P_MKDIFF2, P_QPSRC, P_QPEVP,
  P_QPFALL, P_PRECFLUX,                                    // accumulated over the call
  P_CLD, P_CLDTOP, P_MCUP, P_MCDN, P_MCUUP, P_MCUDN, P_GLIQ, P_GICE, P_MUI, P_MDI,
  NPROF
};

// Per-instance scalars, each [ncrms].
enum Inst {
  I_DZ = 0, I_NDAMP, I_KMIN, I_KMAX, I_Z0, I_COLPREC, I_COLPRECS, I_CLTOT, I_CLHGH, I_CLMED,
  I_CLLOW, NINST
};

// Terminal-velocity coefficients of precipitation (microphysics.F90:499-504).
struct TermVel {
  double vrain, vsnow, vgrau, crain, csnow, cgrau;
};
// REAL(4) Gamma-function values of precip_init.F90:24-45 (computed on the host with the
// reference's own recipe, gammafff.c).
struct Gam {
  double gamr1, gamr2, gamr3, gams1, gams2, gams3, gamg1, gamg2, gamg3;
};

// Everything a kernel needs, passed by value.
struct Dev {
  int ncrms, nzm, nz, plev, ld, nx_rad;
  double dx, dt, dtn, dt3na, at, bt, ct, dtfactor;
  int nstep, icycle, nprec_max_slot;
  // prognostic and diagnostic fields [c][k][i]
  double *u, *w, *t, *qt, *qp, *qn, *tabs, *qv, *qcl, *qci, *qpl, *qpi;
  double *tk, *tkh, *tke;
  double *dudt[3], *dwdt[3];  // AB3 history; na/nb/nc resolved on the host
  double *precsfc, *precssfc; // [c][i]
  double *prof;               // NPROF blocks of [ncrms][nz]
  double *inst;               // NINST blocks of [ncrms]
  double *coef;               // NCOEF blocks of [ncrms][nz]
  double *rad;                // 5 blocks of [ncrms][nz][nx_rad]: t, qv, qc, qi, cld
  const double *Ft, *Gt;      // transform operators [NX][NX]
  const double *eign;         // [NX] horizontal eigenvalues per slot
  unsigned long long *gmax;   // [0] kurant cfl bits, [1] precipitation cfl bits, [2] error flag
  TermVel tv;
  Gam gam;
  double z0c1;                // 3.14159/2 - 3 ln 2 in REAL(4) (crmsurface.F90:125)
};

__device__ __forceinline__ double* prof(const Dev& d, int p, int c) {
  return d.prof + ((size_t)p * d.ncrms + c) * d.nz;
}
__device__ __forceinline__ double& inst(const Dev& d, int p, int c) {
  return d.inst[(size_t)p * d.ncrms + c];
}

// A segment of NX lanes of one wave64 = one CRM instance.
template <int NX>
struct Seg {
  int lane, i, base;
  __device__ __forceinline__ Seg() {
    lane = threadIdx.x & 63;
    i = lane & (NX - 1);
    base = lane - i;
  }
  // value held by column (i + d) mod NX
  __device__ __forceinline__ double rot(double v, int d) const {
    return __shfl(v, base + ((i + d + NX) & (NX - 1)), 64);
  }
  __device__ __forceinline__ int roti(int v, int d) const {
    return __shfl(v, base + ((i + d + NX) & (NX - 1)), 64);
  }
  __device__ __forceinline__ double sum(double v) const {
#pragma unroll
    for (int m = NX / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
  }
  // sum over aligned groups of g lanes (g a power of two dividing NX)
  __device__ __forceinline__ double gsum(double v, int g) const {
#pragma unroll
    for (int m = NX / 2; m >= 1; m >>= 1)
      if (m < g) v += __shfl_xor(v, m, 64);
    return v;
  }
  __device__ __forceinline__ double max(double v) const {
#pragma unroll
    for (int m = NX / 2; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
    return v;
  }
  __device__ __forceinline__ int imax(int v) const {
#pragma unroll
    for (int m = NX / 2; m >= 1; m >>= 1) v = ::max(v, __shfl_xor(v, m, 64));
    return v;
  }
  __device__ __forceinline__ int imin(int v) const {
#pragma unroll
    for (int m = NX / 2; m >= 1; m >>= 1) v = ::min(v, __shfl_xor(v, m, 64));
    return v;
  }
};

// Instance of this thread's segment; false when past ncrms (the segment then idles but still
// joins cross-lane operations of its wave).
template <int NX>
__device__ __forceinline__ bool seg_instance(int ncrms, int& c) {
  int g = (blockIdx.x * blockDim.x + threadIdx.x) / NX;
  c = g < ncrms ? g : ncrms - 1;
  return g < ncrms;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Accumulate into a statistics location that exactly one lane updates per launch: a
// returnless hardware FP64 add, so the sweep does not wait a memory round trip per level.
// Launches are stream-ordered, so the sequence of additions (and the result) is the same as
// a plain read-modify-write.
__device__ __forceinline__ void accum(double* p, double v) { unsafeAtomicAdd(p, v); }

// ---- saturation functions, crm/sat.F90:12-139 (REAL(4) coefficients) ----------------------
__device__ __forceinline__ double poly8(const float* a, double dt) {
  return a[0] + dt * (a[1] + dt * (a[2] + dt * (a[3] + dt * (a[4] + dt * (a[5] + dt * (a[6] + dt * (a[7] + a[8] * dt)))))));
}
__constant__ float c_esatw[9] = {6.105851f, 0.4440316f, 0.1430341e-1f, 0.2641412e-3f, 0.2995057e-5f,
                                 0.2031998e-7f, 0.6936113e-10f, 0.2564861e-13f, -0.3704404e-15f};
__constant__ float c_dtesatw[9] = {0.443956472f, 0.285976452e-1f, 0.794747212e-3f, 0.121167162e-4f,
                                   0.103167413e-6f, 0.385208005e-9f, -0.604119582e-12f, -0.792933209e-14f,
                                   -0.599634321e-17f};
__constant__ float c_esati[9] = {6.11147274f, 0.503160820f, 0.188439774e-1f, 0.420895665e-3f, 0.615021634e-5f,
                                 0.602588177e-7f, 0.385852041e-9f, 0.146898966e-11f, 0.252751365e-14f};
__constant__ float c_dtesati[9] = {0.503223089f, 0.377174432e-1f, 0.126710138e-2f, 0.249065913e-4f,
                                   0.312668753e-6f, 0.255653718e-8f, 0.132073448e-10f, 0.390204672e-13f,
                                   0.497275778e-16f};

__device__ __forceinline__ double esat_low(double t) {
  return exp(R4(9.550426) - R4(5723.265) / t + R4(3.53068) * log(t) - R4(0.00728332) * t);
}
__device__ __forceinline__ double esatw(double t) {
  double dt = t - R4(273.16);
  if (dt > -80.0) return poly8(c_esatw, dt);
  return R4(0.01) * 2.0 * esat_low(t);
}
__device__ __forceinline__ double esati(double t) {
  double dt = t - R4(273.16);
  if (dt > -80.0) return poly8(c_esati, dt);
  return R4(0.01) * esat_low(t);
}
__device__ __forceinline__ double qsatw(double t, double p) {
  double e = esatw(t);
  return R4(0.622) * e / fmax(e, p - e);
}
__device__ __forceinline__ double qsati(double t, double p) {
  double e = esati(t);
  return R4(0.622) * e / fmax(e, p - e);
}
__device__ __forceinline__ double dtqsatw(double t, double p) {
  double dt = t - R4(273.16);
  double de = dt > -80.0 ? poly8(c_dtesatw, dt) : esatw(t + 1) - esatw(t);
  return R4(0.622) * de / p;
}
__device__ __forceinline__ double dtqsati(double t, double p) {
  double dt = t - R4(273.16);
  double de = dt > -80.0 ? poly8(c_dtesati, dt) : esati(t + 1.0) - esati(t);
  return R4(0.622) * de / p;
}

}  // namespace crm
