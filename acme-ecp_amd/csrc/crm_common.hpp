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

#include "crm_pow.hpp"

// The FP64 form is compiled in namespace crm; the mixed-precision form (FP32 bulk fields) is
// compiled from the same sources in its own translation unit and namespace.
#ifndef CRM_NS
#define CRM_NS crm
#endif

namespace CRM_NS {

// Single-precision value of a default-kind (REAL(4)) literal of the reference; widened to
// FP64 in the FP64 translation unit.  The mixed-precision unit keeps it in FP32, which is
// the same value wherever it meets FP64 operands.
#ifdef CRM_MIXED_TU
#define R4(x) (x##f)
#else
#define R4(x) ((double)(x##f))
#endif

// Type of the physical constants: FP64, or FP32 in the mixed-precision unit so that bulk FP32
// arithmetic does not widen.
#ifdef CRM_MIXED_TU
using creal = float;
#else
using creal = double;
#endif

// crm/params.F90:49-69 with cime/src/share/util/shr_const_mod.F90:22-55 under -DCRM
constexpr double SHR_RGAS = 6.02214e26 * 1.38065e-23;
constexpr creal CP = 1.00464e3;
constexpr creal GGR = 9.80616;
constexpr creal LCOND = 2.501e6;
constexpr creal LFUS = 3.337e5;
constexpr creal LSUB = LCOND + LFUS;
constexpr creal RGAS = SHR_RGAS / 28.966;
constexpr creal RV = SHR_RGAS / 18.016;
constexpr creal FAC_COND = LCOND / CP;
constexpr creal FAC_FUS = LFUS / CP;
constexpr creal FAC_SUB = LSUB / CP;
// The same constants in FP64 in every unit (integer ratios are not touched by FP32 literals),
// for the reference-profile set-up, the state transforms at the call boundary and the outputs.
constexpr double CP_D = (double)100464 / 100;
constexpr double GGR_D = (double)980616 / 100000;
constexpr double LCOND_D = (double)2501000;
constexpr double LFUS_D = (double)333700;
constexpr double FAC_COND_D = LCOND_D / CP_D;
constexpr double FAC_FUS_D = LFUS_D / CP_D;
constexpr double FAC_SUB_D = (LCOND_D + LFUS_D) / CP_D;
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
  P_IRHO, P_IADZ, P_IRHOADZ, P_IRHOWADZ, P_DDADZ, P_PRATIO,  // derived, per call
  P_SRDZ, P_SRDX, P_SRDZW, P_SRDXW, P_SBETDZ, P_STKMAX,    // SGS level factors, per call
  P_DRHOI,                                                 // scalar diffusion, per call
  P_T0, P_Q0, P_QV0, P_TABS0, P_U0, P_QN0, P_QP0,          // domain means (diagnose)
  P_V0,                                                    // (3D only)
  P_TTEND, P_QTEND, P_UTEND,                               // large-scale tendencies
  P_VTEND,                                                 // (3D only)
  P_UG0, P_VG0,                                            // large-scale wind (Coriolis, 3D)
  P_UWLE, P_UWSB, P_MKWLE1, P_MKWLE2, P_MKWSB1, P_MKWSB2,  // last-substep fluxes
  P_VWLE, P_VWSB,                                          // (3D only)
  P_MKADV1, P_MKADV2, P_MKDIFF1, P_MKDIFF2, P_QPSRC, P_QPEVP,
  P_QPFALL, P_PRECFLUX,                                    // accumulated over the call
  P_CLD, P_CLDTOP, P_MCUP, P_MCDN, P_MCUUP, P_MCUDN, P_GLIQ, P_GICE, P_MUI, P_MDI,
  P_ACCT, P_ACCQ, P_ACCU, P_ACCV,                          // mean-state acceleration, per subcycle
  NPROF
};

// Per-instance scalars, each [ncrms].
enum Inst {
  I_DZ = 0, I_NDAMP, I_KMIN, I_KMAX, I_Z0, I_COLPREC, I_COLPRECS, I_CLTOT, I_CLHGH, I_CLMED,
  I_CLLOW, I_FCOR, I_FCORZ, NINST
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

// Everything a kernel needs, passed by value.  R is the type of the bulk fields (double, or
// float in the mixed-precision form); level profiles, per-instance scalars, surface
// precipitation and every accumulated statistic stay FP64.
template <typename R>
struct DevT {
  int ncrms, nzm, nz, plev, ld, nx_rad;
  int ny;                     // 1 for 2D; 8 or 4 for the 3D 8 x 8 or 4 x 4 CRM (column nx j + i)
  // radiation column groups (crm_module.F90:979-1015): row length nx, radiation columns
  // nx_rad_x along x, and group sizes nx / nx_rad_x and ny / ny_rad (powers of two)
  int rad_row, rad_nxx, rad_gx, rad_gy;
  double dx, dt, dtn, dt3na, at, bt, ct, dtfactor;
  int nstep, icycle;
  int docoriolis;             // Coriolis terms (params.F90:107; 3D only here)
  // prognostic and diagnostic fields [c][k][i]
  R *u, *w, *t, *qt, *qp, *qn, *tabs, *qv, *qcl, *qci, *qpl, *qpi;
  R *tk, *tkh, *tke;
  R *dudt[3], *dwdt[3];       // AB3 history; na/nb/nc resolved on the host
  R *v, *dvdt[3];             // 3D only (v is identically zero in 2D and not stored)
  double *precsfc, *precssfc; // [c][i]
  double *prof;               // NPROF blocks of [ncrms][nz]
  R *profr;                   // the same profiles in R (== prof when R is double)
  double *inst;               // NINST blocks of [ncrms]
  R *coef;                    // NCOEF blocks of [ncrms][nz]
  R *rad;                     // 5 blocks of [ncrms][nzm][nx_rad]: t, qv, qc, qi, cld (in 3D
                              // nx_rad counts the nx_rad*ny_rad radiation columns); sums in
                              // R: FP64, or FP32 in the mixed engine
  const double *Ft, *Gt;      // transform operators [NX][NX]
  const double *eign;         // [NX] horizontal eigenvalues per slot
  unsigned long long *gmax;   // [0] stop flag of a captured call, [1] largest sedimentation
                              // sub-step count nprec of the call, [2] error flag,
                              // [3] mean-state acceleration cease flag, then the kurant and
                              // precipitation Courant slots
  const unsigned long long *stop;  // &gmax[0] inside a captured call, else null
  TermVel tv;
  Gam gam;
  double z0c1;                // 3.14159/2 - 3 ln 2 in REAL(4) (crmsurface.F90:125)
  // microphysics network: inputs [point][9] and outputs [point][8] of this subcycle (null:
  // plain microphysics everywhere)
  const float *mlp_x, *mlp_y;
  // (input, target) pairs of the plain microphysics, every samp_stride-th point
  float *samp_x, *samp_t;
  unsigned long long *samp_n;
  long long samp_cap;
  int samp_stride;
};

// The kurant maximum and the precipitation Courant maximum are each gathered in KURANT_SLOTS
// words, one 128-byte line apart, after the four flag words of gmax: instance c updates slot
// c % KURANT_SLOTS, so thousands of instances finishing together do not queue on one address;
// the reader takes the maximum of the slots (the host for kurant, k_precip for the other).
// Radiation column of storage column i: i_rad = (i-1)/(nx/crm_nx_rad)+1 and j_rad likewise,
// stored i_rad + nx_rad*j_rad (crm_module.F90:979-1015, crm_rad_module.F90:17-34).
template <typename R>
__device__ __forceinline__ int rad_col(const DevT<R>& d, int i) {
  return (i % d.rad_row) / d.rad_gx + d.rad_nxx * ((i / d.rad_row) / d.rad_gy);
}
// the one lane of each radiation group that owns its sums
template <typename R>
__device__ __forceinline__ bool rad_lead(const DevT<R>& d, int i) {
  return (i % d.rad_row) % d.rad_gx == 0 && (i / d.rad_row) % d.rad_gy == 0;
}

constexpr int GMAX_FLAGS = 4, KURANT_SLOTS = 32, KURANT_STRIDE = 16;
constexpr int GMAX_STOP = 0;  // flag word: the substep that stopped a captured call
constexpr int GMAX_NPREC = 1;  // the call's largest nprec (microphysics.F90:602-662)
constexpr int GMAX_PCFL = GMAX_FLAGS + KURANT_SLOTS * KURANT_STRIDE;  // precipitation slots
constexpr int GMAX_WORDS = GMAX_PCFL + KURANT_SLOTS * KURANT_STRIDE;

template <typename R>
__device__ __forceinline__ double* prof(const DevT<R>& d, int p, int c) {
  return d.prof + ((size_t)p * d.ncrms + c) * d.nz;
}
template <typename R>
__device__ __forceinline__ R* profr(const DevT<R>& d, int p, int c) {
  return d.profr + ((size_t)p * d.ncrms + c) * d.nz;
}
// A profile that the kernel only reads, as a constant-address-space pointer: with a
// wave-uniform instance its loads become scalar loads, which the hardware counts apart from
// the vector loads, so waiting for a profile value does not also wait for the field loads
// issued ahead of it (vector loads are counted in issue order).
template <typename R>
using cprof = const __attribute__((address_space(4))) R*;
template <typename R>
__device__ __forceinline__ cprof<R> profc(const DevT<R>& d, int p, int c) {
  return (cprof<R>)(d.profr + ((size_t)p * d.ncrms + c) * d.nz);
}
template <typename R>
__device__ __forceinline__ cprof<double> profcd(const DevT<R>& d, int p, int c) {
  return (cprof<double>)(d.prof + ((size_t)p * d.ncrms + c) * d.nz);
}
template <typename R>
__device__ __forceinline__ double& inst(const DevT<R>& d, int p, int c) {
  return d.inst[(size_t)p * d.ncrms + c];
}

// ---- cross-lane data movement inside a wave64 ----------------------------------------------
// DPP (data-parallel primitives) move a register between lanes on the VALU, without the LDS
// crossbar that ds_bpermute (what __shfl compiles to) occupies.  Controls, with the lane a
// destination lane l reads:
//   wave_rol:1  l+1 (mod 64)       wave_ror:1  l-1 (mod 64)
//   row_ror:15  l+1 within its 16-lane row      row_ror:1  l-1 within its row
//   quad_perm [1,0,3,2]  l^1       quad_perm [2,3,0,1]  l^2
//   row_half_mirror  7-l within its 8-lane half-row    row_mirror  15-l within its row
// and gfx950's v_permlane16_swap / v_permlane32_swap give the partners l^16 and l^32.
#ifndef CRM_SEG_BPERMUTE
constexpr int DPP_QP_XOR1 = 0xB1, DPP_QP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;
constexpr int DPP_WAVE_ROL1 = 0x134, DPP_WAVE_ROR1 = 0x13C, DPP_ROW_ROR1 = 0x121, DPP_ROW_ROR15 = 0x12F;
// quad_perm [1,2,3,0] and [3,0,1,2]: lane j of each quad reads lane (j+1)%4, (j+3)%4;
// row_ror:12 and row_ror:4: lane j of each row of 16 reads lane (j+4)%16, (j+12)%16
constexpr int DPP_QP_ROT1 = 0x39, DPP_QP_ROT3 = 0x93, DPP_ROW_ROR12 = 0x12C, DPP_ROW_ROR4 = 0x124;

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  return __hiloint2double(dpp_i<CTRL>(__double2hiint(v)), dpp_i<CTRL>(__double2loint(v)));
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL> __device__ __forceinline__ int lane_mv(int v) { return dpp_i<CTRL>(v); }
template <int CTRL> __device__ __forceinline__ float lane_mv(float v) { return dpp_f<CTRL>(v); }
template <int CTRL> __device__ __forceinline__ double lane_mv(double v) { return dpp_d<CTRL>(v); }
// value of lane l^16 (l^32): the swap exchanges odd rows of its first operand with even rows
// of its second (upper half of the first with lower half of the second); with both operands
// equal, a lane in an even row (lower half) finds its partner in the second result.
__device__ __forceinline__ int xor16_i(int v, int lane) {
  auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
  return (int)((lane & 16) ? r[0] : r[1]);
}
__device__ __forceinline__ int xor32_i(int v, int lane) {
  auto r = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
  return (int)((lane & 32) ? r[0] : r[1]);
}
__device__ __forceinline__ double xor16_d(double v, int lane) {
  return __hiloint2double(xor16_i(__double2hiint(v), lane), xor16_i(__double2loint(v), lane));
}
__device__ __forceinline__ double xor32_d(double v, int lane) {
  return __hiloint2double(xor32_i(__double2hiint(v), lane), xor32_i(__double2loint(v), lane));
}
__device__ __forceinline__ int lane_x16(int v, int lane) { return xor16_i(v, lane); }
__device__ __forceinline__ float lane_x16(float v, int lane) {
  return __int_as_float(xor16_i(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_x16(double v, int lane) { return xor16_d(v, lane); }
__device__ __forceinline__ int lane_x32(int v, int lane) { return xor32_i(v, lane); }
__device__ __forceinline__ float lane_x32(float v, int lane) {
  return __int_as_float(xor32_i(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_x32(double v, int lane) { return xor32_d(v, lane); }
#endif

// A segment of NX lanes of one wave64 = one CRM instance.  All lanes of the wave must take
// part in every call (the kernels keep control flow uniform around them).
template <int NX>
struct Seg {
  static constexpr bool Y3 = false;  // 2D: one row of NX columns per instance
  int lane, i, base;
  __device__ __forceinline__ Seg() {
    lane = threadIdx.x & 63;
    i = lane & (NX - 1);
    base = lane - i;
  }
#ifdef CRM_SEG_BPERMUTE
  // value held by column (i + d) mod NX
  template <typename T>
  __device__ __forceinline__ T rot(T v, int d) const {
    return __shfl(v, base + ((i + d + NX) & (NX - 1)), 64);
  }
  template <typename T, typename Op>
  __device__ __forceinline__ T reduce(T v, int g, Op op) const {
#pragma unroll
    for (int m = NX / 2; m >= 1; m >>= 1)
      if (m < g) v = op(v, __shfl_xor(v, m, 64));
    return v;
  }
#else
  // value held by column (i + d) mod NX, d = +1 or -1
  template <typename T>
  __device__ __forceinline__ T rot(T v, int d) const {
    if (d > 0) {
      if constexpr (NX == 64) {
        return lane_mv<DPP_WAVE_ROL1>(v);
      } else if constexpr (NX == 4) {
        return lane_mv<DPP_QP_ROT1>(v);
      } else if constexpr (NX == 32) {
        T r = lane_mv<DPP_WAVE_ROL1>(v);      // lane 31 (63) got lane 32 (0)
        T x = lane_x32(r, lane);              // lane 31 (63) now holds what lane 63 (31) got
        return i == 31 ? x : r;
      } else if constexpr (NX == 16) {
        return lane_mv<DPP_ROW_ROR15>(v);
      } else {
        T r = lane_mv<DPP_ROW_ROR15>(v);      // NX == 8: lane 7 (15) got lane 8 (0)
        T m = lane_mv<DPP_HALF_MIRROR>(v);    // lane 7 (15) reads lane 0 (8)
        return i == 7 ? m : r;
      }
    } else {
      if constexpr (NX == 64) {
        return lane_mv<DPP_WAVE_ROR1>(v);
      } else if constexpr (NX == 4) {
        return lane_mv<DPP_QP_ROT3>(v);
      } else if constexpr (NX == 32) {
        T r = lane_mv<DPP_WAVE_ROR1>(v);
        T x = lane_x32(r, lane);
        return i == 0 ? x : r;
      } else if constexpr (NX == 16) {
        return lane_mv<DPP_ROW_ROR1>(v);
      } else {
        T r = lane_mv<DPP_ROW_ROR1>(v);
        T m = lane_mv<DPP_HALF_MIRROR>(v);
        return i == 0 ? m : r;
      }
    }
  }
  // Reductions over aligned groups of g lanes pair lanes symmetrically at every step, so every
  // lane of a group ends with the bitwise same value.
  template <typename T, typename Op>
  __device__ __forceinline__ T reduce(T v, int g, Op op) const {
    if (g > 1) v = op(v, lane_mv<DPP_QP_XOR1>(v));
    if (g > 2) v = op(v, lane_mv<DPP_QP_XOR2>(v));
    if (g > 4) v = op(v, lane_mv<DPP_HALF_MIRROR>(v));
    if (g > 8) v = op(v, lane_mv<DPP_MIRROR>(v));
    if (g > 16) v = op(v, lane_x16(v, lane));
    if (g > 32) v = op(v, lane_x32(v, lane));
    return v;
  }
#endif
  template <typename T>
  __device__ __forceinline__ T sum(T v) const {
    return reduce(v, NX, [](T a, T b) { return a + b; });
  }
  // sum over aligned groups of g lanes (g a power of two dividing NX)
  template <typename T>
  __device__ __forceinline__ T gsum(T v, int g) const {
    return reduce(v, g, [](T a, T b) { return a + b; });
  }
  // sum over aligned blocks of gx x gy columns of rows of `row` lanes (3D radiation groups;
  // powers of two): the x groups first, then the rows of a block pairwise, so every lane of a
  // block ends with the bitwise same value
  template <typename T>
  __device__ __forceinline__ T gsum2(T v, int gx, int row, int gy) const {
    if (gx == row) return gsum(v, row * gy);  // whole rows: one contiguous group
    v = gsum(v, gx);
    for (int m = row; m < row * gy; m <<= 1) v = v + (T)__shfl_xor(v, m, 64);
    return v;
  }
  // gsum2 of N values at once: each butterfly step is taken (or skipped) for all of them
  // together, one uniform branch per step instead of one per value; per value the additions
  // are gsum2's, so the bits are the same
  template <typename T, int N>
  __device__ __forceinline__ void gsums2(T (&v)[N], int gx, int row, int gy) const {
    const int g = gx == row ? row * gy : gx;
#ifdef CRM_SEG_BPERMUTE
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = gsum(v[j], g);
#else
    auto step = [&](auto mv) {
#pragma unroll
      for (int j = 0; j < N; j++) v[j] = v[j] + mv(v[j]);
    };
    if (g > 1) step([](T x) { return lane_mv<DPP_QP_XOR1>(x); });
    if (g > 2) step([](T x) { return lane_mv<DPP_QP_XOR2>(x); });
    if (g > 4) step([](T x) { return lane_mv<DPP_HALF_MIRROR>(x); });
    if (g > 8) step([](T x) { return lane_mv<DPP_MIRROR>(x); });
    if (g > 16) step([&](T x) { return lane_x16(x, lane); });
    if (g > 32) step([&](T x) { return lane_x32(x, lane); });
#endif
    if (gx != row)
      for (int m = row; m < row * gy; m <<= 1)
#pragma unroll
        for (int j = 0; j < N; j++) v[j] = v[j] + (T)__shfl_xor(v[j], m, 64);
  }
  template <typename T>
  __device__ __forceinline__ T max(T v) const {
    return reduce(v, NX, [](T a, T b) { return fmax(a, b); });
  }
  __device__ __forceinline__ int imax(int v) const {
    return reduce(v, NX, [](int a, int b) { return ::max(a, b); });
  }
  __device__ __forceinline__ int imin(int v) const {
    return reduce(v, NX, [](int a, int b) { return ::min(a, b); });
  }
  // true when p holds in some lane of the wave (a superset of the segment: every segment of
  // the wave takes the branch it guards, so control flow stays uniform)
  __device__ __forceinline__ bool any(bool p) const { return __any(p); }
  // Segment sums of N values at once; afterwards v[j] holds the sum of value j in the lanes
  // where owns(j) is true (one lane per segment).  For NX == 64 the reduction is transposed:
  // v_permlane32_swap on two different values gives each half of the wave the other half's
  // share of one of them, so one swap and one add halve two sums at once; then
  // v_permlane16_swap does the same for rows, and a row butterfly finishes.  Value j ends in
  // every lane of row ROW[j % 4] of slot j / 4 (lanes 0-31 took the first value of each pair
  // at the first step, rows 0 and 2 the first of each pair at the second).
  template <int N>
  __device__ __forceinline__ void multi_sum(double (&v)[N]) const {
#ifndef CRM_SEG_BPERMUTE
    if constexpr (NX == 64 && N > 2) {
      constexpr int P = (N + 1) / 2, Q = (P + 1) / 2;
      double a[2 * Q];
#pragma unroll
      for (int p = 0; p < P; p++) {
        double x = v[2 * p], y = 2 * p + 1 < N ? v[2 * p + 1] : 0.0;
        swap32(x, y);
        a[p] = x + y;
      }
      if (P & 1) a[P] = 0.0;
#pragma unroll
      for (int t = 0; t < Q; t++) {
        double x = a[2 * t], y = a[2 * t + 1];
        swap16(x, y);
        const double r = reduce(x + y, 16, [](double e, double f) { return e + f; });
#pragma unroll
        for (int j = 4 * t; j < 4 * t + 4 && j < N; j++) v[j] = r;
      }
      return;
    }
#endif
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = sum(v[j]);
  }
  __device__ __forceinline__ bool owns(int j) const {
#ifndef CRM_SEG_BPERMUTE
    if constexpr (NX == 64) {
      constexpr int ROW[4] = {0, 2, 1, 3};
      return lane == 16 * ROW[j & 3];
    }
#endif
    return i == 0;
  }
  // N segment sums at once: FP64 values through multi_sum (an FP64 add takes no DPP operand,
  // so each butterfly step of a lone sum costs two moves and an add), FP32 values one
  // reduction each (DPP adds); sum j is then held where sum_owner<T>(j) is true
  template <typename T, int N>
  __device__ __forceinline__ void sums(T (&v)[N]) const {
    if constexpr (std::is_same<T, double>::value) {
      multi_sum(v);
    } else {
#pragma unroll
      for (int j = 0; j < N; j++) v[j] = sum(v[j]);
    }
  }
  template <typename T>
  __device__ __forceinline__ bool sum_owner(int j) const {
    if constexpr (std::is_same<T, double>::value) return owns(j);
    return i == 0;
  }
#ifndef CRM_SEG_BPERMUTE
  // (a, b) -> (a with its upper half replaced by b's lower half, b with its lower half
  // replaced by a's upper half); swap16 does the same for odd/even 16-lane rows
  __device__ __forceinline__ static void swap32(double& a, double& b) {
    auto h = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
    auto l = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
    a = __hiloint2double((int)h[0], (int)l[0]);
    b = __hiloint2double((int)h[1], (int)l[1]);
  }
  __device__ __forceinline__ static void swap16(double& a, double& b) {
    auto h = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false, false);
    auto l = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b), false, false);
    a = __hiloint2double((int)h[0], (int)l[0]);
    b = __hiloint2double((int)h[1], (int)l[1]);
  }
#endif
};

// A 3D CRM of NXY x NXY columns (8 x 8: one instance per wave; 4 x 4: four instances per
// wave): lane = NXY*NXY*instance + NXY*j + i.  i here is the column index within the instance
// (the storage index), as for the 2D segments; x neighbours come from rotations inside each
// aligned group of NXY lanes, y neighbours from a rotation by NXY lanes inside the instance,
// and reductions run over all NXY*NXY columns (nx*ny).
template <int NXY>
struct Seg3 {
  static constexpr bool Y3 = true;
  static constexpr int NN = NXY * NXY;
  static_assert(NXY == 8 || NXY == 4, "3D CRMs of 8 x 8 or 4 x 4 columns");
  Seg<NN> all;
  Seg<NXY> row;
  int lane, i;
  __device__ __forceinline__ Seg3() {
    lane = threadIdx.x & 63;
    i = lane & (NN - 1);
  }
  // value held by the column at x offset d (+1 or -1), periodic in x
  template <typename T>
  __device__ __forceinline__ T rot(T v, int d) const { return row.rot(v, d); }
  // value held by the column at y offset d (+1 or -1), periodic in y
  template <typename T>
  __device__ __forceinline__ T roty(T v, int d) const {
#ifndef CRM_SEG_BPERMUTE
    if constexpr (NXY == 4) return d > 0 ? lane_mv<DPP_ROW_ROR12>(v) : lane_mv<DPP_ROW_ROR4>(v);
#endif
    return __shfl(v, (lane & ~(NN - 1)) | ((lane + NXY * d) & (NN - 1)), 64);
  }
  template <typename T>
  __device__ __forceinline__ T sum(T v) const { return all.sum(v); }
  template <typename T>
  __device__ __forceinline__ T max(T v) const { return all.max(v); }
  __device__ __forceinline__ int imax(int v) const { return all.imax(v); }
  __device__ __forceinline__ int imin(int v) const { return all.imin(v); }
  __device__ __forceinline__ bool any(bool p) const { return __any(p); }
  template <typename T, int N>
  __device__ __forceinline__ void sums(T (&v)[N]) const { all.sums(v); }
  template <typename T>
  __device__ __forceinline__ bool sum_owner(int j) const { return all.template sum_owner<T>(j); }
};

// A 2D CRM of 128 columns on the two waves of a 128-thread workgroup, one instance per
// workgroup: column i = 64 * (wave & 1) + lane.
// - A rotation moves the wave's own 64 lanes with DPP and fetches the one column at the seam
//   from the other wave through LDS.
// - A reduction combines the two waves' butterflies through LDS, wave 0's part first, so both
//   waves hold the bitwise same result.
// Every exchange is one workgroup barrier, so both waves must make every member call in the
// same order; the kernels' control flow around these calls is uniform per instance.  Two LDS
// slots alternate: a slot is written again only after the barrier of the call in between.
struct Seg128 {
  static constexpr bool Y3 = false;
  static constexpr int SLOTS = 16;  // values per wave and exchange
  Seg<64> w;
  int lane, i, half;
  mutable int slot = 0;
  __device__ __forceinline__ Seg128() {
    lane = threadIdx.x & 63;
    half = (threadIdx.x >> 6) & 1;
    i = 64 * half + lane;
  }
  // this call's exchange area: [2 waves][SLOTS] doubles
  __device__ __forceinline__ double* xbuf() const {
    __shared__ double buf[2 * 2 * SLOTS];
    double* p = buf + slot * 2 * SLOTS;
    slot ^= 1;
    return p;
  }
  // value held by column (i + d) mod 128, d = +1 or -1
  template <typename T>
  __device__ __forceinline__ T rot(T v, int d) const {
    T r = w.rot(v, d);
    double* p = xbuf();
    if (lane == (d > 0 ? 0 : 63)) p[half] = (double)v;
    __syncthreads();
    if (lane == (d > 0 ? 63 : 0)) r = (T)p[half ^ 1];
    return r;
  }
  template <typename T, typename Op>
  __device__ __forceinline__ T combine(T v, Op op) const {
    double* p = xbuf();
    if (lane == 0) p[half] = (double)v;
    __syncthreads();
    return op((T)p[0], (T)p[1]);
  }
  template <typename T>
  __device__ __forceinline__ T sum(T v) const {
    return combine(w.sum(v), [](T a, T b) { return a + b; });
  }
  // sum over aligned groups of g lanes (g a power of two dividing 128)
  template <typename T>
  __device__ __forceinline__ T gsum(T v, int g) const {
    return g <= 64 ? w.gsum(v, g) : sum(v);
  }
  // a 2D CRM: the group is gx contiguous columns of one row (gy == 1)
  template <typename T>
  __device__ __forceinline__ T gsum2(T v, int gx, int, int) const {
    return gsum(v, gx);
  }
  template <typename T, int N>
  __device__ __forceinline__ void gsums2(T (&v)[N], int gx, int, int) const {
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = gsum(v[j], gx);
  }
  template <typename T>
  __device__ __forceinline__ T max(T v) const {
    return combine(w.max(v), [](T a, T b) { return fmax(a, b); });
  }
  __device__ __forceinline__ int imax(int v) const {
    return combine(w.imax(v), [](int a, int b) { return ::max(a, b); });
  }
  __device__ __forceinline__ int imin(int v) const {
    return combine(w.imin(v), [](int a, int b) { return ::min(a, b); });
  }
  // true when p holds in some lane of either wave: both waves take the branch it guards
  __device__ __forceinline__ bool any(bool p) const {
    return combine((int)__any(p), [](int a, int b) { return a | b; }) != 0;
  }
  template <int N>
  __device__ __forceinline__ void multi_sum(double (&v)[N]) const {
    static_assert(N <= SLOTS, "too many values for one exchange");
    w.multi_sum(v);
    double* p = xbuf();
#pragma unroll
    for (int j = 0; j < N; j++)
      if (w.owns(j)) p[half * SLOTS + j] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; j++) v[j] = p[j] + p[SLOTS + j];
  }
  __device__ __forceinline__ bool owns(int) const { return i == 0; }
  template <typename T, int N>
  __device__ __forceinline__ void sums(T (&v)[N]) const {
    if constexpr (std::is_same<T, double>::value) {
      multi_sum(v);
    } else {
#pragma unroll
      for (int j = 0; j < N; j++) v[j] = sum(v[j]);
    }
  }
  template <typename T>
  __device__ __forceinline__ bool sum_owner(int) const { return i == 0; }
};

// The lane geometry of a 2D CRM of NX columns.
template <int NX>
struct SegSel {
  using type = Seg<NX>;
};
template <>
struct SegSel<128> {
  using type = Seg128;
};
template <int NX>
using SegFor = typename SegSel<NX>::type;

// Instance of this thread's segment; false when past ncrms (the segment then idles but still
// joins cross-lane operations of its wave).  A segment of 64 or more lanes fills whole waves,
// so its instance is wave-uniform; UNIFORM tells the compiler so, which turns the per-instance
// profile loads into scalar loads.  A kernel whose scalar registers would then spill (k_stats
// measured slower with it) passes false.
template <int NX, bool UNIFORM = true>
__device__ __forceinline__ bool seg_instance(int ncrms, int& c) {
  int g = (blockIdx.x * blockDim.x + threadIdx.x) / NX;
  if constexpr (NX >= 64 && UNIFORM) g = __builtin_amdgcn_readfirstlane(g);
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

// x**y for x >= 0: the library pow in FP64; in FP32 the hardware base-2 logarithm and
// exponential (about 3e-7 relative for the microphysics exponents 0.1-1.0 over the
// mixing-ratio range; the library powf evaluates in extended precision and costs ~10x more).
template <typename T>
__device__ __forceinline__ T fpow(T x, T y);
template <>
__device__ __forceinline__ double fpow<double>(double x, double y) {
  return pow(x, y);
}
template <>
__device__ __forceinline__ float fpow<float>(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// x**y for x >= 0 in the precipitation terminal velocity: in FP64 pow_pos (crm_pow.hpp,
// within 2 ulp of the correctly rounded pow, with far fewer instructions than the library
// pow: the sedimentation sweep measured 83.5 -> 68.2 ms per 240-substep call at 4096 x 64 x
// 58).  The process rates of k_micro keep the library pow: that kernel is at its register
// limit, and the inlined form there spilled more and measured slower (109 -> 125 ms).
template <typename T>
__device__ __forceinline__ T fpow_fall(T x, T y) {
  return fpow<T>(x, y);
}
template <>
__device__ __forceinline__ double fpow_fall<double>(double x, double y) {
  return ::crm::pow_pos(x, y);
}

// ---- saturation functions, crm/sat.F90:12-139 (REAL(4) coefficients) ----------------------
// T is double in the FP64 form and float for the FP32 fields of the mixed form.
template <typename T>
__device__ __forceinline__ T poly8(const float* a, T dt) {
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

template <typename T>
__device__ __forceinline__ T esat_low(T t) {
  return exp((T)R4(9.550426) - (T)R4(5723.265) / t + (T)R4(3.53068) * log(t) - (T)R4(0.00728332) * t);
}
template <typename T>
__device__ __forceinline__ T esatw(T t) {
  T dt = t - (T)R4(273.16);
  if (dt > (T)-80.0) return poly8(c_esatw, dt);
  return (T)R4(0.01) * (T)2.0 * esat_low(t);
}
template <typename T>
__device__ __forceinline__ T esati(T t) {
  T dt = t - (T)R4(273.16);
  if (dt > (T)-80.0) return poly8(c_esati, dt);
  return (T)R4(0.01) * esat_low(t);
}
template <typename T>
__device__ __forceinline__ T qsatw(T t, T p) {
  T e = esatw(t);
  return (T)R4(0.622) * e / fmax(e, p - e);
}
template <typename T>
__device__ __forceinline__ T qsati(T t, T p) {
  T e = esati(t);
  return (T)R4(0.622) * e / fmax(e, p - e);
}
template <typename T>
__device__ __forceinline__ T dtqsatw(T t, T p) {
  T dt = t - (T)R4(273.16);
  T de = dt > (T)-80.0 ? poly8(c_dtesatw, dt) : esatw(t + (T)1) - esatw(t);
  return (T)R4(0.622) * de / p;
}
template <typename T>
__device__ __forceinline__ T dtqsati(T t, T p) {
  T dt = t - (T)R4(273.16);
  T de = dt > (T)-80.0 ? poly8(c_dtesati, dt) : esati(t + (T)1.0) - esati(t);
  return (T)R4(0.622) * de / p;
}

}  // namespace CRM_NS
