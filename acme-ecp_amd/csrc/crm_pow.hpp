// crm_pow.hpp -- x**y in FP64 for the microphysics: finite x >= 0 and |y ln x| < 700.
//
// The microphysics raises mixing ratios and densities to fixed powers between 0.1 and 1.94
// (terminal velocities, microphysics.F90:453-486; process rates, precip_proc.F90:41-161;
// ice fall, ice_fall.F90:113).  The general library pow spends most of its work on cases that
// never occur here (negative bases, integer exponents, extended-precision logarithms), so it
// dominated the precipitation sweeps.  This one follows the textbook route:
//
//   ln x = e ln2 + ln m,  x = m 2^e with m in [sqrt(1/2), sqrt(2));
//   ln m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^21/21),  s = (m - 1)/(m + 1), |s| <= 0.1716,
//          so the first omitted term is below 2e-17 of ln m;
//   ln x is carried as hi + lo: ln 2 is split (LN2_HI has 31 significant bits, so e*LN2_HI is
//          exact) and the sum e*LN2_HI + 2s is split exactly (Dekker's fast two-sum);
//   t = y ln x keeps the rounding error of the product (an fma), so t is good to ~1e-16
//          absolute however large |t| is;
//   exp(t) = 2^n exp(r),  r = t - n ln2 (Cody and Waite's two-constant reduction), |r| <= 0.35,
//          exp(r) summed as its Taylor series to r^13 (first omitted term below 5e-18).
//
// x == 0 gives 0 and x == +inf gives +inf, as pow does for y > 0; negative or NaN x gives NaN.
// Measured against the correctly rounded libm pow over the microphysics' arguments (and over
// x in [1e-300, 1e3], y in (0, 2]): within 2 ulp.  The same source builds for the host so the
// CPU can check it.
#pragma once

#if defined(__HIP__)
#define CRM_POW_HD __host__ __device__
#else
#define CRM_POW_HD
#endif

namespace crm {

CRM_POW_HD inline double pow_pos(double x, double y) {
  constexpr double LN2_HI = 0x1.62e42fee00000p-1;   // ln 2 rounded to 31 bits
  constexpr double LN2_LO = 0x1.a39ef35793c76p-33;  // ln 2 - LN2_HI
  constexpr double INV_LN2 = 0x1.71547652b82fep+0;
  constexpr double SQRT2 = 0x1.6a09e667f3bcdp+0;
  const bool pos = x > 0.0;
  double xs = pos ? x : 1.0;
  const bool sub = xs < 0x1p-1022;  // subnormal: scale into the normal range first
  xs = sub ? xs * 0x1p54 : xs;
  unsigned long long b;
  __builtin_memcpy(&b, &xs, sizeof(b));
  int e = (int)(b >> 52) - 1023 - (sub ? 54 : 0);
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  __builtin_memcpy(&m, &b, sizeof(m));
  const bool big = m > SQRT2;
  m = big ? 0.5 * m : m;
  e += big ? 1 : 0;
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  // sum_{k=1..10} 2 z^k / (2k + 1)
  double p = 2.0 / 21;
  p = __builtin_fma(p, z, 2.0 / 19);
  p = __builtin_fma(p, z, 2.0 / 17);
  p = __builtin_fma(p, z, 2.0 / 15);
  p = __builtin_fma(p, z, 2.0 / 13);
  p = __builtin_fma(p, z, 2.0 / 11);
  p = __builtin_fma(p, z, 2.0 / 9);
  p = __builtin_fma(p, z, 2.0 / 7);
  p = __builtin_fma(p, z, 2.0 / 5);
  p = __builtin_fma(p, z, 2.0 / 3);
  p = p * z;
  const double de = (double)e;
  const double a = de * LN2_HI;
  const double s2 = 2.0 * s;
  const double hi = a + s2;
  const double lo = (s2 - (hi - a)) + __builtin_fma(s, p, de * LN2_LO);
  const double th = y * hi;
  const double tl = __builtin_fma(y, hi, -th) + y * lo;
  const double n = __builtin_rint(th * INV_LN2);
  const double r = __builtin_fma(-n, LN2_LO, __builtin_fma(-n, LN2_HI, th)) + tl;
  // sum_{k=0..13} r^k / k!
  double q = 1.0 / 6227020800.0;
  q = __builtin_fma(q, r, 1.0 / 479001600.0);
  q = __builtin_fma(q, r, 1.0 / 39916800.0);
  q = __builtin_fma(q, r, 1.0 / 3628800.0);
  q = __builtin_fma(q, r, 1.0 / 362880.0);
  q = __builtin_fma(q, r, 1.0 / 40320.0);
  q = __builtin_fma(q, r, 1.0 / 5040.0);
  q = __builtin_fma(q, r, 1.0 / 720.0);
  q = __builtin_fma(q, r, 1.0 / 120.0);
  q = __builtin_fma(q, r, 1.0 / 24.0);
  q = __builtin_fma(q, r, 1.0 / 6.0);
  q = __builtin_fma(q, r, 0.5);
  q = __builtin_fma(q, r, 1.0);
  q = __builtin_fma(q, r, 1.0);
  const double res = __builtin_ldexp(q, (int)n);
  if (pos) return x == __builtin_inf() ? x : res;
  return x == 0.0 ? 0.0 : __builtin_nan("");
}

}  // namespace crm
