// fft991_operator.hpp -- dense matrices of the reference's horizontal transform.
//
// The pressure solve of the reference transforms every level of the Poisson right-hand side
// with fft991_crm (crm/fft.F90:5-153): a half-complex real FFT whose helper routines keep
// their trig tables and pass constants in implicit REAL(4) (fft.F90:247-295, 382-466,
// 469-808), applied in REAL(8) arithmetic.  It is an exact linear map, so it equals a dense
// matrix; the GPU applies that matrix (a batched DFT-as-GEMM) instead of the FFT, and
// reproduces the reference transform to FP64 rounding instead of to the REAL(4) twiddles'
// 1e-7.
//
// This header evaluates the reference transform on unit vectors, once per nx, on the host.
//
// Coefficient order of the reference: a0,b0,a1,b1,...,a_{n/2},b_{n/2} with b0 = b_{n/2} = 0
// (n+2 values).  The device uses the n meaningful ones in a square layout:
//   slot 0 = a0, slot 1 = a_{n/2}, slot 2m = a_m, slot 2m+1 = b_m (m = 1..n/2-1).
#pragma once
#include <cmath>
#include <vector>

namespace crm {

struct Fft991 {
  int n;
  std::vector<int> ifax;
  std::vector<double> trigs;

  explicit Fft991(int n_) : n(n_), ifax(20, 0), trigs(3 * n_ + 8, 0.0) {
    factorize();
    make_trigs();
  }

  // factors of n/2, ascending (fft.F90:181-245, mode 3)
  void factorize() {
    int nn = n / 2, k = 1;
    auto done = [&]() {
      ifax[0] = k - 1;
      for (int a = 1; a < k; a++)
        for (int b = 1; b < k - a; b++)
          if (ifax[b + 1] < ifax[b]) std::swap(ifax[b], ifax[b + 1]);
    };
    for (int f : {4, 2, 3}) {
      bool once = (f == 2);
      while (nn % f == 0) {
        ifax[k++] = f;
        nn /= f;
        if (nn == 1) return done();
        if (once) break;
      }
    }
    for (int l = 5, inc = 2;; l += inc, inc = 6 - inc)
      while (nn % l == 0) {
        ifax[k++] = l;
        nn /= l;
        if (nn == 1) return done();
      }
  }

  // trig table with REAL(4) angle, step and result (fft.F90:247-295)
  void make_trigs() {
    float pi = 2.0f * std::asin(1.0f);
    int nn = n / 2;
    float del = (float)((double)(pi + pi) / nn);
    for (int i = 1; i <= 2 * nn; i += 2) {
      float ang = (float)(0.5 * (i - 1) * (double)del);
      trigs[i - 1] = std::cos(ang);
      trigs[i] = std::sin(ang);
    }
    del = 0.5f * del;
    int nh = (nn + 1) / 2, la = 2 * nn;
    for (int i = 1; i <= 2 * nh; i += 2) {
      float ang = (float)(0.5 * (i - 1) * (double)del);
      trigs[la + i - 1] = std::cos(ang);
      trigs[la + i] = std::sin(ang);
    }
  }

  // One complex pass of factor f over interleaved (re,im) data: in -> out.
  // Index conventions follow vpassm_crm (fft.F90:469-808) with lot = 1.
  void pass(const double* a, const double* b, double* c, double* d, int inc1, int inc2,
            int nh, int f, int la) const {
    const float s36 = 0.587785252292473f, c36 = 0.809016994374947f;
    const float s72 = 0.951056516295154f, c72 = 0.309016994374947f;
    const float s60 = 0.866025403784437f;
    int m = nh / f, iink = m * inc1, jink = la * inc2, jump = (f - 1) * jink;
    int ib = 0, jb = 0;
    for (int kk = 0;; kk++) {
      int k = kk == 0 ? 1 : la + 1 + (kk - 1) * la;
      if (kk > 0 && k > m) return;
      float cs[5] = {1, 1, 1, 1, 1}, sn[5] = {0, 0, 0, 0, 0};
      if (kk > 0) {
        int kb = k + k - 2;
        for (int q = 1; q < f; q++) {
          cs[q] = (float)trigs[q * kb];
          sn[q] = (float)trigs[q * kb + 1];
        }
      }
      for (int l = 0; l < la; l++, ib += inc1, jb += inc2) {
        const double* A[5];
        const double* B[5];
        double* C[5];
        double* D[5];
        for (int q = 0; q < f; q++) {
          A[q] = a + ib + q * iink;
          B[q] = b + ib + q * iink;
          C[q] = c + jb + q * jink;
          D[q] = d + jb + q * jink;
        }
        auto tw = [&](int q, double re, double im) {
          if (kk == 0) {
            *C[q] = re;
            *D[q] = im;
          } else {
            *C[q] = cs[q] * re - sn[q] * im;
            *D[q] = sn[q] * re + cs[q] * im;
          }
        };
        if (f == 2) {
          *C[0] = *A[0] + *A[1];
          *D[0] = *B[0] + *B[1];
          tw(1, *A[0] - *A[1], *B[0] - *B[1]);
        } else if (f == 3) {
          double p1 = *A[0] - 0.5 * (*A[1] + *A[2]), q1 = s60 * (*B[1] - *B[2]);
          double p2 = *B[0] - 0.5 * (*B[1] + *B[2]), q2 = s60 * (*A[1] - *A[2]);
          *C[0] = *A[0] + (*A[1] + *A[2]);
          *D[0] = *B[0] + (*B[1] + *B[2]);
          tw(1, p1 - q1, p2 + q2);
          tw(2, p1 + q1, p2 - q2);
        } else if (f == 4) {
          double ap = *A[0] + *A[2], am = *A[0] - *A[2], bp = *B[0] + *B[2], bm = *B[0] - *B[2];
          double cp = *A[1] + *A[3], cm = *A[1] - *A[3], dp = *B[1] + *B[3], dm = *B[1] - *B[3];
          *C[0] = ap + cp;
          *D[0] = bp + dp;
          tw(2, ap - cp, bp - dp);
          tw(1, am - dm, bm + cm);
          tw(3, am + dm, bm - cm);
        } else {  // f == 5
          double a1 = *A[0] + c72 * (*A[1] + *A[4]) - c36 * (*A[2] + *A[3]);
          double b1 = s72 * (*B[1] - *B[4]) + s36 * (*B[2] - *B[3]);
          double a2 = *B[0] + c72 * (*B[1] + *B[4]) - c36 * (*B[2] + *B[3]);
          double b2 = s72 * (*A[1] - *A[4]) + s36 * (*A[2] - *A[3]);
          double a3 = *A[0] - c36 * (*A[1] + *A[4]) + c72 * (*A[2] + *A[3]);
          double b3 = s36 * (*B[1] - *B[4]) - s72 * (*B[2] - *B[3]);
          double a4 = *B[0] - c36 * (*B[1] + *B[4]) + c72 * (*B[2] + *B[3]);
          double b4 = s36 * (*A[1] - *A[4]) - s72 * (*A[2] - *A[3]);
          *C[0] = *A[0] + (*A[1] + *A[4]) + (*A[2] + *A[3]);
          *D[0] = *B[0] + (*B[1] + *B[4]) + (*B[2] + *B[3]);
          tw(1, a1 - b1, a2 + b2);
          tw(4, a1 + b1, a2 - b2);
          tw(2, a3 - b3, a4 + b4);
          tw(3, a3 + b3, a4 - b4);
        }
      }
      if (kk == 0 && la == m) return;
      jb += jump;
    }
  }

  // a: n+2 values, transformed in place; isign -1 grid->spectral, +1 spectral->grid.
  void apply(double* a, int isign) const {
    std::vector<double> w(n + 4, 0.0);
    int nh = n / 2, nfax = ifax[0];
    bool in_work;  // current data location before the next pass
    if (isign == 1) {
      // pre-processing (fft.F90:297-378)
      w[0] = a[0] + a[n];
      w[1] = a[0] - a[n];
      int ia = 3, ib = n - 1;
      for (int k = 3; k <= nh; k += 2, ia += 2, ib -= 2) {
        float c = (float)trigs[n + k - 1], s = (float)trigs[n + k];
        double A1 = a[ia - 1], B1 = a[ib - 1], A2 = a[ia], B2 = a[ib];
        w[ia - 1] = (A1 + B1) - (s * (A1 - B1) + c * (A2 + B2));
        w[ib - 1] = (A1 + B1) + (s * (A1 - B1) + c * (A2 + B2));
        w[ia] = (c * (A1 - B1) - s * (A2 + B2)) + (A2 - B2);
        w[ib] = (c * (A1 - B1) - s * (A2 + B2)) - (A2 - B2);
      }
      if (ia == ib) {
        w[ia - 1] = 2.0 * a[ia - 1];
        w[ia] = -2.0 * a[ia];
      }
      in_work = true;
    } else {
      in_work = false;
      if (nfax % 2 == 0) {
        for (int m = 0; m < n; m++) w[m] = a[m];
        in_work = true;
      }
    }
    int la = 1;
    for (int k = 1; k <= nfax; k++) {
      if (in_work)
        pass(&w[0], &w[1], &a[0], &a[1], 2, 2, nh, ifax[k], la);
      else
        pass(&a[0], &a[1], &w[0], &w[1], 2, 2, nh, ifax[k], la);
      in_work = !in_work;
      la *= ifax[k];
    }
    if (isign == -1) {
      // post-processing with REAL(4) scale, c, s (fft.F90:380-466); data is in w
      float scale = (float)(1.0 / n);
      a[0] = scale * (w[0] + w[1]);
      a[n] = scale * (w[0] - w[1]);
      a[1] = 0.0;
      a[n + 1] = 0.0;
      scale = 0.5f * scale;
      int ia = 3, ib = n - 1;
      for (int k = 3; k <= nh; k += 2, ia += 2, ib -= 2) {
        float c = (float)trigs[n + k - 1], s = (float)trigs[n + k];
        double W1 = w[ia - 1], W2 = w[ib - 1], W3 = w[ia], W4 = w[ib];
        a[ia - 1] = scale * ((W1 + W2) + (c * (W3 + W4) + s * (W1 - W2)));
        a[ib - 1] = scale * ((W1 + W2) - (c * (W3 + W4) + s * (W1 - W2)));
        a[ia] = scale * ((c * (W1 - W2) - s * (W3 + W4)) + (W4 - W3));
        a[ib] = scale * ((c * (W1 - W2) - s * (W3 + W4)) - (W4 - W3));
      }
      if (ia == ib) {
        scale = 2.0f * scale;
        a[ia - 1] = scale * w[ia - 1];
        a[ia] = -scale * w[ia];
      }
      return;
    }
    if (in_work)
      for (int m = 0; m < n; m++) a[m] = w[m];
    a[n] = 0.0;
    a[n + 1] = 0.0;
  }
};

// Fortran coefficient index (0-based, n+2 entries) of device slot s.
inline int fft991_slot_to_coef(int n, int s) {
  if (s == 0) return 0;
  if (s == 1) return n;
  return s;
}

// Forward operator Ft[i][s] (coefficient slot s from grid value i) and inverse Gt[s][i]
// (grid value i from slot s), both n x n, row-major as written.
inline void fft991_operators(int n, std::vector<double>& Ft, std::vector<double>& Gt) {
  Fft991 t(n);
  Ft.assign((size_t)n * n, 0.0);
  Gt.assign((size_t)n * n, 0.0);
  std::vector<double> a(n + 2);
  for (int i = 0; i < n; i++) {
    std::fill(a.begin(), a.end(), 0.0);
    a[i] = 1.0;
    t.apply(a.data(), -1);
    for (int s = 0; s < n; s++) Ft[(size_t)i * n + s] = a[fft991_slot_to_coef(n, s)];
  }
  for (int s = 0; s < n; s++) {
    std::fill(a.begin(), a.end(), 0.0);
    a[fft991_slot_to_coef(n, s)] = 1.0;
    t.apply(a.data(), 1);
    for (int i = 0; i < n; i++) Gt[(size_t)s * n + i] = a[i];
  }
}

// Horizontal eigenvalue mode number of slot s (pressure.F90:172-198: id = (i-0.1)/2 with i
// the 1-based Fortran coefficient index).
inline int fft991_slot_mode(int n, int s) {
  if (s == 1) return n / 2;
  return s / 2;
}

}  // namespace crm
