// crm_mlp.hip -- the pointwise network that stands in for the bulk microphysics, trained and
// evaluated on the matrix cores: BF16 operands, FP32 accumulation (v_mfma_f32_16x16x16_bf16),
// FP32 master weights and Adam moments.
//
// Every layer, forward and backward, is one GEMM C = A (M x K) * B (K x N) with arbitrary
// strides for A and B, so the transposes of the backward pass cost nothing; the epilogue adds
// the bias and applies ReLU (forward), or applies the ReLU mask of the forward activations
// (backward), or accumulates weight gradients split over the batch (FP32 atomics).
//
// Layouts (row-major, y = x W + b): the 9 inputs are padded to 16 columns and the 6 outputs
// to 16, with zero weights in the padding; batches are padded to a multiple of 16 rows with
// zero inputs and zero output gradients, so padded rows contribute nothing to any gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/crm_mlp.h"
#include "crm_mlp_dev.hpp"

namespace crm_mlp_ns {

enum { EPI_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_BIAS_F32 = 2, EPI_MASK_BF16 = 3, EPI_ATOMIC_F32 = 4 };

// One wave per 16 x 16 tile of C; blockIdx.y selects a chunk of K (EPI_ATOMIC_F32 only).
// A[m][k] = A[m*sam + k*sak], B[k][n] = B[k*sbk + n*sbn]; M, N, K, kchunk multiples of 16.
// MFMA 16x16x16 BF16 operand layout (lane l, r = l % 16, q = l / 16):
//   A: row r, k = 4q..4q+3     B: k = 4q..4q+3, column r     D: rows 4q..4q+3, column r
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(const bf16* A, long sam, long sak, const bf16* B, long sbk, long sbn,
                                              int M, int N, int K, int kchunk, const float* bias,
                                              const bf16* mask, long smm, float* Cf, bf16* Cb, long ldc) {
  const int lane = threadIdx.x & 63;
  const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int tiles_n = N / 16;
  const long tm = wave / tiles_n, tn = wave % tiles_n;
  if (tm * 16 >= M) return;  // the whole wave leaves together
  const int r = lane & 15, q = lane >> 4;
  const int k0 = blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const bf16* a = A + (tm * 16 + r) * sam;
  const bf16* b = B + (tn * 16 + r) * sbn;
  for (int k = k0; k < k1; k += 16) {
    s4 av, bv;
    // K contiguous (forward activations, transposed weights): one 8-byte load per operand
    if (sak == 1) {
      av = *(const s4*)&a[k + 4 * q];
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) av[j] = (short)a[(long)(k + 4 * q + j) * sak];
    }
    if (sbk == 1) {
      bv = *(const s4*)&b[k + 4 * q];
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) bv[j] = (short)b[(long)(k + 4 * q + j) * sbk];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const long m = tm * 16 + 4 * q + j, n = tn * 16 + r;
    const float v = acc[j];
    if constexpr (EPI == EPI_F32) Cf[m * ldc + n] = v;
    if constexpr (EPI == EPI_ATOMIC_F32) unsafeAtomicAdd(&Cf[m * ldc + n], v);
    if constexpr (EPI == EPI_BIAS_F32) Cf[m * ldc + n] = v + bias[n];
    if constexpr (EPI == EPI_BIAS_RELU_BF16) Cb[m * ldc + n] = to_bf16(fmaxf(v + bias[n], 0.f));
    if constexpr (EPI == EPI_MASK_BF16) Cb[m * ldc + n] = to_bf16(from_bf16(mask[m * smm + n]) > 0.f ? v : 0.f);
  }
}

// Raw inputs X[n][9] -> normalised BF16 rows [mp][16] (zero padding).
__global__ void k_inputs(const float* X, int n, int mp, const float* norm, bf16* Xn) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)mp * KP) return;
  const long m = idx / KP;
  const int c = (int)(idx % KP);
  float v = 0.f;
  if (m < n && c < CRM_MLP_NIN) v = (X[m * CRM_MLP_NIN + c] - norm[NORM_MEAN + c]) * norm[NORM_RSCALE + c];
  Xn[idx] = to_bf16(v);
}

// Loss of the normalised outputs and its gradient: dY = 2 (Y - T / oscale) / (6 n) in BF16.
// Grid-stride over the elements; one atomic per workgroup (a few hundred in all, so they do not
// queue on the one address).
__global__ void k_loss(const float* Y, const float* T, int n, int mp, const float* norm, bf16* dY, float* loss) {
  __shared__ float red[4];
  float e2 = 0.f;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < (long)mp * NP;
       idx += (long)gridDim.x * blockDim.x) {
    const long m = idx / NP;
    const int c = (int)(idx % NP);
    float g = 0.f;
    if (m < n && c < CRM_MLP_NOUT) {
      const float d = Y[idx] - T[m * CRM_MLP_NOUT + c] / norm[NORM_OSCALE + c];
      e2 += d * d;
      g = 2.f * d / (6.f * (float)n);
    }
    dY[idx] = to_bf16(g);
  }
  for (int o = 32; o >= 1; o >>= 1) e2 += __shfl_xor(e2, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = red[0] + red[1] + red[2] + red[3];
    if (b != 0.f) unsafeAtomicAdd(loss, b / (6.f * (float)n));
  }
}

// Column sums of a BF16 [mp][ld] matrix over its rows, accumulated into out[ncol] (FP32).  A
// workgroup (256 threads) covers 64 columns with 4 row groups that meet in LDS; blockIdx.y
// splits the rows.
__global__ void k_colsum(const bf16* G, int mp, int ld, int ncol, float* out) {
  __shared__ float red[4][64];
  const int cl = threadIdx.x & 63, grp = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  const long step = (long)gridDim.y * 4;
  float sacc = 0.f;
  if (c < ncol)
    for (long m = (long)blockIdx.y * 4 + grp; m < mp; m += step) sacc += from_bf16(G[m * ld + c]);
  red[grp][cl] = sacc;
  __syncthreads();
  if (grp == 0 && c < ncol) unsafeAtomicAdd(&out[c], (red[0][cl] + red[1][cl]) + (red[2][cl] + red[3][cl]));
}

// Packed FP32 parameters -> padded BF16 weights and FP32 biases used by the GEMMs.
__global__ void k_unpack(const float* p, bf16* W1, bf16* W2, bf16* W3, bf16* W4, float* b1, float* b2, float* b3,
                         float* b4) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  // W1p [16][128]
  if (idx < KP * NH) {
    int k = idx / NH, n = idx % NH;
    W1[idx] = to_bf16(k < CRM_MLP_NIN ? p[OFF_W1 + k * NH + n] : 0.f);
  }
  if (idx < NH * NH) {
    W2[idx] = to_bf16(p[OFF_W2 + idx]);
    W3[idx] = to_bf16(p[OFF_W3 + idx]);
  }
  // W4p [128][16]
  if (idx < NH * NP) {
    int k = idx / NP, n = idx % NP;
    W4[idx] = to_bf16(n < CRM_MLP_NOUT ? p[OFF_W4 + k * CRM_MLP_NOUT + n] : 0.f);
  }
  if (idx < NH) {
    b1[idx] = p[OFF_B1 + idx];
    b2[idx] = p[OFF_B2 + idx];
    b3[idx] = p[OFF_B3 + idx];
  }
  if (idx < NP) b4[idx] = idx < CRM_MLP_NOUT ? p[OFF_B4 + idx] : 0.f;
}

// Padded gradients -> packed gradient vector.
__global__ void k_pack_grad(const float* gW1, const float* gW2, const float* gW3, const float* gW4, const float* gb1,
                            const float* gb2, const float* gb3, const float* gb4, float* g) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= CRM_MLP_NPARAM) return;
  float v;
  if (idx < OFF_B1) {
    v = gW1[idx];  // rows 0..8 of [16][128]
  } else if (idx < OFF_W2) {
    v = gb1[idx - OFF_B1];
  } else if (idx < OFF_B2) {
    v = gW2[idx - OFF_W2];
  } else if (idx < OFF_W3) {
    v = gb2[idx - OFF_B2];
  } else if (idx < OFF_B3) {
    v = gW3[idx - OFF_W3];
  } else if (idx < OFF_W4) {
    v = gb3[idx - OFF_B3];
  } else if (idx < OFF_B4) {
    int e = idx - OFF_W4, k = e / CRM_MLP_NOUT, n = e % CRM_MLP_NOUT;
    v = gW4[k * NP + n];
  } else {
    v = gb4[idx - OFF_B4];
  }
  g[idx] = v;
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, float lr, float c1, float c2) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= CRM_MLP_NPARAM) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float mi = b1 * m[idx] + (1.f - b1) * g[idx];
  float vi = b2 * v[idx] + (1.f - b2) * g[idx] * g[idx];
  m[idx] = mi;
  v[idx] = vi;
  p[idx] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Fused inference, the whole network in registers.  The layers run transposed, H^T = W^T A^T:
// the weights are the A operand (rows = output features) and the points the B operand
// (columns), so the D fragment of one layer -- lane (r, q) holds features 4q..4q+3 of point r
// -- is the B fragment of the next one, and activations never leave the registers.  The
// 128-wide layers use v_mfma_f32_16x16x32_bf16, whose 32 K slots a lane fills with eight
// values; the K order inside a 32-feature block is a free choice as long as A and B agree,
// and we take slot (q, i) = feature 4q + i (i < 4) or 16 + 4q + (i - 4) (i >= 4), which is
// exactly the two D fragments of output blocks 2kb and 2kb + 1.  The weights are stored once
// per network in that fragment order (k_frags: each lane's operand is 8 or 16 contiguous
// bytes) and copied into LDS by every workgroup; a wave owns 64 points (4 blocks of 16) per
// tile and reads each weight fragment once for the 4 blocks.  The grid is persistent: two
// workgroups per CU (the weights take 74 KiB of LDS) loop over the tiles.
constexpr int F_L1 = 0;                     // layer 1: 8 output blocks x 64 lanes x 4 (K = 16)
constexpr int F_L2 = F_L1 + 8 * 64 * 4;     // layer 2: (8 output x 4 K blocks) x 64 lanes x 8
constexpr int F_L3 = F_L2 + 32 * 64 * 8;    // layer 3: as layer 2
constexpr int F_L4 = F_L3 + 32 * 64 * 8;    // layer 4: 4 K blocks x 64 lanes x 8 (16 outputs)
constexpr int F_SIZE = F_L4 + 4 * 64 * 8;   // bf16 elements
constexpr int B_SIZE = 3 * NH + NP;         // biases b1 b2 b3 b4

typedef unsigned u4 __attribute__((ext_vector_type(4)));  // 8 bf16, element 0 in the low half
typedef unsigned u2 __attribute__((ext_vector_type(2)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef __bf16 bh2 __attribute__((ext_vector_type(2)));
typedef short sh2 __attribute__((ext_vector_type(2)));

// Input feature (or K index inside a 32-block) of slot i of lane quarter q.
__device__ __forceinline__ int slot32(int q, int i) { return i < 4 ? 4 * q + i : 16 + 4 * q + (i - 4); }

// Weights in fragment order, from the packed FP32 parameters (one thread per element).
__global__ void k_frags(const float* p, bf16* F) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= F_SIZE) return;
  float w = 0.f;
  if (idx < F_L2) {
    const int ob = idx / 256, l = (idx / 4) % 64, i = idx % 4;
    const int n = 16 * ob + l % 16, k = 4 * (l / 16) + i;
    if (k < CRM_MLP_NIN) w = p[OFF_W1 + k * NH + n];
  } else if (idx < F_L4) {
    const int j = idx < F_L3 ? idx - F_L2 : idx - F_L3, off = idx < F_L3 ? OFF_W2 : OFF_W3;
    const int fr = j / 512, ob = fr / 4, kb = fr % 4, l = (j / 8) % 64, i = j % 8;
    const int n = 16 * ob + l % 16, k = 32 * kb + slot32(l / 16, i);
    w = p[off + k * NH + n];
  } else {
    const int j = idx - F_L4, kb = j / 512, l = (j / 8) % 64, i = j % 8;
    const int n = l % 16, k = 32 * kb + slot32(l / 16, i);
    if (n < CRM_MLP_NOUT) w = p[OFF_W4 + k * CRM_MLP_NOUT + n];
  }
  F[idx] = to_bf16(w);
}

__device__ __forceinline__ f4 mfma32(u4 a, u4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a), __builtin_bit_cast(b8, b), c, 0, 0, 0);
}

// relu of two values rounded to BF16 (nearest, ties to even: v_cvt_pk_bf16_f32), packed.  The
// rounding keeps the sign, so relu after it equals relu before it, and on BF16 bit patterns
// relu is the signed 16-bit max with 0 (negative values and -0 have the sign bit set).
__device__ __forceinline__ unsigned relu_pk(float a, float b) {
  const bh2 h = __builtin_convertvector((f2){a, b}, bh2);
  return __builtin_bit_cast(unsigned, __builtin_elementwise_max(__builtin_bit_cast(sh2, h), (sh2){0, 0}));
}

// D fragment of output block ob -> its half of the next layer's B fragment of block ob / 2.
__device__ __forceinline__ void relu_store(const f4& acc, u4& dst, int half) {
  dst[2 * half] = relu_pk(acc[0], acc[1]);
  dst[2 * half + 1] = relu_pk(acc[2], acc[3]);
}

// A 128 -> 128 hidden layer for the wave's four point blocks.  The epilogue of output block
// ob - 1 is written after the products of block ob, so it runs under them.
__device__ __forceinline__ void hidden128(const bf16* fr, const float* bias, const u4 (&in)[4][4], u4 (&out)[4][4],
                                          int lane) {
  const int q = lane >> 4;
  f4 acc[2][4];
#pragma unroll
  for (int ob = 0; ob < 8; ob++) {
    const f4 b = *(const f4*)&bias[16 * ob + 4 * q];
    f4(&a)[4] = acc[ob & 1];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) a[pb] = b;
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const u4 w = *(const u4*)&fr[((ob * 4 + kb) * 64 + lane) * 8];
#pragma unroll
      for (int pb = 0; pb < 4; pb++) a[pb] = mfma32(w, in[pb][kb], a[pb]);
    }
    if (ob > 0) {
#pragma unroll
      for (int pb = 0; pb < 4; pb++) relu_store(acc[(ob - 1) & 1][pb], out[pb][(ob - 1) >> 1], (ob - 1) & 1);
    }
  }
#pragma unroll
  for (int pb = 0; pb < 4; pb++) relu_store(acc[1][pb], out[pb][3], 1);
}

// X[n][9] raw inputs.  MODE 0 (the CRM): Y[n][8] = the 6 de-normalised rates, then 1 where
// the inputs lie inside the training envelope (else 0), then 0.  MODE 1: Y[n][6], the rates.
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_infer(const float* X, long n, MlpView v,
                                                                                       float* Y) {
  __shared__ __attribute__((aligned(16))) bf16 fl[F_SIZE];
  __shared__ __attribute__((aligned(16))) float bl[B_SIZE];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int i = t; i < F_SIZE / 8; i += 256) ((uint4*)fl)[i] = ((const uint4*)v.F)[i];
  for (int i = t; i < NH; i += 256) {
    bl[i] = v.b1[i];
    bl[NH + i] = v.b2[i];
    bl[2 * NH + i] = v.b3[i];
  }
  if (t < NP) bl[3 * NH + t] = v.b4[t];
  __syncthreads();

  const int r = lane & 15, q = lane >> 4;
  const float* nm = v.norm;
  float mean[4], rs[4], lo[4], hi[4], osc[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int f = 4 * q + i;
    const bool ok = f < CRM_MLP_NIN;
    mean[i] = ok ? nm[NORM_MEAN + f] : 0.f;
    rs[i] = ok ? nm[NORM_RSCALE + f] : 1.f;
    lo[i] = ok ? nm[NORM_LO + f] : -INFINITY;
    hi[i] = ok ? nm[NORM_HI + f] : INFINITY;
    osc[i] = f < CRM_MLP_NOUT ? nm[NORM_OSCALE + f] : 0.f;
  }
  const long ntiles = (n + 63) / 64, stride = (long)gridDim.x * 4;
  // raw inputs of the tile: every lane loads its 16 slots unconditionally from clamped (valid)
  // addresses, so the loads issue together; slots past the 9 features or past n are masked
  float xr[4][4];
  auto load_tile = [&](long tile) {
#pragma unroll
    for (int pb = 0; pb < 4; pb++) {
      const long pt = min(tile * 64 + 16 * pb + r, n - 1);
#pragma unroll
      for (int i = 0; i < 4; i++) xr[pb][i] = X[pt * CRM_MLP_NIN + min(4 * q + i, CRM_MLP_NIN - 1)];
    }
  };
  long tile = (long)blockIdx.x * 4 + wave;
  if (tile < ntiles) load_tile(tile);
  for (; tile < ntiles; tile += stride) {
    const long p0 = tile * 64;
    // an offset the compiler cannot see through, so the weight reads stay inside the loop
    // (hoisted out of it, the whole network would be held in registers and spill)
    int fo = 0;
    asm volatile("" : "+v"(fo));
    const bf16* fw = fl + fo;
    s4 x0[4];
    int in[4];
#pragma unroll
    for (int pb = 0; pb < 4; pb++) {
      const bool live = p0 + 16 * pb + r < n;
      float xn[4];
      in[pb] = 1;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const bool have = 4 * q + i < CRM_MLP_NIN && live;
        const float xv = xr[pb][i];
        if (have) in[pb] &= (xv >= lo[i] && xv <= hi[i]);
        xn[i] = have ? (xv - mean[i]) * rs[i] : 0.f;
      }
      const bh2 h0 = __builtin_convertvector((f2){xn[0], xn[1]}, bh2);
      const bh2 h1 = __builtin_convertvector((f2){xn[2], xn[3]}, bh2);
      x0[pb] = __builtin_bit_cast(s4, (u2){__builtin_bit_cast(unsigned, h0), __builtin_bit_cast(unsigned, h1)});
    }
    // the next tile's inputs load under this tile's products
    if (tile + stride < ntiles) load_tile(tile + stride);
    u4 ha[4][4], hb[4][4];
    // layer 1 (K = 16: the 9 inputs, padded)
#pragma unroll
    for (int ob = 0; ob < 8; ob++) {
      const f4 b = *(const f4*)&bl[16 * ob + 4 * q];
      const s4 w = *(const s4*)&fw[F_L1 + (ob * 64 + lane) * 4];
#pragma unroll
      for (int pb = 0; pb < 4; pb++)
        relu_store(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, x0[pb], b, 0, 0, 0), ha[pb][ob >> 1], ob & 1);
    }
    hidden128(fw + F_L2, bl + NH, ha, hb, lane);
    hidden128(fw + F_L3, bl + 2 * NH, hb, ha, lane);
    // layer 4: outputs 4q..4q+3 of point r (q < 2 hold the 6 rates and padding)
    const f4 b4 = *(const f4*)&bl[3 * NH + 4 * q];
    f4 acc[4] = {b4, b4, b4, b4};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const u4 w = *(const u4*)&fw[F_L4 + (kb * 64 + lane) * 8];
#pragma unroll
      for (int pb = 0; pb < 4; pb++) acc[pb] = mfma32(w, ha[pb][kb], acc[pb]);
    }
#pragma unroll
    for (int pb = 0; pb < 4; pb++) {
      const long pt = p0 + 16 * pb + r;
      if constexpr (MODE == 0) {
        // the point's envelope flag: all four lane quarters agree
        int a = in[pb] & __shfl_xor(in[pb], 16);
        a &= __shfl_xor(a, 32);
        if (q < 2 && pt < n) {
          f4 o;
          o[0] = acc[pb][0] * osc[0];
          o[1] = acc[pb][1] * osc[1];
          o[2] = q == 0 ? acc[pb][2] * osc[2] : (a ? 1.f : 0.f);
          o[3] = q == 0 ? acc[pb][3] * osc[3] : 0.f;
          *(f4*)&Y[pt * 8 + 4 * q] = o;
        }
      } else {
        if (q < 2 && pt < n) {
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (4 * q + j < CRM_MLP_NOUT) Y[pt * CRM_MLP_NOUT + 4 * q + j] = acc[pb][j] * osc[j];
        }
      }
    }
  }
}

}  // namespace crm_mlp_ns

using namespace crm_mlp_ns;

struct crm_mlp {
  int device = 0;
  hipStream_t s = nullptr;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // packed FP32
  bf16 *W1 = nullptr, *W2 = nullptr, *W3 = nullptr, *W4 = nullptr;
  bf16* F = nullptr;  // weights in the fragment order of the fused inference
  int ncu = 0;        // compute units of the device (persistent inference grid)
  float *b1 = nullptr, *b2 = nullptr, *b3 = nullptr, *b4 = nullptr;
  float* gpad = nullptr;  // gW1 gW2 gW3 gW4 gb1 gb2 gb3 gb4, contiguous
  float* norm = nullptr;  // [NORM_SIZE]
  float host_norm[NORM_SIZE] = {};
  float* loss = nullptr;
  long cap = 0;  // rows of the work buffers
  bf16 *Xn = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dY = nullptr, *dA = nullptr, *dB = nullptr;
  float* Y = nullptr;
  long step = 0;
};

namespace {
// Persistent grid of the fused inference: two workgroups per CU, fewer for small batches.
template <int MODE>
void launch_infer(hipStream_t s, const crm_mlp* m, const float* X, long n, float* Y) {
  const long tiles = (n + 63) / 64;
  const long grid = std::min<long>((tiles + 3) / 4, 2L * m->ncu);
  if (n > 0) k_infer<MODE><<<dim3((unsigned)grid), dim3(256), 0, s>>>(X, n, mlp_view(m), Y);
}

constexpr size_t GPAD = (size_t)KP * NH + 2 * NH * NH + NH * NP + 3 * NH + NP;

#define MLPCHECK(x)                                   \
  do {                                                \
    if ((x) != hipSuccess) return CRM_ERR_DEVICE;     \
  } while (0)

int ensure_rows(crm_mlp* m, long rows) {
  if (rows <= m->cap) return CRM_OK;
  bf16* bufs[] = {m->Xn, m->H1, m->H2, m->H3, m->dY, m->dA, m->dB};
  for (bf16* b : bufs)
    if (b) (void)hipFree(b);
  if (m->Y) (void)hipFree(m->Y);
  m->cap = 0;
  MLPCHECK(hipMalloc(&m->Xn, rows * KP * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->H1, rows * NH * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->H2, rows * NH * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->H3, rows * NH * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->dY, rows * NP * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->dA, rows * NH * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->dB, rows * NH * sizeof(bf16)));
  MLPCHECK(hipMalloc(&m->Y, rows * NP * sizeof(float)));
  m->cap = rows;
  return CRM_OK;
}

template <int EPI>
void gemm(hipStream_t s, const bf16* A, long sam, long sak, const bf16* B, long sbk, long sbn, int M, int N, int K,
          const float* bias, const bf16* mask, long smm, float* Cf, bf16* Cb, long ldc, int ksplit = 1) {
  const long waves = (long)(M / 16) * (N / 16);
  int kchunk = K;
  if (ksplit > 1) kchunk = ((K / ksplit + 15) / 16) * 16;
  const int ny = (K + kchunk - 1) / kchunk;
  dim3 grid((unsigned)((waves + 3) / 4), (unsigned)ny);
  k_gemm<EPI><<<grid, dim3(256), 0, s>>>(A, sam, sak, B, sbk, sbn, M, N, K, kchunk, bias, mask, smm, Cf, Cb, ldc);
}

// forward pass of mp (padded) rows already in Xn; leaves H1..H3 and the normalised Y
void forward_rows(crm_mlp* m, int mp) {
  hipStream_t s = m->s;
  gemm<EPI_BIAS_RELU_BF16>(s, m->Xn, KP, 1, m->W1, NH, 1, mp, NH, KP, m->b1, nullptr, 0, nullptr, m->H1, NH);
  gemm<EPI_BIAS_RELU_BF16>(s, m->H1, NH, 1, m->W2, NH, 1, mp, NH, NH, m->b2, nullptr, 0, nullptr, m->H2, NH);
  gemm<EPI_BIAS_RELU_BF16>(s, m->H2, NH, 1, m->W3, NH, 1, mp, NH, NH, m->b3, nullptr, 0, nullptr, m->H3, NH);
  gemm<EPI_BIAS_F32>(s, m->H3, NH, 1, m->W4, NP, 1, mp, NP, NH, m->b4, nullptr, 0, m->Y, nullptr, NP);
}

// loss and gradients into gpad for n rows of X/T (device)
int backward(crm_mlp* m, const float* X, const float* T, int n) {
  hipStream_t s = m->s;
  const int mp = (n + 15) / 16 * 16;
  int rc = ensure_rows(m, mp);
  if (rc) return rc;
  k_inputs<<<dim3((unsigned)(((long)mp * KP + 255) / 256)), dim3(256), 0, s>>>(X, n, mp, m->norm, m->Xn);
  forward_rows(m, mp);
  MLPCHECK(hipMemsetAsync(m->loss, 0, sizeof(float), s));
  MLPCHECK(hipMemsetAsync(m->gpad, 0, GPAD * sizeof(float), s));
  k_loss<<<dim3((unsigned)std::min<long>(((long)mp * NP + 255) / 256, 256)), dim3(256), 0, s>>>(m->Y, T, n, mp, m->norm,
                                                                                               m->dY, m->loss);
  float* gW1 = m->gpad;
  float* gW2 = gW1 + KP * NH;
  float* gW3 = gW2 + NH * NH;
  float* gW4 = gW3 + NH * NH;
  float* gb1 = gW4 + NH * NP;
  float* gb2 = gb1 + NH;
  float* gb3 = gb2 + NH;
  float* gb4 = gb3 + NH;
  const int ks = std::max(1, mp / 256);  // split the batch reduction of the weight gradients
  const dim3 cs(2, 64);
  // layer 4
  gemm<EPI_ATOMIC_F32>(s, m->H3, 1, NH, m->dY, NP, 1, NH, NP, mp, nullptr, nullptr, 0, gW4, nullptr, NP, ks);
  k_colsum<<<dim3(1, 64), dim3(256), 0, s>>>(m->dY, mp, NP, NP, gb4);
  gemm<EPI_MASK_BF16>(s, m->dY, NP, 1, m->W4, 1, NP, mp, NH, NP, nullptr, m->H3, NH, nullptr, m->dA, NH);
  // layer 3
  gemm<EPI_ATOMIC_F32>(s, m->H2, 1, NH, m->dA, NH, 1, NH, NH, mp, nullptr, nullptr, 0, gW3, nullptr, NH, ks);
  k_colsum<<<cs, dim3(256), 0, s>>>(m->dA, mp, NH, NH, gb3);
  gemm<EPI_MASK_BF16>(s, m->dA, NH, 1, m->W3, 1, NH, mp, NH, NH, nullptr, m->H2, NH, nullptr, m->dB, NH);
  // layer 2
  gemm<EPI_ATOMIC_F32>(s, m->H1, 1, NH, m->dB, NH, 1, NH, NH, mp, nullptr, nullptr, 0, gW2, nullptr, NH, ks);
  k_colsum<<<cs, dim3(256), 0, s>>>(m->dB, mp, NH, NH, gb2);
  gemm<EPI_MASK_BF16>(s, m->dB, NH, 1, m->W2, 1, NH, mp, NH, NH, nullptr, m->H1, NH, nullptr, m->dA, NH);
  // layer 1
  gemm<EPI_ATOMIC_F32>(s, m->Xn, 1, KP, m->dA, NH, 1, KP, NH, mp, nullptr, nullptr, 0, gW1, nullptr, NH, ks);
  k_colsum<<<cs, dim3(256), 0, s>>>(m->dA, mp, NH, NH, gb1);
  k_pack_grad<<<dim3((CRM_MLP_NPARAM + 255) / 256), dim3(256), 0, s>>>(gW1, gW2, gW3, gW4, gb1, gb2, gb3, gb4, m->g);
  MLPCHECK(hipGetLastError());
  return CRM_OK;
}

void unpack(crm_mlp* m) {
  k_unpack<<<dim3((NH * NH + 255) / 256), dim3(256), 0, m->s>>>(m->p, m->W1, m->W2, m->W3, m->W4, m->b1, m->b2, m->b3,
                                                                 m->b4);
  k_frags<<<dim3((F_SIZE + 255) / 256), dim3(256), 0, m->s>>>(m->p, m->F);
}
}  // namespace

extern "C" {

int crm_mlp_create(crm_mlp** out, int device) {
  if (!out) return CRM_ERR_ARG;
  *out = nullptr;
  auto* m = new crm_mlp();
  m->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&m->s) != hipSuccess) {
    delete m;
    return CRM_ERR_DEVICE;
  }
  bool ok = hipMalloc(&m->p, CRM_MLP_NPARAM * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->g, CRM_MLP_NPARAM * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->m, CRM_MLP_NPARAM * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->v, CRM_MLP_NPARAM * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->W1, KP * NH * sizeof(bf16)) == hipSuccess &&
            hipMalloc(&m->W2, NH * NH * sizeof(bf16)) == hipSuccess &&
            hipMalloc(&m->W3, NH * NH * sizeof(bf16)) == hipSuccess &&
            hipMalloc(&m->W4, NH * NP * sizeof(bf16)) == hipSuccess &&
            hipMalloc(&m->F, F_SIZE * sizeof(bf16)) == hipSuccess &&
            hipMalloc(&m->b1, NH * sizeof(float)) == hipSuccess && hipMalloc(&m->b2, NH * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->b3, NH * sizeof(float)) == hipSuccess && hipMalloc(&m->b4, NP * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->gpad, GPAD * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->norm, NORM_SIZE * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->loss, sizeof(float)) == hipSuccess;
  if (!ok) {
    crm_mlp_destroy(m);
    return CRM_ERR_NOMEM;
  }
  if (hipDeviceGetAttribute(&m->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || m->ncu < 1) {
    crm_mlp_destroy(m);
    return CRM_ERR_DEVICE;
  }
  (void)hipMemset(m->p, 0, CRM_MLP_NPARAM * sizeof(float));
  (void)hipMemset(m->m, 0, CRM_MLP_NPARAM * sizeof(float));
  (void)hipMemset(m->v, 0, CRM_MLP_NPARAM * sizeof(float));
  for (int c = 0; c < 16; c++) {
    m->host_norm[NORM_SCALE + c] = 1.f;
    m->host_norm[NORM_RSCALE + c] = 1.f;
    m->host_norm[NORM_OSCALE + c] = 1.f;
    m->host_norm[NORM_LO + c] = -INFINITY;
    m->host_norm[NORM_HI + c] = INFINITY;
  }
  (void)hipMemcpy(m->norm, m->host_norm, sizeof(m->host_norm), hipMemcpyHostToDevice);
  unpack(m);
  (void)hipStreamSynchronize(m->s);
  *out = m;
  return CRM_OK;
}

int crm_mlp_destroy(crm_mlp* m) {
  if (!m) return CRM_ERR_ARG;
  (void)hipSetDevice(m->device);
  void* bufs[] = {m->p, m->g, m->m, m->v, m->W1, m->W2, m->W3, m->W4, m->F, m->b1, m->b2, m->b3, m->b4, m->gpad,
                  m->norm, m->loss, m->Xn, m->H1, m->H2, m->H3, m->dY, m->dA, m->dB, m->Y};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (m->s) (void)hipStreamDestroy(m->s);
  delete m;
  return CRM_OK;
}

int crm_mlp_set_params(crm_mlp* m, const float* params) {
  if (!m || !params) return CRM_ERR_ARG;
  MLPCHECK(hipMemcpyAsync(m->p, params, CRM_MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, m->s));
  MLPCHECK(hipMemsetAsync(m->m, 0, CRM_MLP_NPARAM * sizeof(float), m->s));
  MLPCHECK(hipMemsetAsync(m->v, 0, CRM_MLP_NPARAM * sizeof(float), m->s));
  m->step = 0;
  unpack(m);
  MLPCHECK(hipStreamSynchronize(m->s));
  return CRM_OK;
}

int crm_mlp_get_params(const crm_mlp* m, float* params) {
  if (!m || !params) return CRM_ERR_ARG;
  MLPCHECK(hipMemcpyAsync(params, m->p, CRM_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->s));
  MLPCHECK(hipStreamSynchronize(m->s));
  return CRM_OK;
}

int crm_mlp_set_norm(crm_mlp* m, const float* mean, const float* scale, const float* oscale, const float* lo,
                     const float* hi) {
  if (!m) return CRM_ERR_ARG;
  for (int c = 0; c < CRM_MLP_NIN; c++) {
    if (mean) m->host_norm[NORM_MEAN + c] = mean[c];
    if (scale) {
      if (!(scale[c] > 0.f)) return CRM_ERR_ARG;
      m->host_norm[NORM_SCALE + c] = scale[c];
      m->host_norm[NORM_RSCALE + c] = 1.f / scale[c];
    }
    if (lo) m->host_norm[NORM_LO + c] = lo[c];
    if (hi) m->host_norm[NORM_HI + c] = hi[c];
  }
  for (int c = 0; c < CRM_MLP_NOUT; c++)
    if (oscale) {
      if (!(oscale[c] > 0.f)) return CRM_ERR_ARG;
      m->host_norm[NORM_OSCALE + c] = oscale[c];
    }
  MLPCHECK(hipMemcpyAsync(m->norm, m->host_norm, sizeof(m->host_norm), hipMemcpyHostToDevice, m->s));
  MLPCHECK(hipStreamSynchronize(m->s));
  return CRM_OK;
}

int crm_mlp_forward(crm_mlp* m, const float* X, int n, float* Y) {
  if (!m || !X || !Y || n < 0) return CRM_ERR_ARG;
  launch_infer<1>(m->s, m, X, n, Y);
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipStreamSynchronize(m->s));
  return CRM_OK;
}

int crm_mlp_grad(crm_mlp* m, const float* X, const float* T, int n, float* grad, float* loss) {
  if (!m || !X || !T || n < 1) return CRM_ERR_ARG;
  int rc = backward(m, X, T, n);
  if (rc) return rc;
  if (grad) MLPCHECK(hipMemcpyAsync(grad, m->g, CRM_MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->s));
  if (loss) MLPCHECK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->s));
  MLPCHECK(hipStreamSynchronize(m->s));
  return CRM_OK;
}

int crm_mlp_train_step(crm_mlp* m, const float* X, const float* T, int n, float lr, float* loss) {
  if (!m || !X || !T || n < 1 || !(lr > 0.f)) return CRM_ERR_ARG;
  int rc = backward(m, X, T, n);
  if (rc) return rc;
  m->step++;
  const float c1 = 1.f - std::pow(0.9f, (float)m->step), c2 = 1.f - std::pow(0.999f, (float)m->step);
  k_adam<<<dim3((CRM_MLP_NPARAM + 255) / 256), dim3(256), 0, m->s>>>(m->p, m->g, m->m, m->v, lr, c1, c2);
  unpack(m);
  if (loss) MLPCHECK(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->s));
  MLPCHECK(hipGetLastError());
  MLPCHECK(hipStreamSynchronize(m->s));
  return CRM_OK;
}

}  // extern "C"

// for the CRM engines (crm_driver.hpp): the device-side view of a network
namespace crm_mlp_ns {
MlpView mlp_view(const crm_mlp* m) {
  MlpView v;
  v.W1 = m->W1;
  v.W2 = m->W2;
  v.W3 = m->W3;
  v.W4 = m->W4;
  v.F = m->F;
  v.b1 = m->b1;
  v.b2 = m->b2;
  v.b3 = m->b3;
  v.b4 = m->b4;
  v.norm = m->norm;
  return v;
}
int mlp_device(const crm_mlp* m) { return m->device; }
void mlp_infer(hipStream_t s, const crm_mlp* m, const float* X, long n, float* Y) { launch_infer<0>(s, m, X, n, Y); }
}  // namespace crm_mlp_ns
