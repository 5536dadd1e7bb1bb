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
#pragma unroll
    for (int j = 0; j < 4; j++) {
      av[j] = (short)a[(long)(k + 4 * q + j) * sak];
      bv[j] = (short)b[(long)(k + 4 * q + j) * sbk];
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
  if (m < n && c < CRM_MLP_NIN) v = (X[m * CRM_MLP_NIN + c] - norm[NORM_MEAN + c]) / norm[NORM_SCALE + c];
  Xn[idx] = to_bf16(v);
}

// Loss of the normalised outputs and its gradient: dY = 2 (Y - T / oscale) / (6 n) in BF16.
__global__ void k_loss(const float* Y, const float* T, int n, int mp, const float* norm, bf16* dY, float* loss) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (idx < (long)mp * NP) {
    const long m = idx / NP;
    const int c = (int)(idx % NP);
    float g = 0.f;
    if (m < n && c < CRM_MLP_NOUT) {
      const float d = Y[idx] - T[m * CRM_MLP_NOUT + c] / norm[NORM_OSCALE + c];
      e2 = d * d;
      g = 2.f * d / (6.f * (float)n);
    }
    dY[idx] = to_bf16(g);
  }
  // one atomic per wave
  for (int o = 32; o >= 1; o >>= 1) e2 += __shfl_xor(e2, o, 64);
  if ((threadIdx.x & 63) == 0 && e2 != 0.f) unsafeAtomicAdd(loss, e2 / (6.f * (float)n));
}

// Column sums of a BF16 [mp][ld] matrix over its rows, accumulated into out[ncol] (FP32).
__global__ void k_colsum(const bf16* G, int mp, int ld, int ncol, float* out) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int part = blockIdx.y, nparts = gridDim.y;
  if (c >= ncol) return;
  float sacc = 0.f;
  for (long m = part; m < mp; m += nparts) sacc += from_bf16(G[m * ld + c]);
  unsafeAtomicAdd(&out[c], sacc);
}

// De-normalised outputs: Y[n][6] = Yn[m][0..5] * oscale.
__global__ void k_outputs(const float* Yn, int n, const float* norm, float* Y) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * CRM_MLP_NOUT) return;
  const long m = idx / CRM_MLP_NOUT;
  const int c = (int)(idx % CRM_MLP_NOUT);
  Y[idx] = Yn[m * NP + c] * norm[NORM_OSCALE + c];
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

// Inference for the CRM: one workgroup (4 waves) per 64 grid points, the whole network in one
// pass with the activations in LDS.  X[n][9] raw inputs; Y[n][8]: the 6 de-normalised rates,
// then 1 where the inputs lie inside the training envelope (else 0), then 0.
constexpr int HLD = NH + 8;  // LDS row stride of the activations (bf16)

__device__ __forceinline__ void infer_layer(const bf16 (*A)[HLD], int kdim, const bf16* W, const float* b,
                                            bf16 (*H)[HLD], int wave, int lane) {
  const int r = lane & 15, q = lane >> 4;
  for (int tile = wave * 8; tile < wave * 8 + 8; tile++) {
    const int rb = tile >> 3, cb = tile & 7;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < kdim; kk += 16) {
      s4 av, bv;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        av[j] = (short)A[16 * rb + r][kk + 4 * q + j];
        bv[j] = (short)W[(kk + 4 * q + j) * NH + 16 * cb + r];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) H[16 * rb + 4 * q + j][16 * cb + r] = to_bf16(fmaxf(acc[j] + b[16 * cb + r], 0.f));
  }
}

__global__ void __launch_bounds__(256) k_infer(const float* X, long n, MlpView v, float* Y) {
  __shared__ bf16 xs[64][HLD];
  __shared__ bf16 ha[64][HLD];
  __shared__ bf16 hb[64][HLD];
  const long p0 = (long)blockIdx.x * 64;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const float* nm = v.norm;
  if (t < 64) {
    const long p = p0 + t;
    bool in = p < n;
#pragma unroll
    for (int c = 0; c < KP; c++) {
      float f = 0.f;
      if (c < CRM_MLP_NIN && p < n) {
        const float x = X[p * CRM_MLP_NIN + c];
        in = in && x >= nm[NORM_LO + c] && x <= nm[NORM_HI + c];
        f = (x - nm[NORM_MEAN + c]) / nm[NORM_SCALE + c];
      }
      xs[t][c] = to_bf16(f);
    }
    if (p < n) {
      Y[p * 8 + 6] = in ? 1.f : 0.f;
      Y[p * 8 + 7] = 0.f;
    }
  }
  __syncthreads();
  infer_layer(xs, KP, v.W1, v.b1, ha, wave, lane);
  __syncthreads();
  infer_layer(ha, NH, v.W2, v.b2, hb, wave, lane);
  __syncthreads();
  infer_layer(hb, NH, v.W3, v.b3, ha, wave, lane);
  __syncthreads();
  {
    // output layer: wave w computes rows 16w..16w+15
    const int r = lane & 15, q = lane >> 4, rb = wave;
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < NH; kk += 16) {
      s4 av, bv;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        av[j] = (short)ha[16 * rb + r][kk + 4 * q + j];
        bv[j] = (short)v.W4[(kk + 4 * q + j) * NP + r];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const long p = p0 + 16 * rb + 4 * q + j;
      if (r < CRM_MLP_NOUT && p < n) Y[p * 8 + r] = (acc[j] + v.b4[r]) * nm[NORM_OSCALE + r];
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
  k_loss<<<dim3((unsigned)(((long)mp * NP + 255) / 256)), dim3(256), 0, s>>>(m->Y, T, n, mp, m->norm, m->dY, m->loss);
  float* gW1 = m->gpad;
  float* gW2 = gW1 + KP * NH;
  float* gW3 = gW2 + NH * NH;
  float* gW4 = gW3 + NH * NH;
  float* gb1 = gW4 + NH * NP;
  float* gb2 = gb1 + NH;
  float* gb3 = gb2 + NH;
  float* gb4 = gb3 + NH;
  const int ks = std::max(1, mp / 2048);  // split the batch reduction of the weight gradients
  const dim3 cs(2, 64);
  // layer 4
  gemm<EPI_ATOMIC_F32>(s, m->H3, 1, NH, m->dY, NP, 1, NH, NP, mp, nullptr, nullptr, 0, gW4, nullptr, NP, ks);
  k_colsum<<<dim3(1, 64), dim3(64), 0, s>>>(m->dY, mp, NP, NP, gb4);
  gemm<EPI_MASK_BF16>(s, m->dY, NP, 1, m->W4, 1, NP, mp, NH, NP, nullptr, m->H3, NH, nullptr, m->dA, NH);
  // layer 3
  gemm<EPI_ATOMIC_F32>(s, m->H2, 1, NH, m->dA, NH, 1, NH, NH, mp, nullptr, nullptr, 0, gW3, nullptr, NH, ks);
  k_colsum<<<cs, dim3(64), 0, s>>>(m->dA, mp, NH, NH, gb3);
  gemm<EPI_MASK_BF16>(s, m->dA, NH, 1, m->W3, 1, NH, mp, NH, NH, nullptr, m->H2, NH, nullptr, m->dB, NH);
  // layer 2
  gemm<EPI_ATOMIC_F32>(s, m->H1, 1, NH, m->dB, NH, 1, NH, NH, mp, nullptr, nullptr, 0, gW2, nullptr, NH, ks);
  k_colsum<<<cs, dim3(64), 0, s>>>(m->dB, mp, NH, NH, gb2);
  gemm<EPI_MASK_BF16>(s, m->dB, NH, 1, m->W2, 1, NH, mp, NH, NH, nullptr, m->H1, NH, nullptr, m->dA, NH);
  // layer 1
  gemm<EPI_ATOMIC_F32>(s, m->Xn, 1, KP, m->dA, NH, 1, KP, NH, mp, nullptr, nullptr, 0, gW1, nullptr, NH, ks);
  k_colsum<<<cs, dim3(64), 0, s>>>(m->dA, mp, NH, NH, gb1);
  k_pack_grad<<<dim3((CRM_MLP_NPARAM + 255) / 256), dim3(256), 0, s>>>(gW1, gW2, gW3, gW4, gb1, gb2, gb3, gb4, m->g);
  MLPCHECK(hipGetLastError());
  return CRM_OK;
}

void unpack(crm_mlp* m) {
  k_unpack<<<dim3((NH * NH + 255) / 256), dim3(256), 0, m->s>>>(m->p, m->W1, m->W2, m->W3, m->W4, m->b1, m->b2, m->b3,
                                                                 m->b4);
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
            hipMalloc(&m->b1, NH * sizeof(float)) == hipSuccess && hipMalloc(&m->b2, NH * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->b3, NH * sizeof(float)) == hipSuccess && hipMalloc(&m->b4, NP * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->gpad, GPAD * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->norm, NORM_SIZE * sizeof(float)) == hipSuccess &&
            hipMalloc(&m->loss, sizeof(float)) == hipSuccess;
  if (!ok) {
    crm_mlp_destroy(m);
    return CRM_ERR_NOMEM;
  }
  (void)hipMemset(m->p, 0, CRM_MLP_NPARAM * sizeof(float));
  (void)hipMemset(m->m, 0, CRM_MLP_NPARAM * sizeof(float));
  (void)hipMemset(m->v, 0, CRM_MLP_NPARAM * sizeof(float));
  for (int c = 0; c < 16; c++) {
    m->host_norm[NORM_SCALE + c] = 1.f;
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
  void* bufs[] = {m->p, m->g, m->m, m->v, m->W1, m->W2, m->W3, m->W4, m->b1, m->b2, m->b3, m->b4, m->gpad,
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
  const int chunk = 1 << 20;
  for (int r0 = 0; r0 < n; r0 += chunk) {
    const int nr = std::min(chunk, n - r0);
    const int mp = (nr + 15) / 16 * 16;
    int rc = ensure_rows(m, mp);
    if (rc) return rc;
    k_inputs<<<dim3((unsigned)(((long)mp * KP + 255) / 256)), dim3(256), 0, m->s>>>(X + (long)r0 * CRM_MLP_NIN, nr, mp,
                                                                                  m->norm, m->Xn);
    forward_rows(m, mp);
    k_outputs<<<dim3((unsigned)(((long)nr * CRM_MLP_NOUT + 255) / 256)), dim3(256), 0, m->s>>>(
        m->Y, nr, m->norm, Y + (long)r0 * CRM_MLP_NOUT);
  }
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
  v.b1 = m->b1;
  v.b2 = m->b2;
  v.b3 = m->b3;
  v.b4 = m->b4;
  v.norm = m->norm;
  return v;
}
int mlp_device(const crm_mlp* m) { return m->device; }
void mlp_infer(hipStream_t s, const crm_mlp* m, const float* X, long n, float* Y) {
  if (n > 0) k_infer<<<dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s>>>(X, n, mlp_view(m), Y);
}
}  // namespace crm_mlp_ns
