// crm_mlp_dev.hpp -- shared definitions of the microphysics network (crm_mlp.hip) for its own
// kernels and for the CRM engines that evaluate it inside the microphysics sweep.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/crm_mlp.h"

namespace crm_mlp_ns {

typedef unsigned short bf16;
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int KP = 16;   // inputs padded to 16 columns
constexpr int NP = 16;   // outputs padded to 16 columns
constexpr int NH = CRM_MLP_NH;

// packed parameter offsets (crm_mlp.h)
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + CRM_MLP_NIN * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NH * CRM_MLP_NOUT;
static_assert(OFF_B4 + CRM_MLP_NOUT == CRM_MLP_NPARAM, "packed parameter count");

// normalisation block [NORM_SIZE]: mean, scale, oscale, lo, hi, 1 / scale (16 entries each)
enum { NORM_MEAN = 0, NORM_SCALE = 16, NORM_OSCALE = 32, NORM_LO = 48, NORM_HI = 64, NORM_RSCALE = 80, NORM_SIZE = 96 };

__device__ __forceinline__ bf16 to_bf16(float x) {  // round to nearest, ties to even
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float((unsigned)h << 16); }

// Device-side view of a network: padded BF16 weights W1[16][128], W2, W3 [128][128],
// W4[128][16], the same weights in the fragment order of the fused inference (F), FP32 biases
// (b4 padded to 16), the normalisation block.
struct MlpView {
  const bf16 *W1 = nullptr, *W2 = nullptr, *W3 = nullptr, *W4 = nullptr;
  const bf16* F = nullptr;
  const float *b1 = nullptr, *b2 = nullptr, *b3 = nullptr, *b4 = nullptr;
  const float* norm = nullptr;
};

MlpView mlp_view(const crm_mlp* m);
int mlp_device(const crm_mlp* m);
// Fused inference of the whole network on stream s: X[n][9] raw inputs -> Y[n][8] (6 rates,
// in-envelope flag, 0).
void mlp_infer(hipStream_t s, const crm_mlp* m, const float* X, long n, float* Y);

}  // namespace crm_mlp_ns
