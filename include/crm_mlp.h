/*
 * crm_mlp.h -- the learned stand-in for the CRM's bulk microphysics (libcrm_hip.so).
 *
 * No counterpart in the reference: a pointwise multilayer perceptron 9 -> 128 -> 128 -> 128 -> 6
 * (ReLU) that maps the state of one grid point at the entry of micro_proc
 * (MICRO_SAM1MOM/microphysics.F90:292-357) to the rates of change over that call:
 *   inputs  T, p, rho, qv, qcl, qci, qpl, qpi, w           (K, hPa, kg/m3, kg/kg x5, m/s)
 *   outputs d/dt of qv, qcl, qci, qpl, qpi, and of T        (1/s, K/s)
 * BF16 weights and activations on the matrix cores with FP32 accumulation; FP32 master
 * weights and Adam moments for training.  Inputs are normalised as (x - mean) * (1 / scale)
 * (the reciprocal rounded to FP32 once) and
 * outputs de-normalised by oscale (crm_mlp_set_norm); lo/hi bound the training envelope.
 *
 * All array arguments of the compute entry points are device pointers.
 */
#ifndef CRM_MLP_H
#define CRM_MLP_H

#include "crm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { CRM_MLP_NIN = 9, CRM_MLP_NH = 128, CRM_MLP_NOUT = 6 };

/* Packed parameter vector: W1[9][128], b1[128], W2[128][128], b2[128], W3[128][128], b3[128],
 * W4[128][6], b4[6] (row-major, y = x W + b). */
#define CRM_MLP_NPARAM (9 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 6 + 6)

typedef struct crm_mlp crm_mlp;

int crm_mlp_create(crm_mlp **mlp, int device);
int crm_mlp_destroy(crm_mlp *mlp);

/* Host FP32 copies of the packed parameters. */
int crm_mlp_set_params(crm_mlp *mlp, const float *params);
int crm_mlp_get_params(const crm_mlp *mlp, float *params);

/* Normalisation (host arrays): mean[9], scale[9], oscale[6]; training envelope lo[9], hi[9]. */
int crm_mlp_set_norm(crm_mlp *mlp, const float *mean, const float *scale, const float *oscale,
                     const float *lo, const float *hi);

/* Inference: X[n][9] raw inputs -> Y[n][6] de-normalised rates. */
int crm_mlp_forward(crm_mlp *mlp, const float *X, int n, float *Y);

/* Gradient of the loss L = mean over the n x 6 outputs of (y' - t')^2, with y' the network's
 * normalised output and t' = T / oscale, w.r.t. the packed parameters (host grad, may be
 * NULL), and L (host, may be NULL). X[n][9], T[n][6]. */
int crm_mlp_grad(crm_mlp *mlp, const float *X, const float *T, int n, float *grad, float *loss);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on that loss; loss (host, may be NULL)
 * is the value before the step. */
int crm_mlp_train_step(crm_mlp *mlp, const float *X, const float *T, int n, float lr, float *loss);

/* Use the network in place of micro_proc's saturation adjustment and precipitation
 * processes in later calls of a CRM context (mlp NULL: the plain microphysics).  Points whose
 * inputs leave the training envelope use the plain microphysics; every point keeps total water
 * exactly and condensate non-negative (see the kernel's fixer). */
int crm_hip_set_micro_mlp(crm_hip_ctx *ctx, crm_mlp *mlp);

/* Record (input, target) pairs of the plain microphysics during later calls of a CRM context:
 * every stride-th grid point of every subcycle, up to cap pairs, into X[cap][9] and T[cap][6]
 * (device pointers; NULL X stops recording). *count (host) receives the number recorded so
 * far when read with crm_hip_micro_samples_count. */
int crm_hip_record_micro_samples(crm_hip_ctx *ctx, float *X, float *T, long long cap, int stride);
long long crm_hip_micro_samples_count(const crm_hip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* CRM_MLP_H */
