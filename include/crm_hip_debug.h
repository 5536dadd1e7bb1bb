/*
 * crm_hip_debug.h -- inspection and timing entry points of libcrm_hip.so (parity tests and
 * the benchmark).
 *
 * They have no counterpart in the reference's crm() interface; they expose intermediate
 * states of the same computation so that a mismatch can be traced to one phase of the
 * substep loop of crm_module.F90:727-900.
 */
#ifndef CRM_HIP_DEBUG_H
#define CRM_HIP_DEBUG_H

#include "crm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* crm_hip_run, stopping inside the last substep after a phase group:
 *   7  zeroing, buoyancy, forcing, radiative heating, damping, ice sedimentation,
 *      boundaries and surface fluxes          (crm_module.F90:751-805)
 *   9  sgs_proc                               (crm_module.F90:809-813)
 *   11 advect_mom + sgs_mom                   (crm_module.F90:817-821)
 *   13 pressure + adams                       (crm_module.F90:829-838)
 *   15 scalar advection and sedimentation     (crm_module.F90:842)
 *   17 uvw                                    (crm_module.F90:846-850)
 *   18 sgs_scalars                            (crm_module.F90:854)
 *   19 micro_proc, only with mean-state acceleration on: before accelerate_crm
 *                                             (crm_module.F90:865-870)
 *   20 micro_proc + accelerate_crm + diagnose (crm_module.F90:865-880)
 * 0 runs the whole call. Views are host pointers as for crm_hip_run. */
int crm_hip_debug_run(crm_hip_ctx *ctx, const crm_dims *dims, const crm_column_info *col,
                      double dt_gl, const crm_input_view *in, crm_state_view *state,
                      crm_rad_view *rad, crm_output_view *out, int phase_stop);

/* Copy an internal field of the last debug run to buf: grid fields as [c][k][i] (u, w, t,
 * tabs, qv, qcl, qci, qpl, qpi, qt, qp, qn, tk, tkh, tke, dudt, dwdt, precsfc, ...), level
 * profiles as [c][k]. Returns the number of values written, -1 for an unknown name. */
int crm_hip_debug_get(crm_hip_ctx *ctx, const char *name, double *buf);

/* Dense operators of the reference's periodic x-transform (crm/fft.F90:5-153 as called by
 * crm/pressure.F90:117-140): forward Ft[i][s] and inverse Gt[s][i], n x n row-major, slot
 * s = 0 mean, 1 Nyquist cosine, 2m / 2m+1 cosine / sine of mode m. Host-only. */
int crm_hip_fft991_operators(int n, double *Ft, double *Gt);

/* Kernel classes of one call, in launch order. */
enum {
  CRM_K_SETUP = 0,   /* reference profiles            crm_module.F90:320-375           */
  CRM_K_INIT,        /* state in, means, tendencies   crm_module.F90:379-594           */
  CRM_K_SETPERTURB,  /* first-step noise              setperturb.F90:7-86              */
  CRM_K_KURANT,      /* CFL maximum                   kurant.F90:9-83                  */
  CRM_K_PRECIP_INIT, /* microphysics coefficients     precip_init.F90:8-97             */
  CRM_K_PRE,         /* buoyancy..ice_fall            crm_module.F90:751-805           */
  CRM_K_SGS,         /* sgs_proc                      tke_full.F90:11-317              */
  CRM_K_MOM,         /* advect_mom + sgs_mom          crm_module.F90:817-821           */
  CRM_K_PRESS,       /* pressure + adams              crm_module.F90:829-834           */
  CRM_K_MPDATA,      /* advect_all_scalars            advect_all_scalars.F90:7-94      */
  CRM_K_PRECIP,      /* precip_fall                   microphysics.F90:529-833         */
  CRM_K_DIFFUSE,     /* sgs_scalars                   sgs.F90:389-451                  */
  CRM_K_MICRO,       /* micro_proc + diagnose         crm_module.F90:865-880           */
  CRM_K_STATS,       /* per-substep statistics        crm_module.F90:896-1073          */
  CRM_K_FINAL,       /* outputs                       crm_module.F90:1078-1659         */
  CRM_K_COUNT
};

/* Time every launch of the kernel classes in `mask` (bit CRM_K_x) with a HIP event pair on
 * the context's stream, and reset the totals. 0 turns timing off. */
int crm_hip_profile_kernels(crm_hip_ctx *ctx, unsigned mask);

/* Total device time [ms] and launch count of kernel class kid since the last reset. */
int crm_hip_kernel_time(const crm_hip_ctx *ctx, int kid, double *ms_total, long long *launches);

const char *crm_hip_kernel_name(int kid);

/* Number of substeps replayed from captured graphs since the context was created. */
long long crm_hip_graph_launches(const crm_hip_ctx *ctx);

/* Number of captured graphs the context holds (at most 32; least recently used first out). */
int crm_hip_graph_cache_size(const crm_hip_ctx *ctx);

/* Test hook: a captured call (CRM_GRAPHS_CALL) stops at substep nstep (1-based) as if it needed
 * more subcycles, and the host resumes it there; 0 turns the hook off. */
int crm_hip_debug_stop_call_at(crm_hip_ctx *ctx, int nstep);

#ifdef __cplusplus
}
#endif
#endif /* CRM_HIP_DEBUG_H */
