/*
 * crm_hip.h -- C ABI of the MI355X (gfx950) cloud-resolving model library.
 *
 * This library replaces one Fortran entry point of the reference:
 *
 *   subroutine crm(lchnk, icol, ncrms, dt_gl, plev,
 *                  crm_input, crm_state, crm_rad, crm_ecpp_output, crm_output)
 *     components/cam/src/physics/crm/crm_module.F90:45-56
 *
 * called by crm_physics_tend (components/cam/src/physics/cam/crm_physics.F90:1145-1156).
 * The Fortran types it exchanges keep their layouts; here each becomes a "view": a struct
 * of plain pointers into the caller's arrays plus the caller's leading dimension.
 *
 * Layout convention (identical to the Fortran arrays, column-major, pcols leading):
 *   (pcols)                   element icrm              at p[icrm]
 *   (pcols, plev)             element (icrm,k)          at p[icrm + ld*k]
 *   (pcols, plev+1)           element (icrm,k)          at p[icrm + ld*k]
 *   (pcols, nx, ny)           element (icrm,i,j)        at p[icrm + ld*(i + nx*j)]
 *   (pcols, nx, ny, nz)       element (icrm,i,j,k)      at p[icrm + ld*(i + nx*(j + ny*k))]
 *   (pcols, nxr, nyr, nz)     element (icrm,ir,jr,k)    at p[icrm + ld*(ir + nxr*(jr + nyr*k))]
 * All indices above are 0-based; icrm < ncrms <= ld. GCM level k=0 is the model top,
 * k=plev-1 the lowest level (CAM ordering); CRM level 0 is the lowest CRM level.
 *
 * Only instances 0..ncrms-1 are read or written (crm_module.F90:268-281 slices 1:ncrms).
 * The caller owns all memory.
 */
#ifndef CRM_HIP_H
#define CRM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference aborts the whole model (endrun / task_abort / stop); this
 * library returns a code instead and the Fortran shim turns a nonzero code into endrun. */
enum {
  CRM_OK = 0,
  CRM_ERR_ARG = 1,        /* bad dims or null pointer                                       */
  CRM_ERR_CFL = 2,        /* ncycle > 16 (kurant.F90:64-75)                                 */
  CRM_ERR_NXRAD = 3,      /* nx not divisible by nx_rad (crm_module.F90:686-692)            */
  CRM_ERR_DAMPING = 4,    /* tau_min < 2*dt in the sponge (damping.F90:39-42)               */
  CRM_ERR_NAN = 5,        /* non-finite prognostic field detected after the call            */
  CRM_ERR_DEVICE = 6,     /* HIP runtime error                                              */
  CRM_ERR_NOMEM = 7,      /* device allocation failed                                       */
  CRM_ERR_UNSUPPORTED = 8, /* configuration not provided by this build                      */
  CRM_ERR_ACCEL = 9,      /* 1 + crm_accel_factor does not divide nstop (accelerate_crm.F90:78-106) */
  CRM_ERR_SCRATCH = 10    /* batch needs more than CRM_HIP_NOALLOC_SCRATCH_BYTES (noalloc entry) */
};

/* Arithmetic of the bulk stencils. The pressure solve (DFT + Thomas sweep), the hydrostatic
 * reference profiles and the column budget sums stay FP64 in both modes. */
enum { CRM_PREC_FP64 = 0, CRM_PREC_MIXED = 1 };

typedef struct crm_dims {
  int ncrms;        /* instances in this call                         (crm_module.F90:103)  */
  int ld;           /* leading dimension of every caller array (pcols, crm_module.F90:60)   */
  int plev;         /* GCM levels (pver)                              (crm_module.F90:104)  */
  int nx, ny, nz;   /* crm_nx, crm_ny, crm_nz  (cam/crmdims.F90:17-19); ny==1 means 2D      */
  int nx_rad, ny_rad; /* crm_nx_rad, crm_ny_rad (cam/crmdims.F90:21-22)                     */
  double dx, dt;    /* crm_dx [m], crm_dt [s]  (cam/crmdims.F90:24-26)                      */
} crm_dims;

/* crm_input_type (crm/crm_input_module.F90:10-47). */
typedef struct crm_input_view {
  const double *zmid;    /* (ld,plev)   height of mid levels [m]          */
  const double *zint;    /* (ld,plev+1) height of interfaces [m]          */
  const double *tl;      /* (ld,plev)   temperature [K]                   */
  const double *ql;      /* (ld,plev)   water vapour [kg/kg]              */
  const double *qccl;    /* (ld,plev)   cloud liquid [kg/kg]              */
  const double *qiil;    /* (ld,plev)   cloud ice [kg/kg]                 */
  const double *ps;      /* (ld)        surface pressure [Pa]             */
  const double *pmid;    /* (ld,plev)   mid-level pressure [Pa]           */
  const double *pint;    /* (ld,plev+1) interface pressure [Pa]           */
  const double *pdel;    /* (ld,plev)   layer thickness [Pa]              */
  const double *phis;    /* (ld)        surface geopotential              */
  const double *ul;      /* (ld,plev)   u [m/s]                           */
  const double *vl;      /* (ld,plev)   v [m/s]                           */
  const double *ocnfrac; /* (ld)                                          */
  const double *tau00;   /* (ld)        surface stress [N/m2]             */
  const double *wndls;   /* (ld)        surface wind [m/s]                */
  const double *bflxls;  /* (ld)        surface buoyancy flux [K m/s]     */
  const double *fluxu00, *fluxv00, *fluxt00, *fluxq00; /* (ld) unused by crm() */
} crm_input_view;

/* crm_state_type (crm/crm_state_module.F90:12-56), sam1mom members. (ld,nx,ny,nz) each. */
typedef struct crm_state_view {
  double *u_wind, *v_wind, *w_wind, *temperature, *qt, *qp, *qn;
} crm_state_view;

/* crm_rad_type (crm/crm_rad_module.F90:17-34). (ld,nx_rad,ny_rad,nz) each.
 * qrad is read; the other five are overwritten with time averages. */
typedef struct crm_rad_view {
  const double *qrad;
  double *temperature, *qv, *qc, *qi, *cld;
} crm_rad_view;

/* crm_output_type (crm/crm_output_module.F90:7-133): the members crm() writes for the
 * 1-moment microphysics. cltot/clhgh/clmed/cllow are accumulated onto the values the
 * caller passes in (crm_module.F90:1057-1069,1474-1477), as in the reference. */
typedef struct crm_output_view {
  /* (ld,nx,ny,nz) */
  double *qcl, *qci, *qpl, *qpi, *tk, *tkh;
  /* (ld,nx,ny) */
  double *prec_crm;
  /* (ld) */
  double *cltot, *clhgh, *clmed, *cllow;
  double *precc, *precl, *precsc, *precsl;
  double *jt_crm, *mx_crm, *prectend, *precstend, *taux, *tauy, *z0m, *timing_factor;
  /* (ld,plev) */
  double *cldtop, *qc_mean, *qi_mean, *qs_mean, *qg_mean, *qr_mean;
  double *sltend, *qltend, *qcltend, *qiltend;
  double *cld, *gicewp, *gliqwp, *mctot, *mcup, *mcdn, *mcuup, *mcudn;
  double *mu_crm, *md_crm, *du_crm, *eu_crm, *ed_crm;
  double *flux_qt, *fluxsgs_qt, *tkez, *tkesgsz, *tkz, *flux_u, *flux_v, *flux_qp;
  double *precflux, *qt_ls, *qt_trans, *qp_trans, *qp_fall, *qp_src, *qp_evp, *t_ls;
} crm_output_view;

/* Per-call GCM-side context that the reference reads through CAM modules
 * (crm_module.F90:287-291,679): get_rlat_p/get_rlon_p [radians], get_gcol_p, get_nstep. */
typedef struct crm_column_info {
  int lchnk;
  const int *gcol;      /* (ncrms) global column ids, seed = gcol*1000 (crm_module.F90:679) */
  const double *lat;    /* (ncrms) latitude  [radians]                                      */
  const double *lon;    /* (ncrms) longitude [radians]                                      */
  int igstep;           /* GCM step number; setperturb runs when igstep <= 1                */
} crm_column_info;

typedef struct crm_hip_config {
  int device;           /* HIP device ordinal                                               */
  int precision;        /* CRM_PREC_FP64 or CRM_PREC_MIXED                                  */
  int max_ncrms;        /* instances to size device buffers for (0: grow on demand)         */
} crm_hip_config;

typedef struct crm_hip_stats {
  int nstop;            /* substeps of the last call                                        */
  int ncycle_max;       /* largest kurant subcycle count seen in the last call              */
  int nprec_max;        /* largest precipitation sub-step count seen in the last call       */
  double ms_device;     /* device time of the last call's substep loop [ms] (HIP events)    */
  int accel_ceased;     /* mean-state acceleration stopped by its 5 K test (accelerate_crm.F90:255-278) */
} crm_hip_stats;

typedef struct crm_hip_ctx crm_hip_ctx;

/* Create a context: one HIP stream plus persistent device buffers. Calls on one context are
 * serialised by the caller; separate contexts are independent. */
int crm_hip_init(crm_hip_ctx **ctx, const crm_hip_config *cfg);

/* One GCM physics step of the CRM on host (caller) memory -- the drop-in for crm()
 * (crm_module.F90:45-1718). Copies the views to the device, advances nstop=dt_gl/dt substeps,
 * writes crm_state, crm_rad and crm_output back. Stateless across calls (crm_module.F90:727,
 * vars.F90:387-389): device buffers persist, values do not. */
int crm_hip_run(crm_hip_ctx *ctx, const crm_dims *dims, const crm_column_info *col,
                double dt_gl, const crm_input_view *in, crm_state_view *state,
                crm_rad_view *rad, crm_output_view *out);

/* Same as crm_hip_run but every view pointer is a device pointer (caller-resident in HBM),
 * and col->gcol/lat/lon stay host pointers. No host<->device copies of the views. */
int crm_hip_run_device(crm_hip_ctx *ctx, const crm_dims *dims, const crm_column_info *col,
                       double dt_gl, const crm_input_view *in, crm_state_view *state,
                       crm_rad_view *rad, crm_output_view *out);

/* No counterpart in the reference: crm_hip_run_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as crm_hip_run_device, and it calls no hipMalloc, hipFree or other
 * allocating routine: every device buffer of the call (fields, level profiles, transform
 * operators, the call's radiative heating, the batch-global maxima) is carved from one fixed
 * array of CRM_HIP_NOALLOC_SCRATCH_BYTES per GPU that the library holds until the process ends.
 * It runs plain launches on the context's stream (no graphs, no kernel timing) and rejects an
 * attached microphysics network or sample recording (CRM_ERR_UNSUPPORTED). A batch whose
 * buffers do not fit returns CRM_ERR_SCRATCH before anything is launched: about 0.85 MB per
 * 64 x 58 instance in FP64 (about 300 instances), 0.2 MB per 32 x 28 instance; the mixed
 * precision form needs about half. Every call on one GPU shares that GPU's array, and each
 * call clears what it needs: calls that overlap on one GPU (from several contexts or threads)
 * race and must be serialised by the caller. The first call of a size class builds its
 * transform operators on the host (host memory only). */
#define CRM_HIP_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int crm_hip_run_device_noalloc(crm_hip_ctx *ctx, const crm_dims *dims, const crm_column_info *col,
                               double dt_gl, const crm_input_view *in, crm_state_view *state,
                               crm_rad_view *rad, crm_output_view *out);

/* Mean-state acceleration settings of a context, the replacement of crm_accel_init
 * (accelerate_crm.F90:30-38,45-75, which reads them from the physics namelist): with
 * use_crm_accel != 0 every later call integrates nstop / (1 + crm_accel_factor) substeps and
 * amplifies the change of the horizontal-mean t and qt (and u, v when crm_accel_uv != 0) of
 * each subcycle by crm_accel_factor (accelerate_crm.F90:109-366, crm_module.F90:718-720,
 * 872-876). Default: off. */
int crm_hip_accel_init(crm_hip_ctx *ctx, int use_crm_accel, double crm_accel_factor, int crm_accel_uv);

/* The reference's Coriolis switch docoriolis (crm/params.F90:107, .false. there and never set by
 * setparm) for later calls of a context: the f-plane Coriolis terms of coriolis.F90:7-53 with
 * fcor = 4 pi/86400 sin(lat) and fcorz from the instance's latitude (crm_module.F90:287,321-322)
 * and the large-scale wind ug0, vg0 (crm_module.F90:576-577), added to the momentum tendencies
 * after advection and SGS diffusion (crm_module.F90:825). 3D CRMs only: in 2D the reference's
 * Coriolis term would make the v the 2D form does not carry evolve, and a call then returns
 * CRM_ERR_UNSUPPORTED. Default: off. */
int crm_hip_set_coriolis(crm_hip_ctx *ctx, int docoriolis);

/* ---- Batches of uneven size classes, launch schedule, graphs -------------------------------
 * These have no counterpart in the reference, whose crm() runs one compile-time size class per
 * build (cam/crmdims.F90:13-22); they serve callers that advance several classes at once. */

/* Kernel classes of one call, in launch order (CRM_K_* of crm_hip_debug.h). */
#define CRM_HIP_NKERNEL 15

/* One size class of crm_hip_run_many: the arguments of one crm_hip_run_device call. */
typedef struct crm_class_desc {
  crm_dims dims;
  crm_column_info col;
  crm_input_view in;
  crm_state_view state;
  crm_rad_view rad;
  crm_output_view out;
} crm_class_desc;

/* Advance n size classes (uneven nx, nz, 2D and 3D) by dt_gl each, concurrently on the
 * device: class i runs on context ctxs[i] (its own stream, buffers, schedule and graphs) with
 * exactly the semantics of crm_hip_run_device(ctxs[i], &classes[i].dims, ...), including its
 * own batch-global subcycle and sedimentation step counts. status[i] receives each class's
 * code; the return value is the first nonzero one. Contexts must be distinct. */
int crm_hip_run_many(int n, crm_hip_ctx *const *ctxs, const crm_class_desc *classes, double dt_gl,
                     int *status);

/* Workgroup size (64, 128, 192 or 256 threads) of each kernel class of a context. */
int crm_hip_set_schedule(crm_hip_ctx *ctx, const int block[CRM_HIP_NKERNEL]);
int crm_hip_get_schedule(const crm_hip_ctx *ctx, int block[CRM_HIP_NKERNEL]);

/* Launch variant ("tile shape") of each kernel class of a context, 0 the default; results are
 * bitwise the same in every variant. CRM_K_PRESS of 2D 64-column CRMs: 0 one instance per
 * 4-wave workgroup (two workgroups per CU), 1 one instance on 8 waves with the gradient pass's
 * inputs kept in registers from the right-hand side pass (one workgroup per CU), 2 and 3 as 0
 * with two or four levels of each pass unrolled together, 4 as 0 with the Thomas factors in
 * registers (the workgroup's LDS is the slab alone: three workgroups per CU). CRM_K_MPDATA: 0 the default number
 * of levels per trip of the sweep's steady state (two for 2D 64-column CRMs, else one), 1 the
 * other, 2 and 3 as 0 and 1 with the registers bounded for three waves per SIMD, 4 and 5 as 0
 * and 2 with each of the three scalars advected on its own waves. Other classes have variant 0
 * only. */
int crm_hip_set_variants(crm_hip_ctx *ctx, const int variant[CRM_HIP_NKERNEL]);
int crm_hip_get_variants(const crm_hip_ctx *ctx, int variant[CRM_HIP_NKERNEL]);

/* Autotuner: runs the call once per candidate workgroup size (64, 128, 256) and launch variant,
 * with every kernel class timed on the context's stream, keeps the fastest (size, variant) of
 * each class (the default 256 threads and variant 0 unless another is faster by more than 3%)
 * for later calls of this context and writes the sizes to block_out (may be NULL;
 * crm_hip_get_variants gives the variants). The views are device pointers and are advanced by
 * every trial call: pass scratch copies. */
int crm_hip_autotune(crm_hip_ctx *ctx, const crm_dims *dims, const crm_column_info *col, double dt_gl,
                     const crm_input_view *in, crm_state_view *state, crm_rad_view *rad,
                     crm_output_view *out, int block_out[CRM_HIP_NKERNEL]);

/* HIP graphs.  CRM_GRAPHS_SUBSTEP: steady substeps (one subcycle, from the third substep on)
 * replay a captured graph of the whole substep, the subcycle count decided on the host before
 * each.  CRM_GRAPHS_CALL (the default): a call replays one captured graph of all its substeps
 * with one subcycle each; the subcycle count is decided on the device, and a substep that needs
 * more stops the replay there and continues as with CRM_GRAPHS_SUBSTEP; crm_hip_run_many runs
 * each class's graph on its own stream, concurrently.  CRM_GRAPHS_JOINT: as CRM_GRAPHS_CALL, and
 * when every context of a crm_hip_run_many call has this mode, the classes' whole-call graphs
 * are captured together into one graph (class streams forked from and joined into one capture)
 * and launched once.  CRM_GRAPHS_OFF: plain launches.  Results are bitwise the same in every
 * mode. */
enum { CRM_GRAPHS_OFF = 0, CRM_GRAPHS_SUBSTEP = 1, CRM_GRAPHS_CALL = 2, CRM_GRAPHS_JOINT = 3 };
int crm_hip_set_graphs(crm_hip_ctx *ctx, int mode);

int crm_hip_get_stats(const crm_hip_ctx *ctx, crm_hip_stats *stats);
int crm_hip_finalize(crm_hip_ctx *ctx);
const char *crm_hip_errstr(int code);

#ifdef __cplusplus
}
#endif
#endif /* CRM_HIP_H */
