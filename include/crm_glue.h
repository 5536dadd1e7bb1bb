/* crm_glue.h -- the GCM-side code around crm() on the device: building crm_input from the
 * physics state, the surface-flux bypass, the tendencies and radiation conversions after the
 * call, and the column energy and water check.  Replaces the host loops of crm_physics_tend
 * (components/cam/src/physics/cam/crm_physics.F90:1050-1330), crm_surface_flux_bypass_tend
 * (crm_physics.F90:1652-1699), check_energy_chng (cam/check_energy.F90:358-561) and the CRM
 * side of the radiation (rrtmg/radiation.F90:1584-1636, 2393-2398), so a physics state
 * resident in HBM reaches the CRM and returns without host round trips.
 *
 * Every array is a device pointer in the Fortran layout of the physics types: (pcols, pver)
 * element (i, k) at p[i + pcols*k], 0-based, level 0 the model top; (pcols, pver+1)
 * likewise; per-column arrays (pcols).  Only columns 0..ncol-1 are read or written.  Each call
 * runs on the context's stream and returns when the work is done.  Constants are those of
 * physconst / shr_const_mod (cime/src/share/util/shr_const_mod.F90:22-56): gravit 9.80616,
 * cpair 1004.64, latvap 2.501e6, latice 3.337e5.
 */
#ifndef CRM_GLUE_H
#define CRM_GLUE_H

#include "crm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The members of physics_state that crm_physics_tend reads (q split into its constituents). */
typedef struct crm_gcm_state_view {
  const double *zm;     /* (pcols,pver)   */
  const double *zi;     /* (pcols,pver+1) */
  const double *t;      /* (pcols,pver)   */
  const double *qv;     /* (pcols,pver)   q(:,:,1)        */
  const double *qcl;    /* (pcols,pver)   q(:,:,ixcldliq) */
  const double *qci;    /* (pcols,pver)   q(:,:,ixcldice) */
  const double *ps;     /* (pcols)        */
  const double *pmid;   /* (pcols,pver)   */
  const double *pint;   /* (pcols,pver+1) */
  const double *pdel;   /* (pcols,pver)   */
  const double *phis;   /* (pcols)        */
  const double *u;      /* (pcols,pver)   */
  const double *v;      /* (pcols,pver)   */
} crm_gcm_state_view;

/* The members of cam_in_t that crm_physics_tend reads. */
typedef struct crm_gcm_surface_view {
  const double *ocnfrac, *wsx, *wsy, *shf, *lhf; /* (pcols) */
} crm_gcm_surface_view;

/* Before crm() (crm_physics.F90:1005-1010, 1050-1125): crm_input from the physics state and
 * surface fluxes, with the input wind rotated by the CRM orientation crm_angle[ncol] (radians;
 * all zero for a fixed orientation); in points at the caller's crm_input arrays (leading
 * dimension pcols, plev = pver), which are written.  The crm_rad qrad carried as Q dp
 * (pcols,nx_rad,ny_rad,crm_nz; nrad = nx_rad*ny_rad) is divided by pdel of its GCM level in
 * place, back to a heating rate. */
int crm_hip_gcm_input(crm_hip_ctx *ctx, int ncol, int pcols, int pver, int crm_nz, int nrad,
                      const crm_gcm_state_view *state, const crm_gcm_surface_view *srf, const double *crm_angle,
                      const crm_input_view *in, const crm_rad_view *rad);

/* Surface flux bypass (crm_physics.F90:1652-1703): ptend_s and ptend_qv (pcols,pver) are zeroed
 * and their bottom level set to gravit/pdel times shf and cflx_vap (the vapour flux). */
int crm_hip_gcm_flux_bypass(crm_hip_ctx *ctx, int ncol, int pcols, int pver, const double *pdel,
                            const double *shf, const double *cflx_vap, double *ptend_s, double *ptend_qv);

/* After crm() (crm_physics.F90:1158-1169, 1203-1209, 1310-1330):
 *  - ptend_qv, ptend_qcl, ptend_qci, ptend_s (pcols,pver) from qltend, qcltend, qiltend, sltend;
 *  - precc += precl, precsc += precsl, precl = precsl = 0;
 *  - crm_rad qrad (pcols,nx_rad,ny_rad,crm_nz) times pdel of its GCM level (energy
 *    conservation across steps), written in place;
 *  - qrs, qrl (pcols,pver) divided by pdel in place, the radiative heating qrs + qrl put into
 *    ptend_s above the CRM and in its top two levels, and sp_rad_flux[ncol] =
 *    sum over levels of (qrs + qrl) pdel / gravit. */
int crm_hip_gcm_tend(crm_hip_ctx *ctx, int ncol, int pcols, int pver, int crm_nz, int nrad,
                     const double *pdel, crm_output_view *out, const crm_rad_view *rad, double *qrs,
                     double *qrl, double *ptend_s, double *ptend_qv, double *ptend_qcl, double *ptend_qci,
                     double *sp_rad_flux);

/* The GCM column's cloud fields that the radiation reads (pcols,pver): cloud fraction, ice
 * fraction, in-cloud ice and liquid water paths [kg/m2]. */
typedef struct crm_gcm_cloud_view {
  const double *cld, *fice, *cicewp, *cliqwp;
} crm_gcm_cloud_view;

/* The radiation's inputs in each CRM radiation column: (pcols,nrad,pver) each, element
 * (i, r, k) at p[i + pcols*(r + nrad*k)], r = ii + nx_rad*jj (0-based). */
typedef struct crm_rad_column_view {
  double *t, *qv, *qc, *qi, *cld, *fice, *cicewp, *cliqwp;
} crm_rad_column_view;

/* Before the radiation (rrtmg/radiation.F90:1584-1636, sam1mom): its inputs in every CRM
 * radiation column.  Each radiation column starts from the physics state (t, qv, qcl, qci) and
 * the GCM cloud fields; on the CRM levels (GCM level pver-1-m of CRM level m) they are taken
 * from that column's crm_rad averages: cld; where cld > 0 the ice fraction qi/(qc+qi) and the
 * in-cloud water paths q pdel/gravit/max(0.01, cld) of ice and liquid, else 0; qi, qc;
 * vapour max(1e-9, qv); t.  With sam1mom the effective radii are the GCM column's in every
 * radiation column (radiation.F90:1675-1686), so they are not repeated here.  The radiation
 * scheme itself is the caller's. */
int crm_hip_rad_columns(crm_hip_ctx *ctx, int ncol, int pcols, int pver, int crm_nz, int nrad,
                        const crm_gcm_state_view *state, const crm_gcm_cloud_view *cloud,
                        const crm_rad_view *rad, const crm_rad_column_view *cols);

/* After the radiation (radiation.F90:2393-2398): crm_rad qrad (pcols,nx_rad,ny_rad,crm_nz) is
 * the sum qrs_crm + qrl_crm of the shortwave and longwave heating the radiation computed in
 * each radiation column (same layout), written through rad->qrad. */
int crm_hip_rad_qrad(crm_hip_ctx *ctx, int ncol, int pcols, int nrad, int crm_nz, const double *qrs_crm,
                     const double *qrl_crm, const crm_rad_view *rad);

/* check_energy_chng (check_energy.F90:358-561) without rain and snow constituents (absent with
 * sam1mom): the column integrals of total energy te and total water tw of the state, the
 * expected values from te_cur/tw_cur and the boundary fluxes over ztodt, their relative
 * errors te_rer/tw_rer[ncol] (tw_rer 0 where tw_cur <= 0), the cumulative flux tendencies
 * te_tnd/tw_tnd[ncol] (added to), and te_cur/tw_cur[ncol] replaced by te, tw. */
int crm_hip_check_energy_chng(crm_hip_ctx *ctx, int ncol, int pcols, int pver, const crm_gcm_state_view *state,
                              double ztodt, const double *flx_vap, const double *flx_cnd, const double *flx_ice,
                              const double *flx_sen, double *te_cur, double *tw_cur, double *te_tnd,
                              double *tw_tnd, double *te_rer, double *tw_rer);

#ifdef __cplusplus
}
#endif
#endif
