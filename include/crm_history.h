/* crm_history.h -- time-mean history fields of the CRM accumulated on the device.
 *
 * Replaces the outfld calls of crm_physics_tend for the CRM's own fields
 * (components/cam/src/physics/cam/crm_physics.F90:1201-1461, sam1mom without MAML, CLUBB or
 * ECPP): each accumulate call is one outfld pass over every field of one chunk, adding the
 * field's values for columns 0..ncol-1 to a device sum, as CAM's history buffer does for
 * fields averaged over the history interval (avgflag 'A'); crm_hist_mean returns sum / count.
 * Derived fields use the reference's expressions:
 *   CRM_QV   = qt - qcl - qci                                   (crm_physics.F90:1220)
 *   SPDT     = (ptend_s - qrs - qrl) / cpair                    (:1333,1341)
 *   SPQRL, SPQRS = qrl / cpair, qrs / cpair                     (:1317-1318)
 *   ICLDIWP  = gicewp * 1.0e-3 / max(0.01, cld)                 (:1406,1421)
 *   ICLDTWP  = gicewp*1e-3/max(0.01,cld) + gliqwp*1e-3/max(0.01,cld) (:1406-1407,1414,1420)
 *   GCLDLWP  = gicewp + gliqwp                                  (:1413,1416)
 *   TGCLDIWP, TGCLDLWP = sum over levels, top first, of gicewp*1e-3, gliqwp*1e-3 (:1408-1409)
 *   TGCLDCWP = TGCLDIWP + TGCLDLWP                              (:1412,1417)
 * where 1.0e-3 is a 4-byte literal as in the reference build (no -r8) and cpair = 1004.64.
 * The other fields are crm_state / crm_output / physics members copied as they are.
 *
 * Layouts are those of the sources (column-major, pcols leading, 0-based); device pointers
 * throughout.  The work runs on the stream of the crm_hip_ctx passed to each call.
 */
#ifndef CRM_HISTORY_H
#define CRM_HISTORY_H

#include "crm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crm_history crm_history;

/* Inputs of one outfld pass.  A NULL pointer skips the fields that need it (their sample
 * counts do not advance), so one chunk's fields may be accumulated in several calls, e.g.
 * CRM_QRAD with the heating rate before crm_hip_gcm_tend turns qrad into Q dp, and SPDT after
 * it. */
typedef struct crm_hist_inputs {
  int ncol;                       /* columns of this chunk, <= pcols                            */
  const crm_state_view *state;    /* crm_state after crm(): CRM_U/V/W/T/QV                      */
  const crm_output_view *out;     /* crm_output after crm()                                     */
  const double *qrad;             /* (pcols,nx_rad,ny_rad,crm_nz) heating rate the CRM used     */
  const double *pmid, *pdel;      /* (pcols,pver) physics state: PRES, DPRES                    */
  const double *ptend_s, *ptend_qv, *ptend_qcl, *ptend_qci; /* (pcols,pver) after crm_hip_gcm_tend */
  const double *qrs, *qrl;        /* (pcols,pver) after crm_hip_gcm_tend (divided by pdel)      */
} crm_hist_inputs;

/* Sums for every field for chunks of at most pcols columns of one CRM size class. */
int crm_hist_create(crm_history **h, int device, int pcols, int pver, int nx, int ny, int crm_nz, int nx_rad,
                    int ny_rad);
/* One outfld pass over every field whose inputs are given. */
int crm_hist_accumulate(crm_history *h, crm_hip_ctx *ctx, const crm_hist_inputs *in);
/* Number of fields, and a field's history name (e.g. "CRM_QV", "SPDT"), its size per column
 * and the index of a name (-1 if unknown). */
int crm_hist_nfields(void);
const char *crm_hist_name(int field);
long long crm_hist_size(const crm_history *h, int field);
int crm_hist_index(const char *name);
/* Time mean of a field into dst (device, the field's layout with leading dimension pcols):
 * sum / count for every column, 0 where no sample was taken; *count receives the sample count. */
int crm_hist_mean(crm_history *h, crm_hip_ctx *ctx, int field, double *dst, long long *count);
/* Start a new averaging interval. */
int crm_hist_reset(crm_history *h, crm_hip_ctx *ctx);
int crm_hist_destroy(crm_history *h);

#ifdef __cplusplus
}
#endif
#endif /* CRM_HISTORY_H */
