// crm_history.hip -- time-mean history fields of the CRM on the device (include/crm_history.h):
// the outfld calls of crm_physics_tend for the CRM's fields
// (components/cam/src/physics/cam/crm_physics.F90:1201-1461) as sums in HBM, one pass per
// call over every field, with the reference's expressions for the derived ones evaluated in
// the reference's order (floating-point contraction off).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/crm_history.h"
#include "crm_engine.hpp"

#pragma clang fp contract(off)

namespace {

constexpr double CPAIR = 1.00464e3;  // shr_const_mod.F90:47 via physconst
// crm_physics.F90:1406-1409: the literal 1.0e-3 is a default (4-byte) real in the reference build
constexpr double KG_PER_G = (double)1.0e-3f;

// Per-column extents of the fields.
enum Shape { SH_GRID, SH_RAD, SH_SFC, SH_LEV, SH_COL };
// How a field is made from its sources a, b, c (each in the field's layout unless stated).
enum Op {
  OP_COPY,     // a
  OP_QV,       // a - b - c                                        (crm_physics.F90:1220)
  OP_SPDT,     // (a - b - c) / cpair                              (:1333)
  OP_DIVCP,    // a / cpair                                        (:1317-1318)
  OP_ADD,      // a + b                                            (:1413)
  OP_INCLOUD,  // a * 1e-3 / max(0.01, c)                          (:1406)
  OP_INCLOUD2, // a * 1e-3 / max(0.01, c) + b * 1e-3 / max(0.01, c) (:1406-1407,1414)
  OP_COLSUM,   // per column, sum over levels of a * 1e-3 (a: (pcols,pver))   (:1408-1409)
  OP_COLSUM2   // the same sums of a and b, added                  (:1408-1409,1412)
};

// Where a source comes from.
enum Src {
  X_NONE, X_U, X_V, X_W, X_T, X_QT, X_QCL, X_QCI, X_QPL, X_QPI, X_TK, X_TKH, X_PREC, X_QRAD,
  X_PMID, X_PDEL, X_PTS, X_PTQV, X_PTQCL, X_PTQCI, X_QRS, X_QRL,
  X_MCTOT, X_MCUP, X_MCDN, X_MCUUP, X_MCUDN, X_QCM, X_QIM, X_QSM, X_QGM, X_QRM,
  X_FLUXQT, X_FLUXU, X_FLUXV, X_TKEZ, X_TKESGSZ, X_TKZ, X_FLUXSGSQT, X_FLUXQP, X_PRECFLUX,
  X_QTLS, X_QTTRANS, X_QPTRANS, X_QPEVP, X_QPFALL, X_QPSRC, X_TLS, X_CLD, X_CLTOT, X_CLHGH,
  X_CLMED, X_CLLOW, X_CLDTOP, X_TIMING, X_GICEWP, X_GLIQWP, X_MU, X_MD, X_EU, X_DU, X_ED
};

struct Field {
  const char* name;
  Shape shape;
  Op op;
  Src a, b, c;
};

// In the reference's outfld order (crm_physics.F90:1201-1461).
const Field FIELDS[] = {
    {"CRM_QRAD", SH_RAD, OP_COPY, X_QRAD, X_NONE, X_NONE},
    {"PRES", SH_LEV, OP_COPY, X_PMID, X_NONE, X_NONE},
    {"DPRES", SH_LEV, OP_COPY, X_PDEL, X_NONE, X_NONE},
    {"CRM_U", SH_GRID, OP_COPY, X_U, X_NONE, X_NONE},
    {"CRM_V", SH_GRID, OP_COPY, X_V, X_NONE, X_NONE},
    {"CRM_W", SH_GRID, OP_COPY, X_W, X_NONE, X_NONE},
    {"CRM_T", SH_GRID, OP_COPY, X_T, X_NONE, X_NONE},
    {"CRM_QV", SH_GRID, OP_QV, X_QT, X_QCL, X_QCI},
    {"CRM_QC", SH_GRID, OP_COPY, X_QCL, X_NONE, X_NONE},
    {"CRM_QI", SH_GRID, OP_COPY, X_QCI, X_NONE, X_NONE},
    {"CRM_QPC", SH_GRID, OP_COPY, X_QPL, X_NONE, X_NONE},
    {"CRM_QPI", SH_GRID, OP_COPY, X_QPI, X_NONE, X_NONE},
    {"CRM_PREC", SH_SFC, OP_COPY, X_PREC, X_NONE, X_NONE},
    {"CRM_TK", SH_GRID, OP_COPY, X_TK, X_NONE, X_NONE},
    {"CRM_TKH", SH_GRID, OP_COPY, X_TKH, X_NONE, X_NONE},
    {"SPQRL", SH_LEV, OP_DIVCP, X_QRL, X_NONE, X_NONE},
    {"SPQRS", SH_LEV, OP_DIVCP, X_QRS, X_NONE, X_NONE},
    {"SPDQ", SH_LEV, OP_COPY, X_PTQV, X_NONE, X_NONE},
    {"SPDQC", SH_LEV, OP_COPY, X_PTQCL, X_NONE, X_NONE},
    {"SPDQI", SH_LEV, OP_COPY, X_PTQCI, X_NONE, X_NONE},
    {"SPDT", SH_LEV, OP_SPDT, X_PTS, X_QRS, X_QRL},
    {"SPMC", SH_LEV, OP_COPY, X_MCTOT, X_NONE, X_NONE},
    {"SPMCUP", SH_LEV, OP_COPY, X_MCUP, X_NONE, X_NONE},
    {"SPMCDN", SH_LEV, OP_COPY, X_MCDN, X_NONE, X_NONE},
    {"SPMCUUP", SH_LEV, OP_COPY, X_MCUUP, X_NONE, X_NONE},
    {"SPMCUDN", SH_LEV, OP_COPY, X_MCUDN, X_NONE, X_NONE},
    {"SPQC", SH_LEV, OP_COPY, X_QCM, X_NONE, X_NONE},
    {"SPQI", SH_LEV, OP_COPY, X_QIM, X_NONE, X_NONE},
    {"SPQS", SH_LEV, OP_COPY, X_QSM, X_NONE, X_NONE},
    {"SPQG", SH_LEV, OP_COPY, X_QGM, X_NONE, X_NONE},
    {"SPQR", SH_LEV, OP_COPY, X_QRM, X_NONE, X_NONE},
    {"SPQTFLX", SH_LEV, OP_COPY, X_FLUXQT, X_NONE, X_NONE},
    {"SPUFLX", SH_LEV, OP_COPY, X_FLUXU, X_NONE, X_NONE},
    {"SPVFLX", SH_LEV, OP_COPY, X_FLUXV, X_NONE, X_NONE},
    {"SPTKE", SH_LEV, OP_COPY, X_TKEZ, X_NONE, X_NONE},
    {"SPTKES", SH_LEV, OP_COPY, X_TKESGSZ, X_NONE, X_NONE},
    {"SPTK", SH_LEV, OP_COPY, X_TKZ, X_NONE, X_NONE},
    {"SPQTFLXS", SH_LEV, OP_COPY, X_FLUXSGSQT, X_NONE, X_NONE},
    {"SPQPFLX", SH_LEV, OP_COPY, X_FLUXQP, X_NONE, X_NONE},
    {"SPPFLX", SH_LEV, OP_COPY, X_PRECFLUX, X_NONE, X_NONE},
    {"SPQTLS", SH_LEV, OP_COPY, X_QTLS, X_NONE, X_NONE},
    {"SPQTTR", SH_LEV, OP_COPY, X_QTTRANS, X_NONE, X_NONE},
    {"SPQPTR", SH_LEV, OP_COPY, X_QPTRANS, X_NONE, X_NONE},
    {"SPQPEVP", SH_LEV, OP_COPY, X_QPEVP, X_NONE, X_NONE},
    {"SPQPFALL", SH_LEV, OP_COPY, X_QPFALL, X_NONE, X_NONE},
    {"SPQPSRC", SH_LEV, OP_COPY, X_QPSRC, X_NONE, X_NONE},
    {"SPTLS", SH_LEV, OP_COPY, X_TLS, X_NONE, X_NONE},
    {"CLOUD", SH_LEV, OP_COPY, X_CLD, X_NONE, X_NONE},
    {"CLDTOT", SH_COL, OP_COPY, X_CLTOT, X_NONE, X_NONE},
    {"CLDHGH", SH_COL, OP_COPY, X_CLHGH, X_NONE, X_NONE},
    {"CLDMED", SH_COL, OP_COPY, X_CLMED, X_NONE, X_NONE},
    {"CLDLOW", SH_COL, OP_COPY, X_CLLOW, X_NONE, X_NONE},
    {"CLOUDTOP", SH_LEV, OP_COPY, X_CLDTOP, X_NONE, X_NONE},
    {"TIMINGF", SH_COL, OP_COPY, X_TIMING, X_NONE, X_NONE},
    {"GCLDLWP", SH_LEV, OP_ADD, X_GICEWP, X_GLIQWP, X_NONE},
    {"TGCLDCWP", SH_COL, OP_COLSUM2, X_GICEWP, X_GLIQWP, X_NONE},
    {"TGCLDLWP", SH_COL, OP_COLSUM, X_GLIQWP, X_NONE, X_NONE},
    {"TGCLDIWP", SH_COL, OP_COLSUM, X_GICEWP, X_NONE, X_NONE},
    {"ICLDTWP", SH_LEV, OP_INCLOUD2, X_GICEWP, X_GLIQWP, X_CLD},
    {"ICLDIWP", SH_LEV, OP_INCLOUD, X_GICEWP, X_NONE, X_CLD},
    {"MU_CRM", SH_LEV, OP_COPY, X_MU, X_NONE, X_NONE},
    {"MD_CRM", SH_LEV, OP_COPY, X_MD, X_NONE, X_NONE},
    {"EU_CRM", SH_LEV, OP_COPY, X_EU, X_NONE, X_NONE},
    {"DU_CRM", SH_LEV, OP_COPY, X_DU, X_NONE, X_NONE},
    {"ED_CRM", SH_LEV, OP_COPY, X_ED, X_NONE, X_NONE},
};
constexpr int NFIELDS = sizeof(FIELDS) / sizeof(FIELDS[0]);

const double* source(Src x, const crm_hist_inputs* in) {
  const crm_state_view* s = in->state;
  const crm_output_view* o = in->out;
  switch (x) {
    case X_U: return s ? s->u_wind : nullptr;
    case X_V: return s ? s->v_wind : nullptr;
    case X_W: return s ? s->w_wind : nullptr;
    case X_T: return s ? s->temperature : nullptr;
    case X_QT: return s ? s->qt : nullptr;
    case X_QRAD: return in->qrad;
    case X_PMID: return in->pmid;
    case X_PDEL: return in->pdel;
    case X_PTS: return in->ptend_s;
    case X_PTQV: return in->ptend_qv;
    case X_PTQCL: return in->ptend_qcl;
    case X_PTQCI: return in->ptend_qci;
    case X_QRS: return in->qrs;
    case X_QRL: return in->qrl;
    default: break;
  }
  if (!o) return nullptr;
  switch (x) {
    case X_QCL: return o->qcl;
    case X_QCI: return o->qci;
    case X_QPL: return o->qpl;
    case X_QPI: return o->qpi;
    case X_TK: return o->tk;
    case X_TKH: return o->tkh;
    case X_PREC: return o->prec_crm;
    case X_MCTOT: return o->mctot;
    case X_MCUP: return o->mcup;
    case X_MCDN: return o->mcdn;
    case X_MCUUP: return o->mcuup;
    case X_MCUDN: return o->mcudn;
    case X_QCM: return o->qc_mean;
    case X_QIM: return o->qi_mean;
    case X_QSM: return o->qs_mean;
    case X_QGM: return o->qg_mean;
    case X_QRM: return o->qr_mean;
    case X_FLUXQT: return o->flux_qt;
    case X_FLUXU: return o->flux_u;
    case X_FLUXV: return o->flux_v;
    case X_TKEZ: return o->tkez;
    case X_TKESGSZ: return o->tkesgsz;
    case X_TKZ: return o->tkz;
    case X_FLUXSGSQT: return o->fluxsgs_qt;
    case X_FLUXQP: return o->flux_qp;
    case X_PRECFLUX: return o->precflux;
    case X_QTLS: return o->qt_ls;
    case X_QTTRANS: return o->qt_trans;
    case X_QPTRANS: return o->qp_trans;
    case X_QPEVP: return o->qp_evp;
    case X_QPFALL: return o->qp_fall;
    case X_QPSRC: return o->qp_src;
    case X_TLS: return o->t_ls;
    case X_CLD: return o->cld;
    case X_CLTOT: return o->cltot;
    case X_CLHGH: return o->clhgh;
    case X_CLMED: return o->clmed;
    case X_CLLOW: return o->cllow;
    case X_CLDTOP: return o->cldtop;
    case X_TIMING: return o->timing_factor;
    case X_GICEWP: return o->gicewp;
    case X_GLIQWP: return o->gliqwp;
    case X_MU: return o->mu_crm;
    case X_MD: return o->md_crm;
    case X_EU: return o->eu_crm;
    case X_DU: return o->du_crm;
    case X_ED: return o->ed_crm;
    default: return nullptr;
  }
}

// sum[i + pcols*m] += value of element (i, m) for columns i < ncol and m < per (elements per
// column); OP_COLSUM/2 read (pcols, nlev) sources and write per-column sums.
__global__ void k_hist_accum(double* __restrict__ sum, int op, const double* a, const double* b, const double* c,
                             int ncol, int pcols, int per, int nlev) {
  const size_t n = (size_t)ncol * per;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % ncol);
    const size_t m = e / ncol;
    const size_t o = (size_t)i + (size_t)pcols * m;
    double v;
    switch (op) {
      case OP_QV: v = a[o] - b[o] - c[o]; break;
      case OP_SPDT: v = (a[o] - b[o] - c[o]) / CPAIR; break;
      case OP_DIVCP: v = a[o] / CPAIR; break;
      case OP_ADD: v = a[o] + b[o]; break;
      case OP_INCLOUD: v = a[o] * KG_PER_G / fmax(0.01, c[o]); break;
      case OP_INCLOUD2: v = a[o] * KG_PER_G / fmax(0.01, c[o]) + b[o] * KG_PER_G / fmax(0.01, c[o]); break;
      case OP_COLSUM:
      case OP_COLSUM2: {
        double si = 0.0, sl = 0.0;
        for (int k = 0; k < nlev; k++) {
          si = si + a[(size_t)i + (size_t)pcols * k] * KG_PER_G;
          if (op == OP_COLSUM2) sl = sl + b[(size_t)i + (size_t)pcols * k] * KG_PER_G;
        }
        v = op == OP_COLSUM2 ? si + sl : si;
        break;
      }
      default: v = a[o]; break;
    }
    sum[o] += v;
  }
}

__global__ void k_hist_mean(double* __restrict__ dst, const double* __restrict__ sum, double inv, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    dst[e] = sum[e] * inv;
}

__global__ void k_hist_zero(double* __restrict__ p, size_t n) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) p[e] = 0.0;
}

dim3 grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  return dim3((unsigned)(g < 4096 ? (g > 0 ? g : 1) : 4096));
}

}  // namespace

struct crm_history {
  int device, pcols, pver, grid, rad, sfc;
  std::vector<size_t> off;  // per field: offset of its sums (doubles)
  std::vector<long long> count;
  double* sums = nullptr;
  size_t total = 0;
  long long per(int f) const {
    switch (FIELDS[f].shape) {
      case SH_GRID: return grid;
      case SH_RAD: return rad;
      case SH_SFC: return sfc;
      case SH_LEV: return pver;
      default: return 1;
    }
  }
};

extern "C" {

int crm_hist_create(crm_history** out, int device, int pcols, int pver, int nx, int ny, int crm_nz, int nx_rad,
                    int ny_rad) {
  if (!out || pcols < 1 || pver < 1 || nx < 1 || ny < 1 || crm_nz < 1 || nx_rad < 1 || ny_rad < 1)
    return CRM_ERR_ARG;
  *out = nullptr;
  auto* h = new crm_history();
  h->device = device;
  h->pcols = pcols;
  h->pver = pver;
  h->grid = nx * ny * crm_nz;
  h->rad = nx_rad * ny_rad * crm_nz;
  h->sfc = nx * ny;
  h->off.resize(NFIELDS);
  h->count.assign(NFIELDS, 0);
  for (int f = 0; f < NFIELDS; f++) {
    h->off[f] = h->total;
    h->total += (size_t)pcols * h->per(f);
  }
  if (hipSetDevice(device) != hipSuccess) {
    delete h;
    return CRM_ERR_DEVICE;
  }
  if (hipMalloc(&h->sums, h->total * sizeof(double)) != hipSuccess) {
    delete h;
    return CRM_ERR_NOMEM;
  }
  if (hipMemset(h->sums, 0, h->total * sizeof(double)) != hipSuccess) {
    (void)hipFree(h->sums);
    delete h;
    return CRM_ERR_DEVICE;
  }
  *out = h;
  return CRM_OK;
}

int crm_hist_accumulate(crm_history* h, crm_hip_ctx* ctx, const crm_hist_inputs* in) {
  if (!h || !ctx || !in || in->ncol < 1 || in->ncol > h->pcols) return CRM_ERR_ARG;
  if (crm_ctx_device(ctx) != h->device || hipSetDevice(h->device) != hipSuccess) return CRM_ERR_DEVICE;
  hipStream_t s = crm_ctx_stream(ctx);
  for (int f = 0; f < NFIELDS; f++) {
    const Field& F = FIELDS[f];
    const double* a = source(F.a, in);
    const double* b = F.b != X_NONE ? source(F.b, in) : nullptr;
    const double* c = F.c != X_NONE ? source(F.c, in) : nullptr;
    if (!a || (F.b != X_NONE && !b) || (F.c != X_NONE && !c)) continue;
    const size_t n = (size_t)in->ncol * h->per(f);
    k_hist_accum<<<grid_for(n), dim3(256), 0, s>>>(h->sums + h->off[f], F.op, a, b, c, in->ncol, h->pcols,
                                                    (int)h->per(f), h->pver);
    h->count[f]++;
  }
  return hipGetLastError() == hipSuccess ? CRM_OK : CRM_ERR_DEVICE;
}

int crm_hist_nfields(void) { return NFIELDS; }

const char* crm_hist_name(int f) { return (f >= 0 && f < NFIELDS) ? FIELDS[f].name : nullptr; }

long long crm_hist_size(const crm_history* h, int f) {
  if (!h || f < 0 || f >= NFIELDS) return -1;
  return h->per(f);
}

int crm_hist_index(const char* name) {
  if (!name) return -1;
  for (int f = 0; f < NFIELDS; f++)
    if (!std::strcmp(FIELDS[f].name, name)) return f;
  return -1;
}

int crm_hist_mean(crm_history* h, crm_hip_ctx* ctx, int f, double* dst, long long* count) {
  if (!h || !ctx || !dst || f < 0 || f >= NFIELDS) return CRM_ERR_ARG;
  if (crm_ctx_device(ctx) != h->device || hipSetDevice(h->device) != hipSuccess) return CRM_ERR_DEVICE;
  hipStream_t s = crm_ctx_stream(ctx);
  const size_t n = (size_t)h->pcols * h->per(f);
  const long long cnt = h->count[f];
  k_hist_mean<<<grid_for(n), dim3(256), 0, s>>>(dst, h->sums + h->off[f], cnt > 0 ? 1.0 / (double)cnt : 0.0, n);
  if (count) *count = cnt;
  if (hipGetLastError() != hipSuccess) return CRM_ERR_DEVICE;
  return hipStreamSynchronize(s) == hipSuccess ? CRM_OK : CRM_ERR_DEVICE;
}

int crm_hist_reset(crm_history* h, crm_hip_ctx* ctx) {
  if (!h || !ctx) return CRM_ERR_ARG;
  if (crm_ctx_device(ctx) != h->device || hipSetDevice(h->device) != hipSuccess) return CRM_ERR_DEVICE;
  k_hist_zero<<<grid_for(h->total), dim3(256), 0, crm_ctx_stream(ctx)>>>(h->sums, h->total);
  h->count.assign(NFIELDS, 0);
  return hipGetLastError() == hipSuccess ? CRM_OK : CRM_ERR_DEVICE;
}

int crm_hist_destroy(crm_history* h) {
  if (!h) return CRM_OK;
  (void)hipSetDevice(h->device);
  if (h->sums) (void)hipFree(h->sums);
  delete h;
  return CRM_OK;
}

}  // extern "C"
