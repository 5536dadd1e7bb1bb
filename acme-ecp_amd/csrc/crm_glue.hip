// crm_glue.hip -- the GCM-side code around crm() on the device (include/crm_glue.h):
// crm_input from the physics state, the surface-flux bypass, the tendencies and radiation
// conversions after the call, and check_energy_chng.  One thread per GCM column; level sums
// run in the reference's level order and floating-point contraction is off, so every result
// is the reference's expression evaluated in the reference's order.
#include <hip/hip_runtime.h>

#include "../../include/crm_glue.h"
#include "crm_engine.hpp"

#pragma clang fp contract(off)

namespace {

// physconst (cime/src/share/util/shr_const_mod.F90:22,47,53,54)
constexpr double GRAVIT = 9.80616, CPAIR = 1.00464e3, LATVAP = 2.501e6, LATICE = 3.337e5;
// the literal 0.61 of crm_physics.F90:1092 is a default (4-byte) real
constexpr double C061 = (double)0.61f;

struct Cols {
  int pcols;
  __device__ size_t at(int i, int k) const { return (size_t)i + (size_t)pcols * k; }
};

// crm_physics.F90:1005-1010, 1050-1125
__global__ void k_gcm_input(int ncol, int pcols, int pver, int crm_nz, int nrad, crm_gcm_state_view st,
                            crm_gcm_surface_view sf, const double* angle, crm_input_view in, crm_rad_view rad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  const Cols L{pcols};
  // Q dp -> heating rate on the CRM levels (CRM level m sits on GCM level pver-1-m)
  double* qrad = const_cast<double*>(rad.qrad);
  for (int m = 0; m < crm_nz; m++) {
    const double dp = st.pdel[L.at(i, pver - 1 - m)];
    for (int r = 0; r < nrad; r++) {
      const size_t o = (size_t)i + (size_t)pcols * (r + (size_t)nrad * m);
      qrad[o] = qrad[o] / dp;
    }
  }
  auto W = [](const double* p) { return const_cast<double*>(p); };
  const double ca = cos(angle[i]), sa = sin(angle[i]);
  for (int k = 0; k < pver; k++) {
    const size_t o = L.at(i, k);
    W(in.zmid)[o] = st.zm[o];
    W(in.tl)[o] = st.t[o];
    W(in.ql)[o] = st.qv[o];
    W(in.qccl)[o] = st.qcl[o];
    W(in.qiil)[o] = st.qci[o];
    W(in.pmid)[o] = st.pmid[o];
    W(in.pdel)[o] = st.pdel[o];
    W(in.ul)[o] = st.u[o] * ca + st.v[o] * sa;
    W(in.vl)[o] = st.v[o] * ca - st.u[o] * sa;
  }
  for (int k = 0; k <= pver; k++) {
    const size_t o = L.at(i, k);
    W(in.zint)[o] = st.zi[o];
    W(in.pint)[o] = st.pint[o];
  }
  W(in.ps)[i] = st.ps[i];
  W(in.phis)[i] = st.phis[i];
  W(in.ocnfrac)[i] = sf.ocnfrac[i];
  W(in.tau00)[i] = sqrt(sf.wsx[i] * sf.wsx[i] + sf.wsy[i] * sf.wsy[i]);
  const size_t bot = L.at(i, pver - 1);
  W(in.bflxls)[i] = sf.shf[i] / CPAIR + C061 * st.t[bot] * sf.lhf[i] / LATVAP;
  W(in.fluxu00)[i] = sf.wsx[i];
  W(in.fluxv00)[i] = sf.wsy[i];
  W(in.fluxt00)[i] = sf.shf[i] / CPAIR;
  W(in.fluxq00)[i] = sf.lhf[i] / LATVAP;
  W(in.wndls)[i] = sqrt(st.u[bot] * st.u[bot] + st.v[bot] * st.v[bot]);
}

// crm_physics.F90:1690-1703 (rpdel = 1/pdel)
__global__ void k_flux_bypass(int ncol, int pcols, int pver, const double* pdel, const double* shf,
                              const double* cflx, double* ps_, double* pq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  const Cols L{pcols};
  const size_t bot = L.at(i, pver - 1);
  const double g_dp = GRAVIT * (1. / pdel[bot]);
  for (int k = 0; k < pver; k++) {
    ps_[L.at(i, k)] = 0.;
    pq[L.at(i, k)] = 0.;
  }
  ps_[bot] = g_dp * shf[i];
  pq[bot] = g_dp * cflx[i];
}

// crm_physics.F90:1158-1169, 1203-1209, 1310-1330
__global__ void k_gcm_tend(int ncol, int pcols, int pver, int crm_nz, int nrad, const double* pdel,
                           crm_output_view out, crm_rad_view rad, double* qrs, double* qrl, double* pts,
                           double* pqv, double* pqcl, double* pqci, double* sp_rad_flux) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  const Cols L{pcols};
  for (int k = 0; k < pver; k++) {
    const size_t o = L.at(i, k);
    pqv[o] = out.qltend[o];
    pqcl[o] = out.qcltend[o];
    pqci[o] = out.qiltend[o];
    pts[o] = out.sltend[o];
  }
  out.precc[i] = out.precc[i] + out.precl[i];
  out.precsc[i] = out.precsc[i] + out.precsl[i];
  out.precl[i] = 0.;
  out.precsl[i] = 0.;
  // heating rate -> Q dp on the CRM levels (CRM level m sits on GCM level pver-1-m)
  double* qrad = const_cast<double*>(rad.qrad);
  for (int m = 0; m < crm_nz; m++) {
    const double dp = pdel[L.at(i, pver - 1 - m)];
    for (int r = 0; r < nrad; r++) {
      const size_t o = (size_t)i + (size_t)pcols * (r + (size_t)nrad * m);
      qrad[o] = qrad[o] * dp;
    }
  }
  for (int k = 0; k < pver; k++) {
    const size_t o = L.at(i, k);
    qrs[o] = qrs[o] / pdel[o];
    qrl[o] = qrl[o] / pdel[o];
  }
  // radiative heating above the CRM and in its top two levels
  for (int k = 0; k < pver - crm_nz + 2; k++) pts[L.at(i, k)] = qrs[L.at(i, k)] + qrl[L.at(i, k)];
  double f = 0.0;
  for (int k = 0; k < pver; k++) f = f + (qrs[L.at(i, k)] + qrl[L.at(i, k)]) * pdel[L.at(i, k)] / GRAVIT;
  sp_rad_flux[i] = f;
}

// check_energy.F90:358-561 with no rain or snow constituents (wr = ws = 0)
__global__ void k_check_energy(int ncol, int pcols, int pver, crm_gcm_state_view st, double ztodt,
                               const double* flx_vap, const double* flx_cnd, const double* flx_ice,
                               const double* flx_sen, double* te_cur, double* tw_cur, double* te_tnd_c,
                               double* tw_tnd_c, double* te_rer, double* tw_rer) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  const Cols L{pcols};
  double ke = 0., se = 0., wv = 0., wl = 0., wi = 0.;
  const double wr = 0., ws = 0.;
  for (int k = 0; k < pver; k++) {
    const size_t o = L.at(i, k);
    const double u = st.u[o], v = st.v[o], dp = st.pdel[o];
    ke = ke + 0.5 * (u * u + v * v) * dp / GRAVIT;
    se = se + st.t[o] * CPAIR * dp / GRAVIT;
    wv = wv + st.qv[o] * dp / GRAVIT;
  }
  se = se + st.phis[i] * st.ps[i] / GRAVIT;
  for (int k = 0; k < pver; k++) {
    const size_t o = L.at(i, k);
    wl = wl + st.qcl[o] * st.pdel[o] / GRAVIT;
    wi = wi + st.qci[o] * st.pdel[o] / GRAVIT;
  }
  const double te = se + ke + (LATVAP + LATICE) * wv + LATICE * (wl + wr);
  const double tw = wv + wl + wi + wr + ws;
  const double te_tnd = flx_vap[i] * (LATVAP + LATICE) - (flx_cnd[i] - flx_ice[i]) * 1000. * LATICE + flx_sen[i];
  const double tw_tnd = flx_vap[i] - flx_cnd[i] * 1000.;
  te_tnd_c[i] = te_tnd_c[i] + te_tnd;
  tw_tnd_c[i] = tw_tnd_c[i] + tw_tnd;
  const double te_xpd = te_cur[i] + te_tnd * ztodt;
  const double tw_xpd = tw_cur[i] + tw_tnd * ztodt;
  te_rer[i] = (te_xpd - te) / te_cur[i];
  tw_rer[i] = tw_cur[i] > 0. ? (tw_xpd - tw) / tw_cur[i] : 0.;
  te_cur[i] = te;
  tw_cur[i] = tw;
}

dim3 grid(int ncol) { return dim3((unsigned)((ncol + 127) / 128)); }

int finish(const crm_hip_ctx* ctx) {
  if (hipGetLastError() != hipSuccess) return CRM_ERR_DEVICE;
  return hipStreamSynchronize(crm_ctx_stream(ctx)) == hipSuccess ? CRM_OK : CRM_ERR_DEVICE;
}

// rrtmg/radiation.F90:1584-1636: one thread per (column, radiation column, level)
__global__ void k_rad_columns(int ncol, int pcols, int pver, int crm_nz, int nrad, crm_gcm_state_view st,
                              crm_gcm_cloud_view cl, crm_rad_view rad, crm_rad_column_view out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)ncol * nrad * pver;
  if (t >= n) return;
  const int i = (int)(t % ncol);
  const size_t rk = t / ncol;
  const int r = (int)(rk % nrad), k = (int)(rk / nrad);
  const size_t g = (size_t)i + (size_t)pcols * k;                  // (pcols,pver)
  const size_t o = (size_t)i + (size_t)pcols * (r + (size_t)nrad * k);  // (pcols,nrad,pver)
  const int m = pver - 1 - k;  // CRM level on GCM level k
  if (m >= crm_nz) {  // above the CRM: the GCM column's values
    out.t[o] = st.t[g];
    out.qv[o] = st.qv[g];
    out.qc[o] = st.qcl[g];
    out.qi[o] = st.qci[g];
    out.cld[o] = cl.cld[g];
    out.fice[o] = cl.fice[g];
    out.cicewp[o] = cl.cicewp[g];
    out.cliqwp[o] = cl.cliqwp[g];
    return;
  }
  const size_t a = (size_t)i + (size_t)pcols * (r + (size_t)nrad * m);  // (pcols,nrad,crm_nz)
  const double cld = rad.cld[a], qc = rad.qc[a], qi = rad.qi[a];
  out.cld[o] = cld;
  if (cld > 0) {
    const double qtot = qc + qi;
    out.fice[o] = qi / qtot;
    out.cicewp[o] = qi * st.pdel[g] / GRAVIT / fmax(0.01, cld);
    out.cliqwp[o] = qc * st.pdel[g] / GRAVIT / fmax(0.01, cld);
  } else {
    out.fice[o] = 0.0;
    out.cicewp[o] = 0.0;
    out.cliqwp[o] = 0.0;
  }
  out.qi[o] = qi;
  out.qc[o] = qc;
  out.qv[o] = fmax(1.e-9, rad.qv[a]);
  out.t[o] = rad.temperature[a];
}

// radiation.F90:2393-2398
__global__ void k_rad_qrad(int ncol, int pcols, int nrad, int crm_nz, const double* qrs, const double* qrl,
                           double* qrad) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)ncol * nrad * crm_nz) return;
  const size_t o = (size_t)(t % ncol) + (size_t)pcols * (t / ncol);
  qrad[o] = qrs[o] + qrl[o];
}

bool dims_ok(int ncol, int pcols, int pver) { return ncol >= 0 && pcols >= ncol && pver >= 1; }

}  // namespace

extern "C" {

int crm_hip_gcm_input(crm_hip_ctx* ctx, int ncol, int pcols, int pver, int crm_nz, int nrad,
                      const crm_gcm_state_view* state, const crm_gcm_surface_view* srf, const double* crm_angle,
                      const crm_input_view* in, const crm_rad_view* rad) {
  if (!ctx || !state || !srf || !crm_angle || !in || !rad || !dims_ok(ncol, pcols, pver) || crm_nz < 2 ||
      crm_nz > pver || nrad < 1)
    return CRM_ERR_ARG;
  if (ncol == 0) return CRM_OK;
  if (hipSetDevice(crm_ctx_device(ctx)) != hipSuccess) return CRM_ERR_DEVICE;
  k_gcm_input<<<grid(ncol), dim3(128), 0, crm_ctx_stream(ctx)>>>(ncol, pcols, pver, crm_nz, nrad, *state, *srf,
                                                                 crm_angle, *in, *rad);
  return finish(ctx);
}

int crm_hip_gcm_flux_bypass(crm_hip_ctx* ctx, int ncol, int pcols, int pver, const double* pdel,
                            const double* shf, const double* cflx_vap, double* ptend_s, double* ptend_qv) {
  if (!ctx || !pdel || !shf || !cflx_vap || !ptend_s || !ptend_qv || !dims_ok(ncol, pcols, pver))
    return CRM_ERR_ARG;
  if (ncol == 0) return CRM_OK;
  if (hipSetDevice(crm_ctx_device(ctx)) != hipSuccess) return CRM_ERR_DEVICE;
  k_flux_bypass<<<grid(ncol), dim3(128), 0, crm_ctx_stream(ctx)>>>(ncol, pcols, pver, pdel, shf, cflx_vap, ptend_s,
                                                                   ptend_qv);
  return finish(ctx);
}

int crm_hip_gcm_tend(crm_hip_ctx* ctx, int ncol, int pcols, int pver, int crm_nz, int nrad, const double* pdel,
                     crm_output_view* out, const crm_rad_view* rad, double* qrs, double* qrl, double* ptend_s,
                     double* ptend_qv, double* ptend_qcl, double* ptend_qci, double* sp_rad_flux) {
  if (!ctx || !pdel || !out || !rad || !qrs || !qrl || !ptend_s || !ptend_qv || !ptend_qcl || !ptend_qci ||
      !sp_rad_flux || !dims_ok(ncol, pcols, pver) || crm_nz < 2 || crm_nz > pver || nrad < 1)
    return CRM_ERR_ARG;
  if (ncol == 0) return CRM_OK;
  if (hipSetDevice(crm_ctx_device(ctx)) != hipSuccess) return CRM_ERR_DEVICE;
  k_gcm_tend<<<grid(ncol), dim3(128), 0, crm_ctx_stream(ctx)>>>(ncol, pcols, pver, crm_nz, nrad, pdel, *out, *rad,
                                                                qrs, qrl, ptend_s, ptend_qv, ptend_qcl, ptend_qci,
                                                                sp_rad_flux);
  return finish(ctx);
}

int crm_hip_check_energy_chng(crm_hip_ctx* ctx, int ncol, int pcols, int pver, const crm_gcm_state_view* state,
                              double ztodt, const double* flx_vap, const double* flx_cnd, const double* flx_ice,
                              const double* flx_sen, double* te_cur, double* tw_cur, double* te_tnd,
                              double* tw_tnd, double* te_rer, double* tw_rer) {
  if (!ctx || !state || !flx_vap || !flx_cnd || !flx_ice || !flx_sen || !te_cur || !tw_cur || !te_tnd ||
      !tw_tnd || !te_rer || !tw_rer || !dims_ok(ncol, pcols, pver))
    return CRM_ERR_ARG;
  if (ncol == 0) return CRM_OK;
  if (hipSetDevice(crm_ctx_device(ctx)) != hipSuccess) return CRM_ERR_DEVICE;
  k_check_energy<<<grid(ncol), dim3(128), 0, crm_ctx_stream(ctx)>>>(ncol, pcols, pver, *state, ztodt, flx_vap,
                                                                    flx_cnd, flx_ice, flx_sen, te_cur, tw_cur,
                                                                    te_tnd, tw_tnd, te_rer, tw_rer);
  return finish(ctx);
}

int crm_hip_rad_columns(crm_hip_ctx* ctx, int ncol, int pcols, int pver, int crm_nz, int nrad,
                        const crm_gcm_state_view* state, const crm_gcm_cloud_view* cloud, const crm_rad_view* rad,
                        const crm_rad_column_view* cols) {
  if (!ctx || !state || !cloud || !rad || !cols || !dims_ok(ncol, pcols, pver) || crm_nz < 2 || crm_nz > pver ||
      nrad < 1)
    return CRM_ERR_ARG;
  if (ncol == 0) return CRM_OK;
  if (hipSetDevice(crm_ctx_device(ctx)) != hipSuccess) return CRM_ERR_DEVICE;
  const size_t n = (size_t)ncol * nrad * pver;
  k_rad_columns<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, crm_ctx_stream(ctx)>>>(
      ncol, pcols, pver, crm_nz, nrad, *state, *cloud, *rad, *cols);
  return finish(ctx);
}

int crm_hip_rad_qrad(crm_hip_ctx* ctx, int ncol, int pcols, int nrad, int crm_nz, const double* qrs_crm,
                     const double* qrl_crm, const crm_rad_view* rad) {
  if (!ctx || !qrs_crm || !qrl_crm || !rad || !rad->qrad || ncol < 0 || pcols < ncol || nrad < 1 || crm_nz < 1)
    return CRM_ERR_ARG;
  if (ncol == 0) return CRM_OK;
  if (hipSetDevice(crm_ctx_device(ctx)) != hipSuccess) return CRM_ERR_DEVICE;
  const size_t n = (size_t)ncol * nrad * crm_nz;
  k_rad_qrad<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, crm_ctx_stream(ctx)>>>(
      ncol, pcols, nrad, crm_nz, qrs_crm, qrl_crm, const_cast<double*>(rad->qrad));
  return finish(ctx);
}

}  // extern "C"
