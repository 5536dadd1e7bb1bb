// crm_driver.hpp -- the device engine of the HIP CRM: context state, device buffers, the
// substep loop and host staging, for one bulk-field type R.  Included once per translation
// unit: crm_hip.hip instantiates it for FP64 (namespace crm), crm_hip_mixed.hip for the
// mixed-precision form with FP32 bulk fields (namespace crm_mixed).
//
// The loop mirrors crm() (crm/crm_module.F90:712-1075): per GCM substep a CFL check whose
// batch-global maximum decides the subcycle count, then per subcycle the phase-group kernels
// of crm_kernels.hpp, then the per-substep statistics; after the loop the outputs.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/crm_hip.h"
#include "../../include/crm_hip_debug.h"
#include "crm_engine.hpp"
#include "crm_kernels.hpp"

namespace crm {
// Transform operators and eigenvalues of the pressure solve, built on the host in FP64 by the
// FP64 unit for every engine (crm_hip.hip).  ny > 1: the x-y operators of a 3D CRM.
void pressure_transform_host(int nx, int ny, double dx, std::vector<double>& Ft, std::vector<double>& Gt,
                             std::vector<double>& eign);
}  // namespace crm

namespace CRM_NS {

namespace {

struct Transform {
  double *Ft = nullptr, *Gt = nullptr, *eign = nullptr;
};
struct HostTransform {
  std::vector<double> Ft, Gt, eign;
};

enum { NSCAL = 17 };  // u t qt qp qn tabs qv qcl qci qpl qpi tk tkh tke dudt x3
enum { NSCAL3 = 4 };  // 3D only: v dvdt x3
enum { NPOOL = 9 };  // blocks of the field pool (pool_sizes)
constexpr int MAX_VARIANTS = 6;  // launch variants of a kernel class (variants_of)

}  // namespace

template <typename R>
struct Engine : CrmEngine {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_ready = nullptr;  // end of a call's set-up, for a joint capture
  // device state sized for (cap_ncrms, cap_nx, cap_nzm, cap_nxr, cap_y3); cap_nx counts the
  // columns of an instance (nx * ny) and cap_nxr its radiation columns (nx_rad * ny_rad)
  int cap_ncrms = 0, cap_nx = 0, cap_nzm = 0, cap_nxr = 0;
  bool cap_y3 = false;
  char* pool = nullptr;
  size_t pool_bytes = 0;
  unsigned long long* gmax = nullptr;
  int* gcol_dev = nullptr;
  int gcol_cap = 0;
  double* lat_dev = nullptr;  // latitudes of the call's instances (Coriolis)
  int lat_cap = 0;
  int coriolis = 0;           // the reference's docoriolis switch (params.F90:107)
  std::map<std::pair<int, double>, Transform> transforms;
  // host copies of the operators, uploaded into the caller's memory by calls that allocate
  // nothing (crm_hip_run_device_noalloc)
  std::map<std::pair<int, double>, HostTransform> host_transforms;
  // staging of host views
  double* stage = nullptr;
  size_t stage_bytes = 0;
  DevT<R> dev{};
  int na = 0, nb = 1, nc = 2;
  crm_hip_stats stats{};
  bool last_ok = false;
  // per-kernel-class timing with HIP events on `stream` (crm_hip_profile_kernels)
  unsigned prof_mask = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_kid;  // kernel class of each recorded pair
  size_t ev_used = 0;       // pairs recorded in the current call
  double kt_ms[CRM_K_COUNT] = {};
  long long kt_n[CRM_K_COUNT] = {};
  // mean-state acceleration (crm_hip_accel_init)
  int acc_use = 0, acc_uv = 0;
  double acc_factor = 0.0;
  // launch schedule: threads per workgroup of each segment-kernel class (a multiple of 64 up
  // to 256; instances per workgroup = block / columns), set by the autotuner
  int block[CRM_K_COUNT] = {256, 256, 256, 256, 256, 256, 256, 256, 256, 256, 256, 256, 256, 256, 256};
  // launch variant of each kernel class (variants_of), set by the autotuner
  int variant[CRM_K_COUNT] = {};
  // HIP graphs of the steady substep (one subcycle, constant Adams-Bashforth coefficients),
  // keyed by the kernel arguments: the buffer rotation of the Adams-Bashforth scheme repeats
  // every 12 substeps, so a call reuses at most 12 graphs, and later calls reuse them too
  int use_graphs = CRM_GRAPHS_CALL;
  int debug_stop = 0;  // test hook: substep at which a captured call stops (0: none)
  // microphysics network and the recording of training pairs
  const crm_mlp* mlp = nullptr;
  float *mlp_x = nullptr, *mlp_y = nullptr;
  size_t mlp_cap = 0;  // points
  // the call's radiative heating in the sweep layout [instance][level][radiation column]
  double* qrad_t = nullptr;
  size_t qrad_cap = 0;  // elements
  float *samp_x = nullptr, *samp_t = nullptr;
  unsigned long long* samp_n = nullptr;
  long long samp_cap = 0;
  int samp_stride = 1;
  std::map<std::string, hipGraphExec_t> graphs;
  // The cache keeps at most GRAPH_CACHE_MAX graphs and evicts the least recently used: a
  // caller that passes other buffers on every call (per-chunk arrays through
  // crm_hip_run_device) would otherwise keep one whole-call graph per buffer set.  A steady
  // substep rotates through 12 argument sets, so a call needs at most 13.
  static constexpr size_t GRAPH_CACHE_MAX = 32;
  std::list<std::string> graph_lru;  // most recently used first
  long long graph_launches = 0;
  void drop_graphs() {
    if (!graphs.empty() && stream) (void)hipStreamSynchronize(stream);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
    graph_lru.clear();
  }
  hipGraphExec_t graph_get(const std::string& key) {
    auto it = graphs.find(key);
    if (it == graphs.end()) return nullptr;
    graph_lru.remove(key);
    graph_lru.push_front(key);
    return it->second;
  }
  void graph_put(const std::string& key, hipGraphExec_t exec) {
    if (graphs.size() >= GRAPH_CACHE_MAX) {
      (void)hipStreamSynchronize(stream);  // no evicted graph is still running
      while (graphs.size() >= GRAPH_CACHE_MAX && !graph_lru.empty()) {
        auto old = graphs.find(graph_lru.back());
        if (old != graphs.end()) {
          (void)hipGraphExecDestroy(old->second);
          graphs.erase(old);
        }
        graph_lru.pop_back();
      }
    }
    graphs[key] = exec;
    graph_lru.push_front(key);
  }
  int graph_cache_size() const override { return (int)graphs.size(); }

  ~Engine() override;
  int run(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
          crm_state_view* state, crm_rad_view* rad, crm_output_view* out, bool device_views,
          int phase_stop) override;
  int run_joint(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                crm_state_view* state, crm_rad_view* rad, crm_output_view* out, JointCapture* jc,
                int slot) override;
  int debug_get(const char* name, double* buf) override;
  void get_stats(crm_hip_stats* s) const override { *s = stats; }
  int profile_kernels(unsigned mask) override;
  int accel_init(int use, double factor, int uv) override {
    if (use && !(factor >= 0.0)) return CRM_ERR_ARG;
    acc_use = use;
    acc_factor = factor;
    acc_uv = uv;
    return CRM_OK;
  }
  int kernel_time(int kid, double* ms, long long* n) const override;
  int set_schedule(const int* blk) override {
    for (int k = 0; k < CRM_K_COUNT; k++)
      if (blk[k] < 64 || blk[k] > 256 || blk[k] % 64) return CRM_ERR_ARG;
    for (int k = 0; k < CRM_K_COUNT; k++) block[k] = blk[k];
    drop_graphs();
    return CRM_OK;
  }
  void get_schedule(int* blk) const override {
    for (int k = 0; k < CRM_K_COUNT; k++) blk[k] = block[k];
  }
  int set_variants(const int* v) override {
    for (int k = 0; k < CRM_K_COUNT; k++)
      if (v[k] < 0 || v[k] >= MAX_VARIANTS) return CRM_ERR_ARG;
    for (int k = 0; k < CRM_K_COUNT; k++) variant[k] = v[k];
    drop_graphs();
    return CRM_OK;
  }
  void get_variants(int* v) const override {
    for (int k = 0; k < CRM_K_COUNT; k++) v[k] = variant[k];
  }
  void set_graphs(int mode) override {
    use_graphs = mode;
    drop_graphs();
  }
  int get_graphs() const override { return use_graphs; }
  void set_debug_stop(int nstep) override { debug_stop = nstep; }
  void set_coriolis(int on) override {
    coriolis = on ? 1 : 0;
    drop_graphs();
  }
  long long graph_count() const override { return graph_launches; }
  hipStream_t get_stream() const override { return stream; }
  int set_mlp(const crm_mlp* m) override {
    if (m && crm_mlp_ns::mlp_device(m) != device) return CRM_ERR_ARG;
    mlp = m;
    return CRM_OK;
  }
  int record_samples(float* X, float* T, long long cap, int stride) override {
    if (X && (!T || cap < 0 || stride < 1)) return CRM_ERR_ARG;
    if (!samp_n && hipMalloc(&samp_n, sizeof(unsigned long long)) != hipSuccess) return CRM_ERR_NOMEM;
    samp_x = X;
    samp_t = T;
    samp_cap = cap;
    samp_stride = stride;
    return hipMemset(samp_n, 0, sizeof(unsigned long long)) == hipSuccess ? CRM_OK : CRM_ERR_DEVICE;
  }
  long long samples_count() const override {
    unsigned long long n = 0;
    if (!samp_n || hipMemcpy(&n, samp_n, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess) return 0;
    return std::min((long long)n, samp_cap);
  }
  int autotune(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
               crm_state_view* state, crm_rad_view* rad, crm_output_view* out, int* best) override;
  int run_noalloc(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                  crm_state_view* state, crm_rad_view* rad, crm_output_view* out, char* scratch,
                  size_t scratch_bytes) override;
};

// Device memory of a call that allocates nothing: every buffer of the call is carved from
// this caller-supplied region instead of the context's pool.
struct CallScratch {
  char* base;
  size_t bytes;
};

// Start timing one launch of kernel class kid; returns the pair index or -1.
template <typename R>
static long prof_begin(Engine<R>* ctx, int kid) {
  if (!((ctx->prof_mask >> kid) & 1u)) return -1;
  size_t pair = ctx->ev_used;
  while (ctx->ev_pool.size() < 2 * (pair + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    ctx->ev_pool.push_back(e);
  }
  if (ctx->ev_kid.size() < pair + 1) ctx->ev_kid.resize(pair + 1);
  ctx->ev_kid[pair] = kid;
  ctx->ev_used = pair + 1;
  (void)hipEventRecord(ctx->ev_pool[2 * pair], ctx->stream);
  return (long)pair;
}
template <typename R>
static void prof_end(Engine<R>* ctx, long pair) {
  if (pair >= 0) (void)hipEventRecord(ctx->ev_pool[2 * pair + 1], ctx->stream);
}
// After the stream has drained: fold the recorded pairs into the per-class totals.
template <typename R>
static void prof_harvest(Engine<R>* ctx) {
  for (size_t q = 0; q < ctx->ev_used; q++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev_pool[2 * q], ctx->ev_pool[2 * q + 1]) == hipSuccess) {
      ctx->kt_ms[ctx->ev_kid[q]] += ms;
      ctx->kt_n[ctx->ev_kid[q]] += 1;
    }
  }
  ctx->ev_used = 0;
}
#define TIMED(kid, stmt)                   \
  do {                                     \
    long pr_ = prof_begin(ctx, kid);       \
    stmt;                                  \
    prof_end(ctx, pr_);                    \
  } while (0)

#define HIPCHECK(x)                                                                 \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "crm_hip: %s failed: %s\n", #x, hipGetErrorString(e_));       \
      return CRM_ERR_DEVICE;                                                        \
    }                                                                               \
  } while (0)

static bool supported_nx(int nx) { return nx == 8 || nx == 16 || nx == 32 || nx == 64 || nx == 128; }
// level counts: any from 8 to 64 (the pressure solve holds one level per lane of a wave)
static bool supported_nz(int nz) { return nz >= 8 && nz <= 64; }

// REAL(4) Gamma of the reference (gammafff.c:13 via precip_init.F90:24-45)
static double gammafff(double x_in) {
  float x = (float)x_in;
  return (double)(float)std::exp(std::lgamma((double)x));
}

template <typename R>
static void host_constants(DevT<R>& d) {
  const double b_rain = (double)0.8f, b_snow = 0.25, b_grau = 0.5;
  d.gam.gamr1 = gammafff(3. + b_rain);
  d.gam.gamr2 = gammafff((5. + b_rain) / 2.);
  d.gam.gamr3 = gammafff(4. + b_rain);
  d.gam.gams1 = gammafff(3. + b_snow);
  d.gam.gams2 = gammafff((5. + b_snow) / 2.);
  d.gam.gams3 = gammafff(4. + b_snow);
  d.gam.gamg1 = gammafff(3. + b_grau);
  d.gam.gamg2 = gammafff((5. + b_grau) / 2.);
  d.gam.gamg3 = gammafff(4. + b_grau);
  const double pi = (double)3.141592653589793f;
  d.tv.crain = b_rain / 4.;
  d.tv.csnow = b_snow / 4.;
  d.tv.cgrau = b_grau / 4.;
  d.tv.vrain = 842.0 * d.gam.gamr3 / 6. / std::pow(pi * 1000.0 * 8.e6, d.tv.crain);
  d.tv.vsnow = (double)4.84f * d.gam.gams3 / 6. / std::pow(pi * 100.0 * 3.e6, d.tv.csnow);
  d.tv.vgrau = 94.5 * d.gam.gamg3 / 6. / std::pow(pi * 400.0 * 4.e6, d.tv.cgrau);
  volatile float two = 2.f;
  d.z0c1 = (double)(3.14159f / 2.f - 3.f * std::log(two));
}

static size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte sizes of the blocks of the field pool for ncap instances (each block 256-byte aligned).
template <typename R>
static void pool_sizes(int ncap, int nx, int nzm, int nxr, bool y3, size_t sizes[NPOOL]) {
  const int nz = nzm + 1;
  const size_t scal = (size_t)ncap * nzm * nx, wl = (size_t)ncap * nz * nx;
  constexpr bool mirror = !std::is_same<R, double>::value;
  sizes[0] = NSCAL * scal * sizeof(R);                        // bulk scalar fields
  sizes[1] = 4 * wl * sizeof(R);                              // w and dwdt x3
  sizes[2] = 2 * (size_t)ncap * nx * sizeof(double);          // precsfc, precssfc
  sizes[3] = (size_t)NPROF * ncap * nz * sizeof(double);      // profiles
  sizes[4] = mirror ? (size_t)NPROF * ncap * nz * sizeof(R) : 0;
  sizes[5] = (size_t)NINST * ncap * sizeof(double);
  sizes[6] = (size_t)NCOEF * ncap * nz * sizeof(R);
  sizes[7] = 5 * (size_t)ncap * nzm * nxr * sizeof(R);        // radiation sums
  sizes[8] = y3 ? NSCAL3 * scal * sizeof(R) : 0;              // v, dvdt x3
}

template <typename R>
static size_t pool_bytes_for(int ncap, int nx, int nzm, int nxr, bool y3) {
  size_t sizes[NPOOL], bytes = 0;
  pool_sizes<R>(ncap, nx, nzm, nxr, y3, sizes);
  for (size_t b : sizes) bytes += al256(b);
  return bytes;
}

// Point the field pointers of d into the pool at p (laid out by pool_sizes).
template <typename R>
static void carve_pool(DevT<R>& d, char* p, int ncap, int nx, int nzm, int nxr, bool y3) {
  size_t sizes[NPOOL];
  pool_sizes<R>(ncap, nx, nzm, nxr, y3, sizes);
  constexpr bool mirror = !std::is_same<R, double>::value;
  const int nz = nzm + 1;
  const size_t scal = (size_t)ncap * nzm * nx, wl = (size_t)ncap * nz * nx;
  auto take = [&](size_t b) {
    char* q = p;
    p += al256(b);
    return q;
  };
  R* f = (R*)take(sizes[0]);
  R** sc[NSCAL] = {&d.u, &d.t, &d.qt, &d.qp, &d.qn, &d.tabs, &d.qv, &d.qcl, &d.qci,
                   &d.qpl, &d.qpi, &d.tk, &d.tkh, &d.tke, &d.dudt[0], &d.dudt[1], &d.dudt[2]};
  for (int q = 0; q < NSCAL; q++, f += scal) *sc[q] = f;
  R* wf = (R*)take(sizes[1]);
  d.w = wf;
  for (int q = 0; q < 3; q++) d.dwdt[q] = wf + (size_t)(q + 1) * wl;
  double* ps = (double*)take(sizes[2]);
  d.precsfc = ps;
  d.precssfc = ps + (size_t)ncap * nx;
  d.prof = (double*)take(sizes[3]);
  if constexpr (mirror) {
    d.profr = (R*)take(sizes[4]);
  } else {
    take(0);
    d.profr = (R*)d.prof;
  }
  d.inst = (double*)take(sizes[5]);
  d.coef = (R*)take(sizes[6]);
  d.rad = (R*)take(sizes[7]);
  R* vf = (R*)take(sizes[8]);
  d.v = y3 ? vf : nullptr;
  for (int q = 0; q < 3; q++) d.dvdt[q] = y3 ? vf + (size_t)(q + 1) * scal : nullptr;
}

template <typename R>
static int ensure_buffers(Engine<R>* ctx, int ncrms, int nx, int nzm, int nxr, bool y3) {
  if (ncrms <= ctx->cap_ncrms && nx == ctx->cap_nx && nzm == ctx->cap_nzm && nxr == ctx->cap_nxr &&
      y3 == ctx->cap_y3)
    return CRM_OK;
  int ncap = std::max(ncrms, nx == ctx->cap_nx && nzm == ctx->cap_nzm ? ctx->cap_ncrms : 0);
  const size_t bytes = pool_bytes_for<R>(ncap, nx, nzm, nxr, y3);
  ctx->drop_graphs();  // their kernel arguments name the old buffers
  if (ctx->pool) HIPCHECK(hipFree(ctx->pool));
  ctx->pool = nullptr;
  if (hipMalloc(&ctx->pool, bytes) != hipSuccess) {
    ctx->cap_ncrms = 0;
    return CRM_ERR_NOMEM;
  }
  HIPCHECK(hipMemset(ctx->pool, 0, bytes));
  ctx->pool_bytes = bytes;
  ctx->cap_ncrms = ncap;
  ctx->cap_nx = nx;
  ctx->cap_nzm = nzm;
  ctx->cap_nxr = nxr;
  ctx->cap_y3 = y3;
  carve_pool(ctx->dev, ctx->pool, ncap, nx, nzm, nxr, y3);
  return CRM_OK;
}

template <typename R>
static int get_transform(Engine<R>* ctx, int nx, int ny, double dx, Transform& out) {
  auto key = std::make_pair(nx * 1000 + ny, dx);
  auto it = ctx->transforms.find(key);
  if (it != ctx->transforms.end()) {
    out = it->second;
    return CRM_OK;
  }
  std::vector<double> Ft, Gt, eign;
  crm::pressure_transform_host(nx, ny, dx, Ft, Gt, eign);
  const size_t n = (size_t)nx * ny;  // modes = columns
  Transform t;
  HIPCHECK(hipMalloc(&t.Ft, sizeof(double) * n * n));
  HIPCHECK(hipMalloc(&t.Gt, sizeof(double) * n * n));
  HIPCHECK(hipMalloc(&t.eign, sizeof(double) * n));
  HIPCHECK(hipMemcpy(t.Ft, Ft.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(t.Gt, Gt.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(t.eign, eign.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  ctx->transforms[key] = t;
  out = t;
  return CRM_OK;
}

// Offsets of a call's buffers inside a CallScratch region: the field pool first (as
// pool_sizes lays it out for exactly ncrms instances), then the transform operators, the
// column ids, the call's radiative heating in the sweep layout and the batch-global maxima.
struct ScratchPlan {
  size_t pool, ft, gt, eign, gcol, lat, qrad, gmax, total;
};
template <typename R>
static ScratchPlan scratch_plan(int ncrms, int nxa, int nya, int nzm, int nxr) {
  const int nx = nxa * nya;
  ScratchPlan p{};
  size_t o = pool_bytes_for<R>(ncrms, nx, nzm, nxr, nya > 1);
  p.pool = o;
  const size_t n = (size_t)nx;
  p.ft = o;
  o += al256(sizeof(double) * n * n);
  p.gt = o;
  o += al256(sizeof(double) * n * n);
  p.eign = o;
  o += al256(sizeof(double) * n);
  p.gcol = o;
  o += al256(sizeof(int) * (size_t)ncrms);
  p.lat = o;
  o += al256(sizeof(double) * (size_t)ncrms);
  p.qrad = o;
  o += al256(sizeof(double) * (size_t)ncrms * nzm * nxr);
  p.gmax = o;
  o += al256(sizeof(unsigned long long) * GMAX_WORDS);
  p.total = o;
  return p;
}

// The transform operators of (nx, ny, dx) copied into caller memory on stream s, from a host
// copy the context keeps (built on the first call of a shape).
template <typename R>
static int upload_transform(Engine<R>* ctx, int nx, int ny, double dx, Transform dst, hipStream_t s) {
  auto key = std::make_pair(nx * 1000 + ny, dx);
  auto it = ctx->host_transforms.find(key);
  if (it == ctx->host_transforms.end()) {
    HostTransform h;
    crm::pressure_transform_host(nx, ny, dx, h.Ft, h.Gt, h.eign);
    it = ctx->host_transforms.emplace(key, std::move(h)).first;
  }
  const size_t n = (size_t)nx * ny;
  HIPCHECK(hipMemcpyAsync(dst.Ft, it->second.Ft.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(dst.Gt, it->second.Gt.data(), sizeof(double) * n * n, hipMemcpyHostToDevice, s));
  HIPCHECK(hipMemcpyAsync(dst.eign, it->second.eign.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  return CRM_OK;
}

// ---- kernel dispatch over the compile-time column count ------------------------------------
// Launch variants ("tile shapes") the autotuner chooses among, per kernel class: the pressure
// solve of 2D 64-column CRMs has five (0: one instance per 4-wave workgroup, two workgroups per
// CU, the gradient pass reading its inputs again; 1: one instance on 8 waves with every level's
// inputs of the gradient pass kept in registers, one workgroup per CU; 2 and 3: variant 0 with
// two or four levels of each pass unrolled together; 4: variant 0 with the Thomas factors in
// registers, so that the workgroup needs only the slab of LDS), MPDATA six (0, 1: the default
// or the other number of levels per trip; 2, 3: the same with registers bounded for three
// waves per SIMD; 4, 5: the default sweep with each scalar on its own waves, registers
// unbounded or bounded for three waves per SIMD).  Results are bitwise the same in every
// variant of a class.
static int variants_of(int kid, int nx, bool y3) {
  if (kid == CRM_K_PRESS) return (nx == 64 && !y3) ? 5 : 1;
  if (kid == CRM_K_MPDATA) return 6;
  return 1;
}

template <int NX, typename R, typename G>
static void launch_press(int nzm, dim3 g, dim3 b, size_t sh, hipStream_t s, const DevT<R>& d, int na, int nb, int nc,
                         int variant) {
  // 28 and 58 levels (the reference's SP configurations) have compile-time level loops; any
  // other level count up to 64 runs the same kernels with the count read at run time (NZM 0)
  const bool z28 = nzm == 28, z58 = nzm == 58;
  const dim3 gi((unsigned)d.ncrms);
  if constexpr (NX == 64 && !G::Y3) {  // matrix-core transforms
    if (!z28 && !z58) {
      k_press64<0, R, G><<<gi, b, press64_lds_bytes<0>(), s>>>(d, na, nb, nc);
    } else if (variant == 0) {  // one instance per 4-wave block
      if (z28)
        k_press64<28, R, G><<<gi, b, press64_lds_bytes<28>(), s>>>(d, na, nb, nc);
      else
        k_press64<58, R, G><<<gi, b, press64_lds_bytes<58>(), s>>>(d, na, nb, nc);
    } else if (variant == 2) {  // the same, two levels per trip
      if (z28)
        k_press64<28, R, G, 2><<<gi, b, press64_lds_bytes<28>(), s>>>(d, na, nb, nc);
      else
        k_press64<58, R, G, 2><<<gi, b, press64_lds_bytes<58>(), s>>>(d, na, nb, nc);
    } else if (variant == 3) {  // the same, four levels per trip
      if (z28)
        k_press64<28, R, G, 4><<<gi, b, press64_lds_bytes<28>(), s>>>(d, na, nb, nc);
      else
        k_press64<58, R, G, 4><<<gi, b, press64_lds_bytes<58>(), s>>>(d, na, nb, nc);
    } else if (variant == 4) {  // variant 0 with the Thomas factors in registers
      if (z28)
        k_press64<28, R, G, 1, true><<<gi, b, press64_slab_bytes(), s>>>(d, na, nb, nc);
      else
        k_press64<58, R, G, 1, true><<<gi, b, press64_slab_bytes(), s>>>(d, na, nb, nc);
    } else {  // one instance per 8-wave block, inputs of the gradient pass in registers
      if (z28)
        k_press64r<28, R><<<gi, dim3(512), press64_lds_bytes<28>(), s>>>(d, na, nb, nc);
      else
        k_press64r<58, R><<<gi, dim3(512), press64_lds_bytes<58>(), s>>>(d, na, nb, nc);
    }
  } else if constexpr (NX == 64) {  // 3D 8 x 8: one instance per 4-wave block
    if (z28)
      k_press64<28, R, G><<<gi, b, press64_lds_bytes<28>(), s>>>(d, na, nb, nc);
    else if (z58)
      k_press64<58, R, G><<<gi, b, press64_lds_bytes<58>(), s>>>(d, na, nb, nc);
    else
      k_press64<0, R, G><<<gi, b, press64_lds_bytes<0>(), s>>>(d, na, nb, nc);
  } else if constexpr (NX == 128) {  // one instance per block of 4 waves
    if (z28)
      k_press128<28, R><<<gi, b, (size_t)28 * 128 * sizeof(double), s>>>(d, na, nb, nc);
    else if (z58)
      k_press128<58, R><<<gi, b, (size_t)58 * 128 * sizeof(double), s>>>(d, na, nb, nc);
    else  // the Thomas factors in LDS after the slab
      k_press128<0, R><<<gi, b, (size_t)2 * nzm * 128 * sizeof(double), s>>>(d, na, nb, nc);
  } else {
    if (z28)
      k_press<NX, 28, R, G><<<g, b, sh, s>>>(d, na, nb, nc);
    else if (z58)
      k_press<NX, 58, R, G><<<g, b, sh, s>>>(d, na, nb, nc);
    else  // the Thomas factors in LDS after the slab
      k_press<NX, 0, R, G><<<g, b, 2 * sh, s>>>(d, na, nb, nc);
  }
}

template <int NX, typename R, typename G>
static void launch_mpdata(dim3 g, dim3 b, hipStream_t s, const DevT<R>& d, int variant) {
  constexpr bool PAIR = (NX == 64 && !G::Y3);
  if (variant == 1)
    k_mpdata<NX, R, G, !PAIR><<<g, b, 0, s>>>(d);
  else if (variant == 2)
    k_mpdata_w3<NX, R, G, PAIR><<<g, b, 0, s>>>(d);
  else if (variant == 3)
    k_mpdata_w3<NX, R, G, !PAIR><<<g, b, 0, s>>>(d);
  else if (variant == 4)
    k_mpdata_split<NX, R, G, PAIR, false><<<dim3(g.x * 3), b, 0, s>>>(d, (int)g.x);
  else if (variant == 5)
    k_mpdata_split<NX, R, G, PAIR, true><<<dim3(g.x * 3), b, 0, s>>>(d, (int)g.x);
  else
    k_mpdata<NX, R, G, PAIR><<<g, b, 0, s>>>(d);
}

struct Launcher {
  int nx, nzm, ncrms;
  hipStream_t s;
  const int* block;  // threads per workgroup of each kernel class
  dim3 grid_inst(int b) const { return dim3((unsigned)((ncrms + b - 1) / b)); }
  // threads per workgroup of segment-kernel class kid: 128-column instances span two waves
  // whose exchanges are workgroup barriers, so they run one instance per workgroup
  int bs(int kid) const { return nx == 128 ? 128 : block[kid]; }
  // grid and workgroup of segment-kernel class kid
  dim3 g(int kid) const { return dim3((unsigned)(((size_t)ncrms * nx + bs(kid) - 1) / bs(kid))); }
  dim3 b(int kid) const { return dim3((unsigned)bs(kid)); }
};

// NXV: columns per instance; GEO: the lane geometry (a 2D row of NXV columns, or the 8 x 8
// or 4 x 4 columns of a 3D CRM; nx here counts nx*ny).
#define NX_DISPATCH(nx, CALL)       \
  if (y3) {                         \
    if ((nx) == 16) {               \
      constexpr int NXV = 16;       \
      using GEO = Seg3<4>;          \
      CALL;                         \
    } else {                        \
      constexpr int NXV = 64;       \
      using GEO = Seg3<8>;          \
      CALL;                         \
    }                               \
  } else                            \
    switch (nx) {                   \
      case 8: {                     \
        constexpr int NXV = 8;      \
        using GEO = Seg<8>;         \
        CALL;                       \
      } break;                      \
      case 16: {                    \
        constexpr int NXV = 16;     \
        using GEO = Seg<16>;        \
        CALL;                       \
      } break;                      \
      case 32: {                    \
        constexpr int NXV = 32;     \
        using GEO = Seg<32>;        \
        CALL;                       \
      } break;                      \
      case 128: {                   \
        constexpr int NXV = 128;    \
        using GEO = Seg128;         \
        CALL;                       \
      } break;                      \
      default: {                    \
        constexpr int NXV = 64;     \
        using GEO = Seg<64>;        \
        CALL;                       \
      } break;                      \
    }

static void abcoefs(int nstep, const double* dt3, int na, int nb, int nc, double& at, double& bt, double& ct) {
  if (nstep >= 3) {
    double alpha = dt3[nb] / dt3[na], beta = dt3[nc] / dt3[na];
    ct = (2. + 3. * alpha) / (6. * (alpha + beta) * beta);
    bt = -(1. + 2. * (alpha + beta) * ct) / (2. * alpha);
    at = 1. - bt - ct;
  } else if (nstep >= 2) {
    at = 3. / 2.;
    bt = -1. / 2.;
    ct = 0.;
  } else {
    at = 1.;
    bt = 0.;
    ct = 0.;
  }
}

static int validate(const crm_dims* dm) {
  if (!dm || dm->ncrms < 1 || dm->ld < dm->ncrms || dm->plev < dm->nz + 1) return CRM_ERR_ARG;
  if (!supported_nz(dm->nz)) return CRM_ERR_UNSUPPORTED;
  if (dm->ny != 1) {
    // 3D: 8 x 8 or 4 x 4 columns (dy == dx); radiation groups of any power-of-two shape
    if (!((dm->nx == 8 && dm->ny == 8) || (dm->nx == 4 && dm->ny == 4))) return CRM_ERR_UNSUPPORTED;
    if (dm->nx_rad < 1 || dm->ny_rad < 1 || dm->nx % dm->nx_rad || dm->ny % dm->ny_rad) return CRM_ERR_NXRAD;
    const int gx = dm->nx / dm->nx_rad, gy = dm->ny / dm->ny_rad;
    if ((gx & (gx - 1)) || (gy & (gy - 1))) return CRM_ERR_UNSUPPORTED;
    if (60. < 2 * dm->dt) return CRM_ERR_DAMPING;
    return CRM_OK;
  }
  if (dm->ny_rad != 1) return CRM_ERR_NXRAD;
  if (!supported_nx(dm->nx)) return CRM_ERR_UNSUPPORTED;
  if (dm->nx_rad < 1 || dm->nx % dm->nx_rad != 0) return CRM_ERR_NXRAD;
  if ((dm->nx / dm->nx_rad) & ((dm->nx / dm->nx_rad) - 1)) return CRM_ERR_UNSUPPORTED;
  if (60. < 2 * dm->dt) return CRM_ERR_DAMPING;
  return CRM_OK;
}

// The whole call on device-resident views.  phase_stop > 0 stops after that phase group of
// the last substep (debug entry only).
template <typename R>
static int run_core(Engine<R>* ctx, const crm_dims* dm, const crm_column_info* col, double dt_gl,
                    const crm_input_view* in, crm_state_view* st, crm_rad_view* rad, crm_output_view* out,
                    int phase_stop, const CallScratch* sc = nullptr, JointCapture* jc = nullptr,
                    int slot = -1) {
  // as a class of a joint capture: arrive exactly once, without a sequence on every path that
  // does not replay a whole-call graph (so the last class to arrive is never left waiting)
  struct Arrival {
    JointCapture* jc;
    int slot;
    bool done = false;
    int alone() {
      if (!jc || done) return CRM_OK;
      done = true;
      return jc->arrive(slot, false, nullptr, nullptr, std::string(), nullptr);
    }
    ~Arrival() { (void)alone(); }
  } arrival{jc, slot};
  int rc = validate(dm);
  if (rc) return rc;
  const bool y3 = dm->ny > 1;
  const int nx = dm->nx * dm->ny, nzm = dm->nz, ncrms = dm->ncrms;  // nx: columns per instance
  const int nxr = dm->nx_rad * dm->ny_rad;
  // set-up (allocations, null-stream clears and uploads) excludes captures of other threads
  hipStream_t s = ctx->stream;
  std::unique_lock<std::mutex> setup_lock(crm_device_mutex());
  Transform tf;
  DevT<R> dscr{};  // the kernel arguments of a call in caller memory (sc)
  unsigned long long* gmax = ctx->gmax;
  int* gcol_dev = nullptr;
  double* lat_buf = nullptr;
  double* qrad_buf = nullptr;
  // Coriolis terms (coriolis.F90) need the instances' latitudes; in 2D they would also need the
  // v the 2D form does not carry
  if (ctx->coriolis && (dm->ny == 1 || !col || !col->lat)) return dm->ny == 1 ? CRM_ERR_UNSUPPORTED : CRM_ERR_ARG;
  // a call in caller memory runs plain launches without events: no graph or event is created
  struct MaskRestore {
    unsigned& m;
    unsigned saved;
    ~MaskRestore() { m = saved; }
  } mask_restore{ctx->prof_mask, ctx->prof_mask};
  if (sc) {
    if (ctx->mlp || ctx->samp_x) return CRM_ERR_UNSUPPORTED;
    const ScratchPlan pl = scratch_plan<R>(ncrms, dm->nx, dm->ny, nzm, nxr);
    if (pl.total > sc->bytes) return CRM_ERR_SCRATCH;
    ctx->prof_mask = 0;
    char* b = sc->base;
    carve_pool(dscr, b, ncrms, nx, nzm, nxr, y3);
    tf.Ft = (double*)(b + pl.ft);
    tf.Gt = (double*)(b + pl.gt);
    tf.eign = (double*)(b + pl.eign);
    gcol_dev = (int*)(b + pl.gcol);
    lat_buf = (double*)(b + pl.lat);
    qrad_buf = (double*)(b + pl.qrad);
    gmax = (unsigned long long*)(b + pl.gmax);
    // the region keeps the last call's contents: the pool and the maxima start from zero, as
    // in a freshly allocated context
    k_zero_u64<<<dim3(2048), dim3(256), 0, s>>>((unsigned long long*)b, pl.pool / 8);
    k_zero_u64<<<dim3(1), dim3(256), 0, s>>>(gmax, GMAX_WORDS);
    rc = upload_transform(ctx, dm->nx, dm->ny, dm->dx, tf, s);
    if (rc) return rc;
  } else {
    rc = ensure_buffers(ctx, ncrms, nx, nzm, nxr, y3);
    if (rc) return rc;
    rc = get_transform(ctx, dm->nx, dm->ny, dm->dx, tf);
    if (rc) return rc;
  }
  DevT<R>& d = sc ? dscr : ctx->dev;
  d.ncrms = ncrms;
  d.nzm = nzm;
  d.nz = nzm + 1;
  d.plev = dm->plev;
  d.ld = dm->ld;
  d.nx_rad = nxr;
  d.ny = dm->ny;
  d.rad_row = dm->nx;
  d.rad_nxx = dm->nx_rad;
  d.rad_gx = dm->nx / dm->nx_rad;
  d.rad_gy = dm->ny / dm->ny_rad;
  d.dx = dm->dx;
  d.dt = dm->dt;
  d.Ft = tf.Ft;
  d.Gt = tf.Gt;
  d.eign = tf.eign;
  d.gmax = gmax;
  d.docoriolis = ctx->coriolis;
  host_constants(d);
  Launcher L{nx, nzm, ncrms, s, ctx->block};
  const dim3 b256(256);
  ctx->last_ok = false;
  ctx->stats = crm_hip_stats{};

  if (ctx->coriolis) {
    if (!sc) {
      if (ctx->lat_cap < ncrms) {
        if (ctx->lat_dev) HIPCHECK(hipFree(ctx->lat_dev));
        ctx->lat_dev = nullptr;
        ctx->lat_cap = 0;
        if (hipMalloc(&ctx->lat_dev, sizeof(double) * ncrms) != hipSuccess) return CRM_ERR_NOMEM;
        ctx->lat_cap = ncrms;
      }
      lat_buf = ctx->lat_dev;
    }
    HIPCHECK(hipMemcpyAsync(lat_buf, col->lat, sizeof(double) * ncrms, hipMemcpyHostToDevice, s));
  }
  TIMED(CRM_K_SETUP, (k_setup<R><<<L.grid_inst(64), dim3(64), 0, s>>>(d, *in, ctx->coriolis ? lat_buf : nullptr)));
  TIMED(CRM_K_INIT, NX_DISPATCH(nx, (k_init<NXV, R><<<L.g(CRM_K_INIT), L.b(CRM_K_INIT), 0, s>>>(d, *in, *st, *out, dt_gl))));
  if (col && col->igstep <= 1) {
    if (!sc) {
      if (ctx->gcol_cap < ncrms) {
        if (ctx->gcol_dev) HIPCHECK(hipFree(ctx->gcol_dev));
        HIPCHECK(hipMalloc(&ctx->gcol_dev, sizeof(int) * ncrms));
        ctx->gcol_cap = ncrms;
      }
      gcol_dev = ctx->gcol_dev;
    }
    HIPCHECK(hipMemcpyAsync(gcol_dev, col->gcol, sizeof(int) * ncrms, hipMemcpyHostToDevice, s));
    TIMED(CRM_K_SETPERTURB, NX_DISPATCH(nx, (k_setperturb<NXV, R><<<L.grid_inst(64), dim3(64), 0, s>>>(d, gcol_dev))));
  }
  if constexpr (!std::is_same<R, double>::value) {
    // FP32 copies of the level profiles for the bulk kernels
    const size_t np = (size_t)NPROF * ncrms * (nzm + 1);
    k_prof_mirror<R><<<dim3((unsigned)((np + 255) / 256)), b256, 0, s>>>(d.prof, d.profr, np);
  }
  // crm_rad%qrad is constant over the call: one transpose from the caller's (pcols, nx_rad,
  // ny_rad, nz) order, so every substep's k_pre reads it contiguously
  {
    const size_t nq = (size_t)ncrms * nzm * d.nx_rad;
    if (!sc) {
      if (ctx->qrad_cap < nq) {
        if (ctx->qrad_t) HIPCHECK(hipFree(ctx->qrad_t));
        ctx->qrad_t = nullptr;
        ctx->qrad_cap = 0;
        if (hipMalloc(&ctx->qrad_t, nq * sizeof(double)) != hipSuccess) return CRM_ERR_NOMEM;
        ctx->qrad_cap = nq;
      }
      qrad_buf = ctx->qrad_t;
    }
    k_qrad_order<<<dim3((unsigned)((nq + 255) / 256)), b256, 0, s>>>(rad->qrad, qrad_buf, ncrms, d.ld, nzm,
                                                                     d.nx_rad);
  }
  HIPCHECK(hipGetLastError());
  k_zero_u64<<<dim3(1), dim3(64), 0, s>>>(gmax + GMAX_NPREC, 3);

  if (!(dt_gl >= dm->dt) || !(dt_gl / dm->dt < 1.0e6)) return CRM_ERR_ARG;
  int nstop = (int)(dt_gl / dm->dt);
  const double dt = dt_gl / nstop;
  // mean-state acceleration: crm_accel_nstop (accelerate_crm.F90:78-106, crm_module.F90:718-720)
  const bool acc = ctx->acc_use != 0;
  const double acc_f = ctx->acc_factor;
  bool ceased = false;
  if (acc) {
    const int a = (int)(1 + acc_f);
    if (a < 1 || nstop % a != 0) return CRM_ERR_ACCEL;
    nstop = (int)(nstop / (1 + acc_f));
  }
  d.dt = dt;
  // microphysics network / training-pair recording: per-point inputs (and outputs) of every
  // subcycle
  const size_t npts = (size_t)ncrms * nzm * nx;
  const bool net_inputs = ctx->mlp != nullptr || ctx->samp_x != nullptr;
  if (net_inputs && ctx->mlp_cap < npts) {
    if (ctx->mlp_x) HIPCHECK(hipFree(ctx->mlp_x));
    if (ctx->mlp_y) HIPCHECK(hipFree(ctx->mlp_y));
    ctx->mlp_x = ctx->mlp_y = nullptr;
    ctx->mlp_cap = 0;
    if (hipMalloc(&ctx->mlp_x, npts * CRM_MLP_NIN * sizeof(float)) != hipSuccess ||
        hipMalloc(&ctx->mlp_y, npts * 8 * sizeof(float)) != hipSuccess)
      return CRM_ERR_NOMEM;
    ctx->mlp_cap = npts;
  }
  d.mlp_x = net_inputs ? ctx->mlp_x : nullptr;
  d.mlp_y = ctx->mlp ? ctx->mlp_y : nullptr;
  d.samp_x = ctx->samp_x;
  d.samp_t = ctx->samp_t;
  d.samp_n = ctx->samp_n;
  d.samp_cap = ctx->samp_cap;
  d.samp_stride = ctx->samp_stride;
  double dt3[3] = {0.0, 0.0, 0.0};
  int na = 0, nb = 1, nc = 2;
  // k_press: the 64/nx instances of one wave per block (k_press64, k_press128: one per block)
  const int press_per_block = nx < 64 ? 64 / nx : 1;
  const dim3 gpress((unsigned)((ncrms + press_per_block - 1) / press_per_block));
  const size_t press_lds = (size_t)nzm * 64 * sizeof(double);
  unsigned long long err_flag = 0;
  unsigned long long tail[2] = {0, 0};  // read back after the call: the largest nprec, the error flag
  HIPCHECK(hipEventRecord(ctx->ev0, s));
  // A mean t change above 5 K flagged by the acceleration during substep n stops it for the
  // rest of the call and lengthens it (accelerate_crm.F90:255-278); the flag is read with the
  // next CFL maximum, or after the last substep.
  auto cease = [&](unsigned long long flag, int nstep_flagged) {
    if (acc && !ceased && flag) {
      ceased = true;
      nstop = (int)(nstop + (nstop - nstep_flagged + 1) * acc_f);
    }
  };
  setup_lock.unlock();
  int nstep = 0;
  d.stop = nullptr;
  // Whole-call graph: the call's substeps captured once with one subcycle each and replayed;
  // k_decide stops the replay at the first substep that needs more subcycles (every later
  // kernel returns at entry), and the loop below resumes that substep from the host's record
  // of its bookkeeping before it (slot rotation, buffer swaps, Adams-Bashforth history).
  if (ctx->use_graphs >= CRM_GRAPHS_CALL && !sc && !acc && !net_inputs && ctx->prof_mask == 0 && phase_stop == 0 &&
      nstop <= 2048) {
    struct Snap {
      int na, nb, nc;
      DevT<R> d;
      double dt3[3];
    };
    std::vector<Snap> snap((size_t)nstop + 1);
    const double* qrad = qrad_buf;
    d.stop = gmax + GMAX_STOP;
    DevT<R> dk;  // the key: the call's starting arguments
    std::memcpy(&dk, &d, sizeof(d));
    std::string key = "call";
    key.append((const char*)&dk, sizeof(dk));
    const int ints[6] = {na, nb, nc, nx, nstop, ctx->debug_stop};
    key.append((const char*)ints, sizeof(ints));
    key.append((const char*)dt3, sizeof(dt3));
    key.append((const char*)&qrad, sizeof(qrad));
    key.append((const char*)ctx->block, sizeof(ctx->block));
    key.append((const char*)ctx->variant, sizeof(ctx->variant));
    // the call's sequence; with launch false only the host bookkeeping is replayed
    auto sequence = [&](bool launch) {
      for (int n = 1; n <= nstop; n++) {
        snap[n] = Snap{na, nb, nc, d, {dt3[0], dt3[1], dt3[2]}};
        d.nstep = n;
        d.icycle = 1;
        d.dtn = dt;
        dt3[na] = d.dtn;
        d.dtfactor = 1.0;
        abcoefs(n, dt3, na, nb, nc, d.at, d.bt, d.ct);
        d.dt3na = dt3[na];
        if (launch) {
          // the kurant maximum of substep n: from k_kurant at the first substep, after it from
          // the previous substep's k_micro (the same state, the same maximum); the slots are
          // cleared once k_decide has read them, for this substep's k_micro (the last
          // substep's maximum is not read)
          if (n == 1) {
            k_clear_u64<<<1, 64, 0, s>>>(gmax + GMAX_FLAGS, d.stop);
            NX_DISPATCH(nx, (k_kurant<NXV, R><<<L.g(CRM_K_KURANT), L.b(CRM_K_KURANT), 0, s>>>(d)));
          }
          k_decide<<<1, 64, 0, s>>>(gmax, n, ctx->debug_stop);
          k_clear_u64<<<1, 64, 0, s>>>(gmax + GMAX_FLAGS, d.stop);
          k_precip_init<R><<<dim3((unsigned)(((size_t)ncrms * nzm + 127) / 128)), dim3(128), 0, s>>>(d);
          NX_DISPATCH(nx, (k_pre<NXV, R><<<L.g(CRM_K_PRE), L.b(CRM_K_PRE), 0, s>>>(d, na, qrad)));
          NX_DISPATCH(nx, (k_sgs<NXV, R, GEO><<<L.g(CRM_K_SGS), L.b(CRM_K_SGS), 0, s>>>(d)));
          NX_DISPATCH(nx, (k_mom<NXV, R, GEO><<<L.g(CRM_K_MOM), L.b(CRM_K_MOM), 0, s>>>(d, na)));
          NX_DISPATCH(nx, (launch_press<NXV, R, GEO>(nzm, gpress, b256, press_lds, s, d, na, nb, nc, ctx->variant[CRM_K_PRESS])));
          k_clear_u64<<<1, 64, 0, s>>>(gmax + GMAX_PCFL, d.stop);
          NX_DISPATCH(nx, (launch_mpdata<NXV, R, GEO>(L.g(CRM_K_MPDATA), L.b(CRM_K_MPDATA), s, d, ctx->variant[CRM_K_MPDATA])));
          NX_DISPATCH(nx, (k_precip<NXV, R><<<L.g(CRM_K_PRECIP), L.b(CRM_K_PRECIP), 0, s>>>(d)));
        }
        std::swap(d.u, d.dudt[nc]);
        std::swap(d.w, d.dwdt[nc]);
        if (y3) std::swap(d.v, d.dvdt[nc]);
        if (launch) {
          NX_DISPATCH(nx, (k_diffuse<NXV, R, GEO><<<L.g(CRM_K_DIFFUSE), L.b(CRM_K_DIFFUSE), 0, s>>>(d)));
          NX_DISPATCH(nx, (k_micro<NXV, R, true><<<L.g(CRM_K_MICRO), L.b(CRM_K_MICRO), 0, s>>>(d, 0)));
        }
        int nn = na;
        na = nc;
        nc = nb;
        nb = nn;
        if (launch) {
          NX_DISPATCH(nx, (k_stats<NXV, R><<<L.g(CRM_K_STATS), L.b(CRM_K_STATS), 0, s>>>(d)));
        }
      }
    };
    HIPCHECK(hipMemsetAsync(gmax + GMAX_STOP, 0, sizeof(unsigned long long), s));
    if (jc) {  // one graph with the other classes of crm_hip_run_many
      HIPCHECK(hipEventRecord(ctx->ev_ready, s));
      arrival.done = true;
      rc = jc->arrive(slot, true, s, ctx->ev_ready, key, [&](bool launch) { sequence(launch); });
      if (rc) return rc;
    } else {
      hipGraphExec_t exec = ctx->graph_get(key);
      if (!exec) {
        std::lock_guard<std::mutex> capture_lock(crm_device_mutex());
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        sequence(true);
        HIPCHECK(hipStreamEndCapture(s, &graph));
        HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        ctx->graph_put(key, exec);
      } else {
        sequence(false);
      }
      HIPCHECK(hipGraphLaunch(exec, s));
    }
    ctx->graph_launches += nstop;
    unsigned long long stopped_at = 0;
    HIPCHECK(hipMemcpyAsync(&stopped_at, gmax + GMAX_STOP, sizeof(stopped_at), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (stopped_at == 0) {
      nstep = nstop;
    } else {  // resume substep stopped_at as it was before it
      const Snap& sn = snap[stopped_at];
      na = sn.na;
      nb = sn.nb;
      nc = sn.nc;
      d = sn.d;
      for (int j = 0; j < 3; j++) dt3[j] = sn.dt3[j];
      nstep = (int)stopped_at - 1;
      ctx->graph_launches -= nstop - nstep;
    }
    if (nstep >= 1) ctx->stats.ncycle_max = std::max(ctx->stats.ncycle_max, 1);
    d.stop = nullptr;
  }
  rc = arrival.alone();
  if (rc) return rc;
  while (true) {
    if (nstep >= nstop) {
      if (acc && !ceased) {
        unsigned long long flag = 0;
        HIPCHECK(hipMemcpyAsync(&flag, gmax + 3, sizeof(flag), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        cease(flag, nstep);
      }
      if (nstep >= nstop) break;
    }
    nstep++;
    d.nstep = nstep;
    unsigned long long bits[GMAX_PCFL];
    k_zero_u64<<<dim3(2), dim3(256), 0, s>>>(gmax + GMAX_FLAGS, GMAX_PCFL - GMAX_FLAGS);
    TIMED(CRM_K_KURANT, NX_DISPATCH(nx, (k_kurant<NXV, R><<<L.g(CRM_K_KURANT), L.b(CRM_K_KURANT), 0, s>>>(d))));
    HIPCHECK(hipMemcpyAsync(bits, gmax, sizeof(bits), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    cease(bits[3], nstep - 1);
    // non-negative doubles order as their bit patterns
    unsigned long long cbits = 0;
    for (int j = 0; j < KURANT_SLOTS; j++) cbits = std::max(cbits, bits[GMAX_FLAGS + j * KURANT_STRIDE]);
    double cfl;
    std::memcpy(&cfl, &cbits, sizeof(cfl));
    // ncycle = max(1, ceiling(cfl / 0.7)) > 16 is an abort in the reference (kurant.F90:53-75);
    // the test is made before the conversion so that huge or non-finite values cannot overflow
    const double q = cfl / (double)0.7f;
    if (q != q) return CRM_ERR_NAN;
    if (!(q <= 16.0)) return CRM_ERR_CFL;
    int ncycle = std::max(1, (int)std::ceil(q));
    ctx->stats.ncycle_max = std::max(ctx->stats.ncycle_max, ncycle);
    const bool last = (nstep == nstop);
    // Steady substeps (one subcycle, constant Adams-Bashforth coefficients from the third
    // substep on) replay a captured HIP graph of the whole substep, statistics included.
    if (ctx->use_graphs != CRM_GRAPHS_OFF && !sc && ncycle == 1 && nstep >= 3 && !acc && !net_inputs &&
        ctx->prof_mask == 0 && phase_stop == 0) {
      d.icycle = 1;
      d.dtn = dt;
      dt3[na] = d.dtn;
      d.dtfactor = 1.0;
      abcoefs(nstep, dt3, na, nb, nc, d.at, d.bt, d.ct);
      d.dt3na = dt3[na];
      const double* qrad = qrad_buf;
      auto body = [&]() {
        k_precip_init<R><<<dim3((unsigned)(((size_t)ncrms * nzm + 127) / 128)), dim3(128), 0, s>>>(d);
        NX_DISPATCH(nx, (k_pre<NXV, R><<<L.g(CRM_K_PRE), L.b(CRM_K_PRE), 0, s>>>(d, na, qrad)));
        NX_DISPATCH(nx, (k_sgs<NXV, R, GEO><<<L.g(CRM_K_SGS), L.b(CRM_K_SGS), 0, s>>>(d)));
        NX_DISPATCH(nx, (k_mom<NXV, R, GEO><<<L.g(CRM_K_MOM), L.b(CRM_K_MOM), 0, s>>>(d, na)));
        NX_DISPATCH(nx, (launch_press<NXV, R, GEO>(nzm, gpress, b256, press_lds, s, d, na, nb, nc, ctx->variant[CRM_K_PRESS])));
        k_clear_u64<<<1, 64, 0, s>>>(gmax + GMAX_PCFL, d.stop);
        NX_DISPATCH(nx, (launch_mpdata<NXV, R, GEO>(L.g(CRM_K_MPDATA), L.b(CRM_K_MPDATA), s, d, ctx->variant[CRM_K_MPDATA])));
        NX_DISPATCH(nx, (k_precip<NXV, R><<<L.g(CRM_K_PRECIP), L.b(CRM_K_PRECIP), 0, s>>>(d)));
        std::swap(d.u, d.dudt[nc]);
        std::swap(d.w, d.dwdt[nc]);
        if (y3) std::swap(d.v, d.dvdt[nc]);
        NX_DISPATCH(nx, (k_diffuse<NXV, R, GEO><<<L.g(CRM_K_DIFFUSE), L.b(CRM_K_DIFFUSE), 0, s>>>(d)));
        NX_DISPATCH(nx, (k_micro<NXV, R><<<L.g(CRM_K_MICRO), L.b(CRM_K_MICRO), 0, s>>>(d, 0)));
        int nn = na;
        na = nc;
        nc = nb;
        nb = nn;
        NX_DISPATCH(nx, (k_stats<NXV, R><<<L.g(CRM_K_STATS), L.b(CRM_K_STATS), 0, s>>>(d)));
      };
      DevT<R> dk;  // the key: every kernel argument of the substep
      std::memcpy(&dk, &d, sizeof(d));
      dk.nstep = 0;
      dk.icycle = 0;
      std::string key((const char*)&dk, sizeof(dk));
      const int ints[4] = {na, nb, nc, nx};
      key.append((const char*)ints, sizeof(ints));
      key.append((const char*)&qrad, sizeof(qrad));
      key.append((const char*)ctx->block, sizeof(ctx->block));
      key.append((const char*)ctx->variant, sizeof(ctx->variant));
    key.append((const char*)ctx->variant, sizeof(ctx->variant));
      hipGraphExec_t exec = ctx->graph_get(key);
      if (!exec) {
        std::lock_guard<std::mutex> capture_lock(crm_device_mutex());
        hipGraph_t graph = nullptr;
        HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        body();
        HIPCHECK(hipStreamEndCapture(s, &graph));
        HIPCHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        ctx->graph_put(key, exec);
      } else {
        // the host side of the substep: buffer swap of uvw and the index rotation
        std::swap(d.u, d.dudt[nc]);
        std::swap(d.w, d.dwdt[nc]);
        if (y3) std::swap(d.v, d.dvdt[nc]);
        int nn = na;
        na = nc;
        nc = nb;
        nb = nn;
      }
      HIPCHECK(hipGraphLaunch(exec, s));
      ctx->graph_launches++;
      continue;
    }
    for (int icyc = 1; icyc <= ncycle; icyc++) {
      d.icycle = icyc;
      d.dtn = dt / ncycle;
      dt3[na] = d.dtn;
      d.dtfactor = d.dtn / dt;
      abcoefs(nstep, dt3, na, nb, nc, d.at, d.bt, d.ct);
      d.dt3na = dt3[na];
      if (icyc == 1)
        TIMED(CRM_K_PRECIP_INIT,
              (k_precip_init<R><<<dim3((unsigned)(((size_t)ncrms * nzm + 127) / 128)), dim3(128), 0, s>>>(d)));
      const double* qrad = qrad_buf;
      TIMED(CRM_K_PRE, NX_DISPATCH(nx, (k_pre<NXV, R><<<L.g(CRM_K_PRE), L.b(CRM_K_PRE), 0, s>>>(d, na, qrad))));
      if (last && phase_stop == 7) goto stopped;
      TIMED(CRM_K_SGS, NX_DISPATCH(nx, (k_sgs<NXV, R, GEO><<<L.g(CRM_K_SGS), L.b(CRM_K_SGS), 0, s>>>(d))));
      if (last && phase_stop == 9) goto stopped;
      TIMED(CRM_K_MOM, NX_DISPATCH(nx, (k_mom<NXV, R, GEO><<<L.g(CRM_K_MOM), L.b(CRM_K_MOM), 0, s>>>(d, na))));
      if (last && phase_stop == 11) goto stopped;
      TIMED(CRM_K_PRESS, NX_DISPATCH(nx, (launch_press<NXV, R, GEO>(nzm, gpress, b256, press_lds, s, d, na, nb, nc, ctx->variant[CRM_K_PRESS]))));
      if (last && phase_stop == 13) goto stopped;
      k_clear_u64<<<1, 64, 0, s>>>(gmax + GMAX_PCFL, d.stop);
      TIMED(CRM_K_MPDATA, NX_DISPATCH(nx, (launch_mpdata<NXV, R, GEO>(L.g(CRM_K_MPDATA), L.b(CRM_K_MPDATA), s, d, ctx->variant[CRM_K_MPDATA]))));
      TIMED(CRM_K_PRECIP, NX_DISPATCH(nx, (k_precip<NXV, R><<<L.g(CRM_K_PRECIP), L.b(CRM_K_PRECIP), 0, s>>>(d))));
      if (last && phase_stop == 15) goto stopped;
      std::swap(d.u, d.dudt[nc]);
      std::swap(d.w, d.dwdt[nc]);
      if (y3) std::swap(d.v, d.dvdt[nc]);
      if (last && phase_stop == 17) goto stopped;
      TIMED(CRM_K_DIFFUSE, NX_DISPATCH(nx, (k_diffuse<NXV, R, GEO><<<L.g(CRM_K_DIFFUSE), L.b(CRM_K_DIFFUSE), 0, s>>>(d))));
      if (last && phase_stop == 18) goto stopped;
      {
        const int accel = (acc && !ceased) ? 1 : 0;
        if (net_inputs)
          TIMED(CRM_K_MICRO, NX_DISPATCH(nx, (k_mlp_inputs<NXV, R><<<dim3((unsigned)((npts + 255) / 256)), b256, 0, s>>>(d, ctx->mlp_x))));
        if (ctx->mlp) TIMED(CRM_K_MICRO, crm_mlp_ns::mlp_infer(s, ctx->mlp, ctx->mlp_x, (long)npts, ctx->mlp_y));
        TIMED(CRM_K_MICRO, NX_DISPATCH(nx, (k_micro<NXV, R><<<L.g(CRM_K_MICRO), L.b(CRM_K_MICRO), 0, s>>>(d, accel))));
        if (accel && last && phase_stop == 19) goto stopped;  // before the acceleration
        if (accel)  // crm_module.F90:872-876
          TIMED(CRM_K_MICRO, NX_DISPATCH(nx, (k_accel<NXV, R><<<L.g(CRM_K_MICRO), L.b(CRM_K_MICRO), 0, s>>>(d, acc_f, ctx->acc_uv))));
      }
      if (last && phase_stop == 20) goto stopped;
      int nn = na;
      na = nc;
      nc = nb;
      nb = nn;
    }
    TIMED(CRM_K_STATS, NX_DISPATCH(nx, (k_stats<NXV, R><<<L.g(CRM_K_STATS), L.b(CRM_K_STATS), 0, s>>>(d))));
    HIPCHECK(hipGetLastError());
  }
  ctx->stats.nstop = nstop;
  ctx->stats.accel_ceased = ceased ? 1 : 0;
  TIMED(CRM_K_FINAL, NX_DISPATCH(nx, (k_final<NXV, R><<<L.g(CRM_K_FINAL), L.b(CRM_K_FINAL), 0, s>>>(d, *in, *st, *rad, *out, dt_gl, nstop))));
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipEventRecord(ctx->ev1, s));
  HIPCHECK(hipMemcpyAsync(tail, gmax + GMAX_NPREC, sizeof(tail), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  err_flag = tail[1];
  ctx->stats.nprec_max = (int)tail[0];
  {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.ms_device = ms;
  }
  prof_harvest(ctx);
  ctx->na = na;
  ctx->nb = nb;
  ctx->nc = nc;
  if (err_flag) return err_flag == ERR_FLAG_NAN ? CRM_ERR_NAN : CRM_ERR_CFL;
  ctx->last_ok = true;
  return CRM_OK;
stopped:
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(s));
  prof_harvest(ctx);
  ctx->na = na;
  ctx->nb = nb;
  ctx->nc = nc;
  return CRM_OK;
}

// ---- staging of host views ---------------------------------------------------------------
namespace {
struct ViewField {
  const double* host;
  double* dev;
  size_t n;
  bool in, outw;
};
}  // namespace

template <typename R>
static int stage_views(Engine<R>* ctx, const crm_dims* dm, const crm_input_view* in, crm_state_view* st,
                       crm_rad_view* rad, crm_output_view* out, crm_input_view& din, crm_state_view& dst,
                       crm_rad_view& drad, crm_output_view& dout, std::vector<ViewField>& fields) {
  const size_t ld = dm->ld, plev = dm->plev, g4 = ld * dm->nx * dm->ny * dm->nz, g3 = ld * dm->nx * dm->ny;
  const size_t r4 = ld * dm->nx_rad * dm->ny_rad * dm->nz;
  fields.clear();
  auto add = [&](const double* h, size_t n, bool inb, bool outw) {
    fields.push_back({h, nullptr, n, inb, outw});
  };
  const double* ip[] = {in->zmid, in->zint, in->tl, in->ql, in->qccl, in->qiil, in->ps, in->pmid, in->pint,
                        in->pdel, in->phis, in->ul, in->vl, in->ocnfrac, in->tau00, in->wndls, in->bflxls,
                        in->fluxu00, in->fluxv00, in->fluxt00, in->fluxq00};
  size_t isz[] = {ld * plev, ld * (plev + 1), ld * plev, ld * plev, ld * plev, ld * plev, ld, ld * plev,
                  ld * (plev + 1), ld * plev, ld, ld * plev, ld * plev, ld, ld, ld, ld, ld, ld, ld, ld};
  for (int q = 0; q < 21; q++) add(ip[q], ip[q] ? isz[q] : 0, true, false);
  double* sp[] = {st->u_wind, st->v_wind, st->w_wind, st->temperature, st->qt, st->qp, st->qn};
  for (double* p : sp) add(p, g4, true, true);
  add(rad->qrad, r4, true, false);
  double* rp[] = {rad->temperature, rad->qv, rad->qc, rad->qi, rad->cld};
  for (double* p : rp) add(p, r4, true, true);
  double* og[] = {out->qcl, out->qci, out->qpl, out->qpi, out->tk, out->tkh};
  for (double* p : og) add(p, g4, true, true);
  add(out->prec_crm, g3, true, true);
  double* oc[] = {out->cltot, out->clhgh, out->clmed, out->cllow, out->precc, out->precl, out->precsc,
                  out->precsl, out->jt_crm, out->mx_crm, out->prectend, out->precstend, out->taux,
                  out->tauy, out->z0m, out->timing_factor};
  for (double* p : oc) add(p, ld, true, true);
  double* opl[] = {out->cldtop, out->qc_mean, out->qi_mean, out->qs_mean, out->qg_mean, out->qr_mean,
                   out->sltend, out->qltend, out->qcltend, out->qiltend, out->cld, out->gicewp,
                   out->gliqwp, out->mctot, out->mcup, out->mcdn, out->mcuup, out->mcudn, out->mu_crm,
                   out->md_crm, out->du_crm, out->eu_crm, out->ed_crm, out->flux_qt, out->fluxsgs_qt,
                   out->tkez, out->tkesgsz, out->tkz, out->flux_u, out->flux_v, out->flux_qp,
                   out->precflux, out->qt_ls, out->qt_trans, out->qp_trans, out->qp_fall, out->qp_src,
                   out->qp_evp, out->t_ls};
  for (double* p : opl) add(p, ld * plev, true, true);
  size_t total = 0;
  for (auto& f : fields) {
    if (!f.host && f.n) {
      // a missing optional input member (fluxu00.. are unused by crm()) stays absent
      if (&f - &fields[0] >= 17 && &f - &fields[0] <= 20) {
        f.n = 0;
        continue;
      }
      return CRM_ERR_ARG;
    }
    total += (f.n + 31) & ~size_t(31);
  }
  if (total * sizeof(double) > ctx->stage_bytes) {
    if (ctx->stage) HIPCHECK(hipFree(ctx->stage));
    ctx->stage = nullptr;
    if (hipMalloc(&ctx->stage, total * sizeof(double)) != hipSuccess) {
      ctx->stage_bytes = 0;
      return CRM_ERR_NOMEM;
    }
    ctx->stage_bytes = total * sizeof(double);
  }
  double* p = ctx->stage;
  for (auto& f : fields) {
    if (!f.n) continue;
    f.dev = p;
    p += (f.n + 31) & ~size_t(31);
    HIPCHECK(hipMemcpyAsync(f.dev, f.host, f.n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  }
  size_t q = 0;
  const double** di[] = {&din.zmid, &din.zint, &din.tl, &din.ql, &din.qccl, &din.qiil, &din.ps, &din.pmid,
                         &din.pint, &din.pdel, &din.phis, &din.ul, &din.vl, &din.ocnfrac, &din.tau00,
                         &din.wndls, &din.bflxls, &din.fluxu00, &din.fluxv00, &din.fluxt00, &din.fluxq00};
  for (auto pp : di) *pp = fields[q++].dev;
  double** ds[] = {&dst.u_wind, &dst.v_wind, &dst.w_wind, &dst.temperature, &dst.qt, &dst.qp, &dst.qn};
  for (auto pp : ds) *pp = fields[q++].dev;
  drad.qrad = fields[q++].dev;
  double** dr[] = {&drad.temperature, &drad.qv, &drad.qc, &drad.qi, &drad.cld};
  for (auto pp : dr) *pp = fields[q++].dev;
  double** dg[] = {&dout.qcl, &dout.qci, &dout.qpl, &dout.qpi, &dout.tk, &dout.tkh, &dout.prec_crm,
                   &dout.cltot, &dout.clhgh, &dout.clmed, &dout.cllow, &dout.precc, &dout.precl,
                   &dout.precsc, &dout.precsl, &dout.jt_crm, &dout.mx_crm, &dout.prectend,
                   &dout.precstend, &dout.taux, &dout.tauy, &dout.z0m, &dout.timing_factor,
                   &dout.cldtop, &dout.qc_mean, &dout.qi_mean, &dout.qs_mean, &dout.qg_mean,
                   &dout.qr_mean, &dout.sltend, &dout.qltend, &dout.qcltend, &dout.qiltend, &dout.cld,
                   &dout.gicewp, &dout.gliqwp, &dout.mctot, &dout.mcup, &dout.mcdn, &dout.mcuup,
                   &dout.mcudn, &dout.mu_crm, &dout.md_crm, &dout.du_crm, &dout.eu_crm, &dout.ed_crm,
                   &dout.flux_qt, &dout.fluxsgs_qt, &dout.tkez, &dout.tkesgsz, &dout.tkz, &dout.flux_u,
                   &dout.flux_v, &dout.flux_qp, &dout.precflux, &dout.qt_ls, &dout.qt_trans,
                   &dout.qp_trans, &dout.qp_fall, &dout.qp_src, &dout.qp_evp, &dout.t_ls};
  for (auto pp : dg) *pp = fields[q++].dev;
  return CRM_OK;
}

template <typename R>
static int unstage(Engine<R>* ctx, std::vector<ViewField>& fields) {
  for (auto& f : fields)
    if (f.outw && f.n) HIPCHECK(hipMemcpyAsync((void*)f.host, f.dev, f.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHECK(hipStreamSynchronize(ctx->stream));
  return CRM_OK;
}

template <typename R>
static int run_host(Engine<R>* ctx, const crm_dims* dims, const crm_column_info* col, double dt_gl,
                    const crm_input_view* in, crm_state_view* state, crm_rad_view* rad, crm_output_view* out,
                    int phase_stop) {
  if (!ctx || !dims || !in || !state || !rad || !out) return CRM_ERR_ARG;
  int rc = validate(dims);
  if (rc) return rc;
  crm_input_view din{};
  crm_state_view dst{};
  crm_rad_view drad{};
  crm_output_view dout{};
  std::vector<ViewField> fields;
  rc = stage_views(ctx, dims, in, state, rad, out, din, dst, drad, dout, fields);
  if (rc) return rc;
  rc = run_core(ctx, dims, col, dt_gl, &din, &dst, &drad, &dout, phase_stop);
  if (rc) return rc;
  if (phase_stop == 0) rc = unstage(ctx, fields);
  return rc;
}


template <typename R>
Engine<R>::~Engine() {
  (void)hipSetDevice(device);
  for (auto& kv : transforms) {
    (void)hipFree(kv.second.Ft);
    (void)hipFree(kv.second.Gt);
    (void)hipFree(kv.second.eign);
  }
  drop_graphs();
  if (mlp_x) (void)hipFree(mlp_x);
  if (qrad_t) (void)hipFree(qrad_t);
  if (mlp_y) (void)hipFree(mlp_y);
  if (samp_n) (void)hipFree(samp_n);
  if (pool) (void)hipFree(pool);
  if (stage) (void)hipFree(stage);
  if (gmax) (void)hipFree(gmax);
  if (gcol_dev) (void)hipFree(gcol_dev);
  if (lat_dev) (void)hipFree(lat_dev);
  for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev_ready) (void)hipEventDestroy(ev_ready);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

template <typename R>
int Engine<R>::run(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                   crm_state_view* state, crm_rad_view* rad, crm_output_view* out, bool device_views,
                   int phase_stop) {
  if (!dims || !in || !state || !rad || !out) return CRM_ERR_ARG;
  (void)hipSetDevice(device);
  if (device_views) return run_core(this, dims, col, dt_gl, in, state, rad, out, phase_stop);
  return run_host(this, dims, col, dt_gl, in, state, rad, out, phase_stop);
}

template <typename R>
int Engine<R>::run_joint(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                         crm_state_view* state, crm_rad_view* rad, crm_output_view* out, JointCapture* jc,
                         int slot) {
  if (!dims || !in || !state || !rad || !out) {
    if (jc) (void)jc->arrive(slot, false, nullptr, nullptr, std::string(), nullptr);
    return CRM_ERR_ARG;
  }
  (void)hipSetDevice(device);
  if (!ev_ready && hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming) != hipSuccess) {
    if (jc) (void)jc->arrive(slot, false, nullptr, nullptr, std::string(), nullptr);
    return CRM_ERR_DEVICE;
  }
  return run_core(this, dims, col, dt_gl, in, state, rad, out, 0, nullptr, jc, slot);
}

template <typename R>
int Engine<R>::run_noalloc(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                           crm_state_view* state, crm_rad_view* rad, crm_output_view* out, char* scratch,
                           size_t scratch_bytes) {
  if (!dims || !in || !state || !rad || !out || !scratch) return CRM_ERR_ARG;
  (void)hipSetDevice(device);
  const CallScratch sc{scratch, scratch_bytes};
  return run_core(this, dims, col, dt_gl, in, state, rad, out, 0, &sc);
}

template <typename R>
int Engine<R>::profile_kernels(unsigned mask) {
  prof_mask = mask;
  for (int k = 0; k < CRM_K_COUNT; k++) {
    kt_ms[k] = 0.0;
    kt_n[k] = 0;
  }
  return CRM_OK;
}

template <typename R>
int Engine<R>::kernel_time(int kid, double* ms, long long* n) const {
  if (kid < 0 || kid >= CRM_K_COUNT || !ms || !n) return CRM_ERR_ARG;
  *ms = kt_ms[kid];
  *n = kt_n[kid];
  return CRM_OK;
}

template <typename R>
int Engine<R>::debug_get(const char* name, double* buf) {
  Engine<R>* ctx = this;
  if (!ctx->pool || !name || !buf) return -1;
  const DevT<R>& d = ctx->dev;
  const int ncrms = d.ncrms, nzm = d.nzm, nz = d.nz, nx = ctx->cap_nx;
  std::string n(name);
  const R* grid = nullptr;
  const double* grid64 = nullptr;
  int nk = nzm;
  if (n == "u") grid = d.u;
  else if (n == "w") { grid = d.w; nk = nz; }
  else if (n == "t") grid = d.t;
  else if (n == "tabs") grid = d.tabs;
  else if (n == "qv") grid = d.qv;
  else if (n == "qcl") grid = d.qcl;
  else if (n == "qci") grid = d.qci;
  else if (n == "qpl") grid = d.qpl;
  else if (n == "qpi") grid = d.qpi;
  else if (n == "qt") grid = d.qt;
  else if (n == "qp") grid = d.qp;
  else if (n == "qn") grid = d.qn;
  else if (n == "tk") grid = d.tk;
  else if (n == "tkh") grid = d.tkh;
  else if (n == "tke") grid = d.tke;
  else if (n == "dudt_na") grid = d.dudt[ctx->na];
  else if (n == "dudt_nb") grid = d.dudt[ctx->nb];
  else if (n == "dudt_nc") grid = d.dudt[ctx->nc];
  else if (n == "v" && d.v) grid = d.v;
  else if (n == "dvdt_na" && d.v) grid = d.dvdt[ctx->na];
  else if (n == "dvdt_nb" && d.v) grid = d.dvdt[ctx->nb];
  else if (n == "dvdt_nc" && d.v) grid = d.dvdt[ctx->nc];
  else if (n == "dwdt_na") { grid = d.dwdt[ctx->na]; nk = nz; }
  else if (n == "dwdt_nb") { grid = d.dwdt[ctx->nb]; nk = nz; }
  else if (n == "dwdt_nc") { grid = d.dwdt[ctx->nc]; nk = nz; }
  else if (n == "precsfc") { grid64 = d.precsfc; nk = 1; }
  else if (n == "precssfc") { grid64 = d.precssfc; nk = 1; }
  if (grid64) {
    size_t cnt = (size_t)ncrms * nk * nx;
    if (hipMemcpy(buf, grid64, cnt * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return (int)cnt;
  }
  if (grid) {
    size_t cnt = (size_t)ncrms * nk * nx;
    std::vector<R> tmp(cnt);
    if (hipMemcpy(tmp.data(), grid, cnt * sizeof(R), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    for (size_t q = 0; q < cnt; q++) buf[q] = (double)tmp[q];
    return (int)cnt;
  }
  static const std::map<std::string, std::pair<int, int>> profs = {
      {"t0", {P_T0, 0}}, {"q0", {P_Q0, 0}}, {"qv0", {P_QV0, 0}}, {"tabs0", {P_TABS0, 0}}, {"u0", {P_U0, 0}},
      {"qn0", {P_QN0, 0}}, {"qp0", {P_QP0, 0}}, {"rho", {P_RHO, 0}}, {"rhow", {P_RHOW, 1}},
      {"adz", {P_ADZ, 0}}, {"adzw", {P_ADZW, 1}}, {"pres", {P_PRES, 0}}, {"presi", {P_PRESI, 1}},
      {"z", {P_Z, 1}}, {"zi", {P_ZI, 1}}, {"bet", {P_BET, 0}}, {"gamaz", {P_GAMAZ, 0}},
      {"ttend", {P_TTEND, 0}}, {"qtend", {P_QTEND, 0}}, {"utend", {P_UTEND, 0}}, {"uwle", {P_UWLE, 1}},
      {"v0", {P_V0, 0}}, {"vtend", {P_VTEND, 0}}, {"vwle", {P_VWLE, 1}}, {"vwsb", {P_VWSB, 1}},
      {"uwsb", {P_UWSB, 1}}, {"mkwle1", {P_MKWLE1, 1}}, {"mkwsb1", {P_MKWSB1, 1}}, {"mkadv1", {P_MKADV1, 1}},
      {"mkdiff1", {P_MKDIFF1, 1}}, {"qpsrc", {P_QPSRC, 1}}, {"qpevp", {P_QPEVP, 1}},
      {"qpfall", {P_QPFALL, 1}}, {"precflux", {P_PRECFLUX, 1}}};
  auto it = profs.find(n);
  if (it != profs.end()) {
    int len = it->second.second ? nz : nzm;
    std::vector<double> tmp((size_t)ncrms * nz);
    if (hipMemcpy(tmp.data(), d.prof + (size_t)it->second.first * ncrms * nz, tmp.size() * sizeof(double),
                  hipMemcpyDeviceToHost) != hipSuccess)
      return -1;
    size_t m = 0;
    for (int c = 0; c < ncrms; c++)
      for (int k = 0; k < len; k++) buf[m++] = tmp[(size_t)c * nz + k];
    return (int)m;
  }
  if (n == "dz") {
    if (hipMemcpy(buf, d.inst + (size_t)I_DZ * ncrms, ncrms * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return -1;
    return ncrms;
  }
  return -1;
}


// Workgroup sizes per kernel class: one call per candidate size with every class timed on
// the stream, then the fastest size of each class.  The views are scratch: each call advances
// them.
template <typename R>
int Engine<R>::autotune(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                        crm_state_view* state, crm_rad_view* rad, crm_output_view* out, int* best) {
  // every workgroup size (64, 128, 256) with every launch variant; each trial call times
  // every kernel class
  static const int cand[] = {64, 128, 256};
  const unsigned saved_mask = prof_mask;
  const bool y3 = dims && dims->ny > 1;
  const int nxc = dims ? dims->nx * dims->ny : 0;
  constexpr int NV = MAX_VARIANTS;
  double t[3][NV][CRM_K_COUNT] = {};
  for (int v = 0; v < NV; v++) {
    for (int c = 0; c < 3; c++) {
      bool any = false;  // a class whose variant v exists
      for (int k = 0; k < CRM_K_COUNT; k++) {
        block[k] = cand[c];
        const int nv = variants_of(k, nxc, y3);
        variant[k] = v < nv ? v : 0;
        any |= v < nv;
      }
      if (!any) continue;
      drop_graphs();
      profile_kernels((1u << CRM_K_COUNT) - 1);
      int rc = run(dims, col, dt_gl, in, state, rad, out, true, 0);
      if (rc) {
        profile_kernels(saved_mask);
        for (int k = 0; k < CRM_K_COUNT; k++) variant[k] = 0;
        return rc;
      }
      for (int k = 0; k < CRM_K_COUNT; k++)
        if (v < variants_of(k, nxc, y3)) t[c][v][k] = kt_n[k] ? kt_ms[k] / (double)kt_n[k] : 0.0;
    }
  }
  for (int k = 0; k < CRM_K_COUNT; k++) {
    // 256 threads and the default variant unless another measured faster by more than 3%
    // (timing noise of a short trial call is a few percent)
    int bc = 2, bv = 0;
    double bt = t[2][0][k];
    for (int v = 0; v < variants_of(k, nxc, y3); v++)
      for (int c = 0; c < 3; c++)
        if (t[c][v][k] > 0.0 && t[c][v][k] < bt) {
          bc = c;
          bv = v;
          bt = t[c][v][k];
        }
    if (!(bt < 0.97 * t[2][0][k])) {
      bc = 2;
      bv = 0;
    }
    block[k] = cand[bc];
    variant[k] = bv;
    if (best) best[k] = block[k];
  }
  drop_graphs();
  profile_kernels(saved_mask);
  return CRM_OK;
}

template <typename R>
static CrmEngine* make_engine(int device, int* rc) {
  auto* e = new Engine<R>();
  e->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess ||
      hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
      hipMalloc(&e->gmax, GMAX_WORDS * sizeof(unsigned long long)) != hipSuccess) {
    delete e;
    *rc = CRM_ERR_DEVICE;
    return nullptr;
  }
  *rc = CRM_OK;
  return e;
}

}  // namespace CRM_NS
