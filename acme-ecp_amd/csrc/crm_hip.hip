// crm_hip.hip -- the C ABI of include/crm_hip.h and include/crm_hip_debug.h, and the FP64
// engine (crm_driver.hpp instantiated with FP64 bulk fields).  The mixed-precision engine is
// compiled separately (crm_hip_mixed.hip); a context owns one engine of its precision.
#include <atomic>
#include <list>
#include <map>
#include <thread>
#include <vector>

#include "crm_driver.hpp"
#include "fft991_operator.hpp"

namespace crm {
CrmEngine* make_engine_fp64(int device, int* rc) { return make_engine<double>(device, rc); }

void pressure_transform_host(int nx, int ny, double dx, std::vector<double>& Ft, std::vector<double>& Gt,
                             std::vector<double>& eign) {
  // pressure.F90:172-198
  const double pii = 3.14159265358979323846, xnx = pii / nx, ddx2 = 1.0 / (dx * dx);
  if (ny <= 1) {
    fft991_operators(nx, Ft, Gt);
    eign.assign(nx, 0.0);
    for (int s = 0; s < nx; s++) {
      double xi = fft991_slot_mode(nx, s);
      eign[s] = (2.0 * std::cos(2.0 * xnx * xi) - 2.0) * ddx2 + 0.0;
    }
    return;
  }
  // 3D (dy == dx): the x transform of every row, then the y transform of every x
  // coefficient (pressure.F90:122-147), is one linear map of the nx*ny slab: the Kronecker
  // product of the two one-dimensional operators.  Column q = nx*jq + iq, mode s = nx*sy + sx.
  std::vector<double> fx, gx, fy, gy;
  fft991_operators(nx, fx, gx);
  fft991_operators(ny, fy, gy);
  const int n = nx * ny;
  const double xny = pii / ny, ddy2 = 1.0 / (dx * dx);
  Ft.assign((size_t)n * n, 0.0);
  Gt.assign((size_t)n * n, 0.0);
  eign.assign(n, 0.0);
  for (int q = 0; q < n; q++) {
    const int iq = q % nx, jq = q / nx;
    for (int s = 0; s < n; s++) {
      const int sx = s % nx, sy = s / nx;
      Ft[(size_t)q * n + s] = fx[(size_t)iq * nx + sx] * fy[(size_t)jq * ny + sy];
      Gt[(size_t)s * n + q] = gx[(size_t)sx * nx + iq] * gy[(size_t)sy * ny + jq];
    }
  }
  for (int s = 0; s < n; s++) {
    const double xi = fft991_slot_mode(nx, s % nx), xj = fft991_slot_mode(ny, s / nx);
    eign[s] = (2.0 * std::cos(2.0 * xnx * xi) - 2.0) * ddx2 + (2.0 * std::cos(2.0 * xny * xj) - 2.0) * ddy2;
  }
}
}  // namespace crm

struct crm_hip_ctx {
  int device = 0, precision = CRM_PREC_FP64;
  CrmEngine* eng = nullptr;
};

// ---- device memory of crm_hip_run_device_noalloc ------------------------------------------
// One fixed array per GPU holds every device buffer of a call of the entry point that
// allocates nothing.  In host code the array's name is only a stand-in: its device address on
// each GPU is looked up once per device by crm_hip_init (hipGetSymbolAddress with that device
// current) and kept in noalloc_base[device]; every later use goes through that address.
static __device__ __align__(256) char scratch[CRM_HIP_NOALLOC_SCRATCH_BYTES];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::atomic<char*> noalloc_base[NOALLOC_MAX_DEVICES];
static std::atomic<int> noalloc_state[NOALLOC_MAX_DEVICES];  // 0 not looked up, 1 ok, -1 failed

// Called with `device` current.
static void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(crm_device_mutex());
  if (noalloc_state[device].load() != 0) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    fprintf(stderr, "crm_hip: device %d: address of the no-allocation scratch array: %s\n", device,
            hipGetErrorString(e));
    noalloc_state[device].store(-1);
    return;
  }
  noalloc_base[device].store((char*)ptr);
  noalloc_state[device].store(1);
}

std::mutex& crm_device_mutex() {
  static std::mutex m;
  return m;
}

int crm_ctx_device(const crm_hip_ctx* ctx) { return ctx->device; }
hipStream_t crm_ctx_stream(const crm_hip_ctx* ctx) { return ctx->eng->get_stream(); }

extern "C" {

int crm_hip_init(crm_hip_ctx** out, const crm_hip_config* cfg) {
  if (!out) return CRM_ERR_ARG;
  *out = nullptr;
  auto* ctx = new crm_hip_ctx();
  if (cfg) {
    ctx->device = cfg->device;
    ctx->precision = cfg->precision;
  }
  int rc = CRM_OK;
  if (ctx->precision == CRM_PREC_FP64)
    ctx->eng = crm::make_engine_fp64(ctx->device, &rc);
  else if (ctx->precision == CRM_PREC_MIXED)
    ctx->eng = crm_mixed::make_engine_mixed(ctx->device, &rc);
  else
    rc = CRM_ERR_ARG;
  if (rc != CRM_OK) {
    delete ctx;
    return rc;
  }
  noalloc_setup(ctx->device);  // the engine has made ctx->device current
  *out = ctx;
  return CRM_OK;
}

int crm_hip_run(crm_hip_ctx* ctx, const crm_dims* dims, const crm_column_info* col, double dt_gl,
                const crm_input_view* in, crm_state_view* state, crm_rad_view* rad, crm_output_view* out) {
  if (!ctx) return CRM_ERR_ARG;
  return ctx->eng->run(dims, col, dt_gl, in, state, rad, out, false, 0);
}

int crm_hip_run_device(crm_hip_ctx* ctx, const crm_dims* dims, const crm_column_info* col, double dt_gl,
                       const crm_input_view* in, crm_state_view* state, crm_rad_view* rad, crm_output_view* out) {
  if (!ctx) return CRM_ERR_ARG;
  return ctx->eng->run(dims, col, dt_gl, in, state, rad, out, true, 0);
}

int crm_hip_run_device_noalloc(crm_hip_ctx* ctx, const crm_dims* dims, const crm_column_info* col, double dt_gl,
                               const crm_input_view* in, crm_state_view* state, crm_rad_view* rad,
                               crm_output_view* out) {
  if (!ctx || !ctx->eng || ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return CRM_ERR_ARG;
  if (noalloc_state[ctx->device].load() != 1) return CRM_ERR_DEVICE;
  return ctx->eng->run_noalloc(dims, col, dt_gl, in, state, rad, out, noalloc_base[ctx->device].load(),
                               CRM_HIP_NOALLOC_SCRATCH_BYTES);
}

int crm_hip_accel_init(crm_hip_ctx* ctx, int use_crm_accel, double crm_accel_factor, int crm_accel_uv) {
  if (!ctx || !ctx->eng) return CRM_ERR_ARG;
  return ctx->eng->accel_init(use_crm_accel, crm_accel_factor, crm_accel_uv);
}

static_assert(CRM_HIP_NKERNEL == CRM_K_COUNT, "kernel class count");

// ---- one graph for several classes (crm_hip_run_many) ---------------------------------------
namespace {
// Captured joint graphs, keyed by the classes' own whole-call keys in class order; least
// recently used first out.  Replayed only under joint_mutex, so no evicted graph is running.
std::mutex joint_mutex;
std::map<std::string, hipGraphExec_t> joint_graphs;
std::list<std::string> joint_lru;
constexpr size_t JOINT_GRAPHS_MAX = 8;
}  // namespace

int JointCapture::arrive(int slot, bool joins, hipStream_t stream, hipEvent_t ready, std::string key,
                         std::function<void(bool)> sequence) {
  std::unique_lock<std::mutex> lk(m);
  Part& p = parts[slot];
  p.joins = joins;
  p.stream = stream;
  p.ready = ready;
  p.key = std::move(key);
  p.sequence = std::move(sequence);
  arrived++;
  if (arrived == (int)parts.size()) {
    rc = launch_locked();
    released = true;
    cv.notify_all();
    return rc;
  }
  if (!joins) return CRM_OK;  // runs on its own; nothing of it is in the joint graph
  cv.wait(lk, [&] { return released; });
  return rc;
}

int JointCapture::launch_locked() {
  std::string key;
  int njoin = 0;
  for (size_t i = 0; i < parts.size(); i++) {
    const Part& p = parts[i];
    if (!p.joins) continue;
    njoin++;
    const size_t len = p.key.size();
    key.append((const char*)&i, sizeof(i));
    key.append((const char*)&len, sizeof(len));
    key.append(p.key);
  }
  if (njoin == 0) return CRM_OK;
  if (hipSetDevice(device) != hipSuccess) return CRM_ERR_DEVICE;
  std::lock_guard<std::mutex> jl(joint_mutex);
  hipGraphExec_t exec = nullptr;
  auto it = joint_graphs.find(key);
  if (it != joint_graphs.end()) {
    exec = it->second;
    joint_lru.remove(key);
    joint_lru.push_front(key);
    for (Part& p : parts)
      if (p.joins) p.sequence(false);
  } else {
    // capture every class's sequence on its own stream, forked from and joined into master
    std::lock_guard<std::mutex> capture_lock(crm_device_mutex());
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(master, hipStreamCaptureModeThreadLocal) != hipSuccess) return CRM_ERR_DEVICE;
    bool ok = hipEventRecord(fork, master) == hipSuccess;
    for (size_t i = 0; i < parts.size() && ok; i++) {
      Part& p = parts[i];
      if (!p.joins) continue;
      ok = hipStreamWaitEvent(p.stream, fork, 0) == hipSuccess;
      if (!ok) break;
      p.sequence(true);
      ok = hipEventRecord(join[i], p.stream) == hipSuccess && hipStreamWaitEvent(master, join[i], 0) == hipSuccess;
    }
    const hipError_t e = hipStreamEndCapture(master, &graph);
    if (!ok || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return CRM_ERR_DEVICE;
    }
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return CRM_ERR_DEVICE;
    while (joint_graphs.size() >= JOINT_GRAPHS_MAX && !joint_lru.empty()) {
      auto old = joint_graphs.find(joint_lru.back());
      if (old != joint_graphs.end()) {
        (void)hipGraphExecDestroy(old->second);
        joint_graphs.erase(old);
      }
      joint_lru.pop_back();
    }
    joint_graphs[key] = exec;
    joint_lru.push_front(key);
  }
  // the graph runs after every class's set-up, and every class continues after the graph
  for (Part& p : parts)
    if (p.joins && hipStreamWaitEvent(master, p.ready, 0) != hipSuccess) return CRM_ERR_DEVICE;
  if (hipGraphLaunch(exec, master) != hipSuccess || hipEventRecord(done, master) != hipSuccess) return CRM_ERR_DEVICE;
  for (Part& p : parts)
    if (p.joins && hipStreamWaitEvent(p.stream, done, 0) != hipSuccess) return CRM_ERR_DEVICE;
  return CRM_OK;
}

int crm_hip_run_many(int n, crm_hip_ctx* const* ctxs, const crm_class_desc* classes, double dt_gl, int* status) {
  if (n < 1 || !ctxs || !classes) return CRM_ERR_ARG;
  for (int i = 0; i < n; i++) {
    if (!ctxs[i] || !ctxs[i]->eng) return CRM_ERR_ARG;
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return CRM_ERR_ARG;
  }
  std::vector<int> rcs(n, CRM_OK);
  // The classes' whole-call graphs are captured together into one graph when every context is
  // on one device with graph mode CRM_GRAPHS_JOINT; each class still runs its set-up, any
  // resumed substeps and its outputs on its own stream and host thread.  Otherwise each class
  // replays its own graph on its own stream (measured faster: the device overlaps the classes'
  // streams more than the branches of one graph).
  bool one_device = true;
  for (int i = 1; i < n; i++) one_device &= ctxs[i]->device == ctxs[0]->device;
  for (int i = 0; i < n; i++) one_device &= ctxs[i]->eng->get_graphs() == CRM_GRAPHS_JOINT;
  JointCapture jc;
  bool joint = one_device && hipSetDevice(ctxs[0]->device) == hipSuccess;
  if (joint) {
    jc.device = ctxs[0]->device;
    jc.parts.resize(n);
    jc.join.assign(n, nullptr);
    joint = hipStreamCreate(&jc.master) == hipSuccess &&
            hipEventCreateWithFlags(&jc.fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&jc.done, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < n && joint; i++) joint = hipEventCreateWithFlags(&jc.join[i], hipEventDisableTiming) == hipSuccess;
  }
  auto one = [&](int i) {
    if (hipSetDevice(ctxs[i]->device) != hipSuccess) {
      if (joint) (void)jc.arrive(i, false, nullptr, nullptr, std::string(), nullptr);
      rcs[i] = CRM_ERR_DEVICE;
      return;
    }
    crm_class_desc c = classes[i];
    if (joint)
      rcs[i] = ctxs[i]->eng->run_joint(&c.dims, &c.col, dt_gl, &c.in, &c.state, &c.rad, &c.out, &jc, i);
    else
      rcs[i] = ctxs[i]->eng->run(&c.dims, &c.col, dt_gl, &c.in, &c.state, &c.rad, &c.out, true, 0);
  };
  // one host thread per class: each drives its own stream, so the classes' set-up and any
  // host-driven substeps overlap on the device
  std::vector<std::thread> th;
  for (int i = 1; i < n; i++) th.emplace_back(one, i);
  one(0);
  for (auto& t : th) t.join();
  if (jc.master) {
    (void)hipStreamSynchronize(jc.master);
    (void)hipStreamDestroy(jc.master);
  }
  for (hipEvent_t e : jc.join)
    if (e) (void)hipEventDestroy(e);
  if (jc.fork) (void)hipEventDestroy(jc.fork);
  if (jc.done) (void)hipEventDestroy(jc.done);
  int first = CRM_OK;
  for (int i = 0; i < n; i++) {
    if (status) status[i] = rcs[i];
    if (!first && rcs[i]) first = rcs[i];
  }
  return first;
}

int crm_hip_set_schedule(crm_hip_ctx* ctx, const int block[CRM_HIP_NKERNEL]) {
  if (!ctx || !ctx->eng || !block) return CRM_ERR_ARG;
  return ctx->eng->set_schedule(block);
}

int crm_hip_get_schedule(const crm_hip_ctx* ctx, int block[CRM_HIP_NKERNEL]) {
  if (!ctx || !ctx->eng || !block) return CRM_ERR_ARG;
  ctx->eng->get_schedule(block);
  return CRM_OK;
}

int crm_hip_set_coriolis(crm_hip_ctx* ctx, int docoriolis) {
  if (!ctx || !ctx->eng) return CRM_ERR_ARG;
  ctx->eng->set_coriolis(docoriolis);
  return CRM_OK;
}

int crm_hip_set_variants(crm_hip_ctx* ctx, const int variant[CRM_HIP_NKERNEL]) {
  if (!ctx || !ctx->eng || !variant) return CRM_ERR_ARG;
  return ctx->eng->set_variants(variant);
}

int crm_hip_get_variants(const crm_hip_ctx* ctx, int variant[CRM_HIP_NKERNEL]) {
  if (!ctx || !ctx->eng || !variant) return CRM_ERR_ARG;
  ctx->eng->get_variants(variant);
  return CRM_OK;
}

int crm_hip_autotune(crm_hip_ctx* ctx, const crm_dims* dims, const crm_column_info* col, double dt_gl,
                     const crm_input_view* in, crm_state_view* state, crm_rad_view* rad, crm_output_view* out,
                     int block_out[CRM_HIP_NKERNEL]) {
  if (!ctx || !ctx->eng) return CRM_ERR_ARG;
  if (hipSetDevice(ctx->device) != hipSuccess) return CRM_ERR_DEVICE;
  return ctx->eng->autotune(dims, col, dt_gl, in, state, rad, out, block_out);
}

int crm_hip_set_graphs(crm_hip_ctx* ctx, int on) {
  if (!ctx || !ctx->eng) return CRM_ERR_ARG;
  if (on < CRM_GRAPHS_OFF || on > CRM_GRAPHS_JOINT) return CRM_ERR_ARG;
  ctx->eng->set_graphs(on);
  return CRM_OK;
}

int crm_hip_set_micro_mlp(crm_hip_ctx* ctx, crm_mlp* mlp) {
  if (!ctx || !ctx->eng) return CRM_ERR_ARG;
  return ctx->eng->set_mlp(mlp);
}

int crm_hip_record_micro_samples(crm_hip_ctx* ctx, float* X, float* T, long long cap, int stride) {
  if (!ctx || !ctx->eng) return CRM_ERR_ARG;
  if (hipSetDevice(ctx->device) != hipSuccess) return CRM_ERR_DEVICE;
  return ctx->eng->record_samples(X, T, cap, stride);
}

long long crm_hip_micro_samples_count(const crm_hip_ctx* ctx) {
  if (!ctx || !ctx->eng) return -1;
  return ctx->eng->samples_count();
}

long long crm_hip_graph_launches(const crm_hip_ctx* ctx) {
  if (!ctx || !ctx->eng) return -1;
  return ctx->eng->graph_count();
}

int crm_hip_graph_cache_size(const crm_hip_ctx* ctx) {
  if (!ctx || !ctx->eng) return -1;
  return ctx->eng->graph_cache_size();
}

int crm_hip_debug_stop_call_at(crm_hip_ctx* ctx, int nstep) {
  if (!ctx || !ctx->eng || nstep < 0) return CRM_ERR_ARG;
  ctx->eng->set_debug_stop(nstep);
  return CRM_OK;
}

int crm_hip_get_stats(const crm_hip_ctx* ctx, crm_hip_stats* stats) {
  if (!ctx || !stats) return CRM_ERR_ARG;
  ctx->eng->get_stats(stats);
  return CRM_OK;
}

int crm_hip_finalize(crm_hip_ctx* ctx) {
  if (!ctx) return CRM_OK;
  delete ctx->eng;
  delete ctx;
  return CRM_OK;
}

const char* crm_hip_errstr(int code) {
  switch (code) {
    case CRM_OK: return "ok";
    case CRM_ERR_ARG: return "invalid argument (dims or null view pointer)";
    case CRM_ERR_CFL: return "Courant number too large: more than 16 subcycles (kurant) or a runaway precipitation sub-step count";
    case CRM_ERR_NXRAD: return "crm_nx_rad and crm_ny_rad need to divide nx and ny";
    case CRM_ERR_DAMPING: return "damping: tau_min is too small for dt";
    case CRM_ERR_NAN: return "non-finite prognostic field";
    case CRM_ERR_DEVICE: return "HIP runtime error";
    case CRM_ERR_NOMEM: return "device allocation failed";
    case CRM_ERR_UNSUPPORTED: return "size class or option not provided by this build";
    case CRM_ERR_ACCEL: return "bad crm_accel_factor and nstop pair: (1+crm_accel_factor) does not divide nstop";
    case CRM_ERR_SCRATCH: return "batch too large for the fixed device memory of crm_hip_run_device_noalloc";
    default: return "unknown error";
  }
}

int crm_hip_debug_run(crm_hip_ctx* ctx, const crm_dims* dims, const crm_column_info* col, double dt_gl,
                      const crm_input_view* in, crm_state_view* state, crm_rad_view* rad, crm_output_view* out,
                      int phase_stop) {
  if (!ctx) return CRM_ERR_ARG;
  return ctx->eng->run(dims, col, dt_gl, in, state, rad, out, false, phase_stop);
}

int crm_hip_debug_get(crm_hip_ctx* ctx, const char* name, double* buf) {
  if (!ctx) return -1;
  return ctx->eng->debug_get(name, buf);
}

static const char* const kernel_names[CRM_K_COUNT] = {
    "setup", "init", "setperturb", "kurant", "precip_init", "pre", "sgs", "mom",
    "press", "mpdata", "precip", "diffuse", "micro", "stats", "final"};

int crm_hip_profile_kernels(crm_hip_ctx* ctx, unsigned mask) {
  if (!ctx) return CRM_ERR_ARG;
  return ctx->eng->profile_kernels(mask);
}

int crm_hip_kernel_time(const crm_hip_ctx* ctx, int kid, double* ms_total, long long* launches) {
  if (!ctx) return CRM_ERR_ARG;
  return ctx->eng->kernel_time(kid, ms_total, launches);
}

const char* crm_hip_kernel_name(int kid) {
  return (kid >= 0 && kid < CRM_K_COUNT) ? kernel_names[kid] : "unknown";
}

// Host-side transform operators for one nx (tests compare them with the reference transform).
int crm_hip_fft991_operators(int n, double* Ft, double* Gt) {
  if (n < 4 || (n & 1) || !Ft || !Gt) return CRM_ERR_ARG;
  std::vector<double> F, G;
  crm::fft991_operators(n, F, G);
  std::memcpy(Ft, F.data(), sizeof(double) * n * n);
  std::memcpy(Gt, G.data(), sizeof(double) * n * n);
  return CRM_OK;
}

}  // extern "C"
