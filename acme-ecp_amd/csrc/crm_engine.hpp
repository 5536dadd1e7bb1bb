// crm_engine.hpp -- the interface between the C ABI (crm_hip.hip) and the device engines,
// one per bulk-field precision, each compiled in its own translation unit.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/crm_hip.h"
#include "../../include/crm_mlp.h"

#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

// Several contexts' calls of crm_hip_run_many captured into one HIP graph.  Each class runs its
// call on its own host thread up to the replay of its whole-call graph; there it hands over
// its sequence of kernels (enqueued with launch = true, or only its host bookkeeping with
// launch = false when the joint graph is cached) and waits.  The last class to arrive captures
// every class's sequence on the class's own stream, forked from and joined back into one
// stream, launches the one graph there and releases the classes, which each continue on their
// own stream after it (the resume of a replay that stopped, the outputs).  A class that does
// not replay a whole-call graph (acceleration, a network, timing, more substeps than a graph
// holds) arrives without a sequence and runs on its own; so does one that returns early.
struct JointCapture {
  struct Part {
    bool joins = false;
    hipStream_t stream = nullptr;
    hipEvent_t ready = nullptr;  // recorded on the class stream after the call's set-up
    std::string key;
    std::function<void(bool)> sequence;
  };
  int device = 0;
  std::vector<Part> parts;
  int arrived = 0;
  bool released = false;
  int rc = 0;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr, done = nullptr;
  std::vector<hipEvent_t> join;
  std::mutex m;
  std::condition_variable cv;
  // Arrive as class `slot`; with joins, wait until the joint graph has been launched (its work
  // then follows on `stream` through the done event) and return its status.
  int arrive(int slot, bool joins, hipStream_t stream, hipEvent_t ready, std::string key,
             std::function<void(bool)> sequence);
  // run by the last class to arrive, with m held
  int launch_locked();
};

struct CrmEngine {
  virtual ~CrmEngine() {}
  // One crm() call; device_views selects crm_hip_run_device semantics, phase_stop > 0 stops
  // inside the last substep (inspection only).
  virtual int run(const crm_dims* dims, const crm_column_info* col, double dt_gl,
                  const crm_input_view* in, crm_state_view* state, crm_rad_view* rad,
                  crm_output_view* out, bool device_views, int phase_stop) = 0;
  // crm_hip_run_device semantics as class `slot` of a joint capture (crm_hip_run_many)
  virtual int run_joint(const crm_dims* dims, const crm_column_info* col, double dt_gl,
                        const crm_input_view* in, crm_state_view* state, crm_rad_view* rad,
                        crm_output_view* out, JointCapture* jc, int slot) = 0;
  virtual int debug_get(const char* name, double* buf) = 0;
  virtual void get_stats(crm_hip_stats* s) const = 0;
  virtual int profile_kernels(unsigned mask) = 0;
  virtual int kernel_time(int kid, double* ms, long long* n) const = 0;
  virtual int accel_init(int use, double factor, int uv) = 0;
  virtual int set_schedule(const int* block) = 0;
  virtual void get_schedule(int* block) const = 0;
  virtual int set_variants(const int* variant) = 0;
  virtual void get_variants(int* variant) const = 0;
  virtual void set_graphs(int mode) = 0;
  virtual int get_graphs() const = 0;
  virtual void set_debug_stop(int nstep) = 0;
  virtual void set_coriolis(int on) = 0;
  virtual long long graph_count() const = 0;
  virtual int graph_cache_size() const = 0;
  virtual int set_mlp(const crm_mlp* m) = 0;
  virtual int record_samples(float* X, float* T, long long cap, int stride) = 0;
  virtual long long samples_count() const = 0;
  virtual int autotune(const crm_dims* dims, const crm_column_info* col, double dt_gl, const crm_input_view* in,
                       crm_state_view* state, crm_rad_view* rad, crm_output_view* out, int* best) = 0;
  virtual hipStream_t get_stream() const = 0;
  // One crm() call on device views whose every device buffer is carved from the caller's
  // region [scratch, scratch + scratch_bytes): plain launches, no graphs, no events.
  virtual int run_noalloc(const crm_dims* dims, const crm_column_info* col, double dt_gl,
                          const crm_input_view* in, crm_state_view* state, crm_rad_view* rad,
                          crm_output_view* out, char* scratch, size_t scratch_bytes) = 0;
};

// Serialises stream captures with the library's null-stream work (buffer set-up, operator
// uploads) across threads: the engine streams are blocking streams, so null-stream work of one
// context would join, and invalidate, another context's capture in progress
// (crm_hip_run_many runs classes on concurrent threads).
std::mutex& crm_device_mutex();

// The device and stream of a context (for the GCM-side kernels of crm_glue.hip).
int crm_ctx_device(const crm_hip_ctx* ctx);
hipStream_t crm_ctx_stream(const crm_hip_ctx* ctx);

namespace crm {
CrmEngine* make_engine_fp64(int device, int* rc);
}
namespace crm_mixed {
CrmEngine* make_engine_mixed(int device, int* rc);
}
