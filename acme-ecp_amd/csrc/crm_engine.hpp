// crm_engine.hpp -- the interface between the C ABI (crm_hip.hip) and the device engines,
// one per bulk-field precision, each compiled in its own translation unit.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/crm_hip.h"
#include "../../include/crm_mlp.h"

struct CrmEngine {
  virtual ~CrmEngine() {}
  // One crm() call; device_views selects crm_hip_run_device semantics, phase_stop > 0 stops
  // inside the last substep (inspection only).
  virtual int run(const crm_dims* dims, const crm_column_info* col, double dt_gl,
                  const crm_input_view* in, crm_state_view* state, crm_rad_view* rad,
                  crm_output_view* out, bool device_views, int phase_stop) = 0;
  virtual int debug_get(const char* name, double* buf) = 0;
  virtual void get_stats(crm_hip_stats* s) const = 0;
  virtual int profile_kernels(unsigned mask) = 0;
  virtual int kernel_time(int kid, double* ms, long long* n) const = 0;
  virtual int accel_init(int use, double factor, int uv) = 0;
  virtual int set_schedule(const int* block) = 0;
  virtual void get_schedule(int* block) const = 0;
  virtual void set_graphs(int mode) = 0;
  virtual void set_debug_stop(int nstep) = 0;
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
#include <mutex>
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
