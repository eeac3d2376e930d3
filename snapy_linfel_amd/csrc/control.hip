// CFL time loop with redo, run on the device.
//
// The loop of examples/shock.cpp:59-79 around MeshBlockImpl::max_time_step
// (src/mesh/meshblock.cpp:511-530) and MeshBlockImpl::check_redo
// (src/mesh/meshblock.cpp:974-1028) keeps its state (time, cycle, redo
// count, status) in a device record, snapx_step_control.  Each attempt is a
// fixed sequence of launches: single-thread kernels for the decisions, the
// library's own CFL reduction, step and positivity check for the work, and a
// conditional copy for the restore.  Nothing waits on the host and nothing
// clears memory with a memset, so a caller can capture any number of
// attempts in one HIP graph.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.hpp"

namespace snapx {

namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// loop test and cycle count at the top of the loop body (shock.cpp:59-60)
__global__ void attempt_begin_kernel(snapx_step_control* c) {
  const bool stop =
      c->status != 0 || (c->nlim >= 0 && c->cycle >= c->nlim) || c->time >= c->tlim;
  if (stop && c->status == 0) c->status = 1;
  c->active = stop ? 0 : 1;
  if (!stop) c->cycle += 1;
}

// dt = 2^-redo * cfl * HydroImpl::max_time_step (meshblock.cpp:529)
__global__ void attempt_dt_kernel(snapx_step_control* c, double* log, int k) {
  const double dt = ldexp(1.0, -c->redo) * c->cfl * c->dt_raw;
  c->dt = dt;
  if (log) log[k] = c->active ? dt : 0.0;
}

// check_redo's decision (meshblock.cpp:1006-1027) and the loop's time update
// (shock.cpp:70-75).  restore: bit 0 the step's input u, bit 1 the attempt's
// input w (attempts after the loop stopped).
__global__ void attempt_end_kernel(snapx_step_control* c) {
  int restore = 0;
  if (!c->active) {
    restore = 3;
  } else if (c->verdict & c->redo_mask) {
    c->redo += 1;
    c->redos += 1;
    if (c->redo > c->max_redo) {
      c->status = -1;  // the reference returns -1 before restoring u
    } else {
      restore = 1;
      c->cycle -= 1;
    }
  } else {
    c->redo = 0;
    c->time += c->dt;
    c->steps += 1;
  }
  c->restore = restore;
}

template <typename T>
__global__ void attempt_restore_kernel(T* __restrict__ u, const T* __restrict__ ub,
                                       T* __restrict__ w, const T* __restrict__ wb, int64_t n,
                                       const snapx_step_control* __restrict__ c) {
  const int r = c->restore;
  if (!r) return;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    if (r & 1) u[q] = ub[q];
    if ((r & 2) && w) w[q] = wb[q];
  }
}

constexpr int kThreads = 256;

int launch_small(void (*k)(snapx_step_control*), snapx_step_control* c, hipStream_t st) {
  hipLaunchKernelGGL(k, dim3(1), dim3(1), 0, st, c);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_restore(void* u, const void* ub, void* w, const void* wb, int64_t n,
                   const snapx_step_control* c, hipStream_t st) {
  hipLaunchKernelGGL(attempt_restore_kernel<T>, blocks_for(n, kThreads), dim3(kThreads), 0, st,
                     reinterpret_cast<T*>(u), reinterpret_cast<const T*>(ub),
                     reinterpret_cast<T*>(w), reinterpret_cast<const T*>(wb), n, c);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

// The attempt loop around a step callback step(dt_device).
template <typename Step>
int run_attempts(const snapx_config* cfg, snapx_step_control* ctrl, int nattempts,
                 bool from_u, void* u, void* w, void* ub, void* wb, double* dt_log, int dtype,
                 hipStream_t st, Step step) {
  const Grid g = make_grid(cfg);
  const int64_t n = 5 * (int64_t)g.ncell;
  const size_t bytes = (dtype == SNAPX_F64 ? sizeof(double) : sizeof(float)) * (size_t)n;
  void* sv = reinterpret_cast<void*>(st);
  for (int k = 0; k < nattempts; ++k) {
    TraceRange range("snapx CFL attempt");
    if (int e = launch_small(attempt_begin_kernel, ctrl, st)) return e;
    if (wb) SNAPX_HIP(hipMemcpyAsync(wb, w, bytes, hipMemcpyDeviceToDevice, st));
    if (from_u)
      if (int e = snapx_ideal_gas_cons2prim(cfg, u, w, dtype, sv)) return e;
    SNAPX_HIP(hipMemcpyAsync(ub, u, bytes, hipMemcpyDeviceToDevice, st));
    if (int e = snapx_max_time_step(cfg, w, 1.0, &ctrl->dt_raw, dtype, sv)) return e;
    hipLaunchKernelGGL(attempt_dt_kernel, dim3(1), dim3(1), 0, st, ctrl, dt_log, k);
    SNAPX_LAUNCH_CHECK();
    if (int e = step(&ctrl->dt)) return e;
    if (int e = snapx_check_state(cfg, u, &ctrl->verdict, dtype, sv)) return e;
    if (int e = launch_small(attempt_end_kernel, ctrl, st)) return e;
    if (int e = dtype == SNAPX_F64 ? launch_restore<double>(u, ub, wb ? w : nullptr, wb, n, ctrl, st)
                                   : launch_restore<float>(u, ub, wb ? w : nullptr, wb, n, ctrl, st))
      return e;
  }
  return 0;
}

}  // namespace

}  // namespace snapx

using namespace snapx;

extern "C" size_t snapx_rk3_attempts_workspace_bytes(const snapx_config* cfg, int dtype) {
  if (validate_config(cfg)) return 0;
  const Grid g = make_grid(cfg);
  const size_t es = dtype == SNAPX_F64 ? sizeof(double) : sizeof(float);
  return 2 * align256(es * 5 * (size_t)g.ncell) + snapx_rk3_workspace_bytes(cfg, dtype);
}

extern "C" int snapx_rk3_attempts(const snapx_config* cfg, snapx_step_control* ctrl,
                                  int nattempts, int dt_source, void* u, void* w, double* dt_log,
                                  void* workspace, size_t workspace_bytes, int dtype,
                                  void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(ctrl && u && w && workspace && nattempts >= 0,
                  "snapx_rk3_attempts: bad arguments");
  SNAPX_CHECK_ARG(dt_source == SNAPX_DT_FROM_W || dt_source == SNAPX_DT_FROM_U,
                  "snapx_rk3_attempts: unknown dt source");
  SNAPX_CHECK_ARG(dtype == SNAPX_F64 || dtype == SNAPX_F32, "snapx_rk3_attempts: unknown dtype");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_rk3_attempts_workspace_bytes(cfg, dtype),
                  "snapx_rk3_attempts: workspace too small");
  const Grid g = make_grid(cfg);
  const size_t es = dtype == SNAPX_F64 ? sizeof(double) : sizeof(float);
  const size_t field = align256(es * 5 * (size_t)g.ncell);
  char* base = reinterpret_cast<char*>(workspace);
  void* ub = base;
  void* wb = base + field;
  void* sws = base + 2 * field;
  const size_t sbytes = workspace_bytes - 2 * field;
  return run_attempts(cfg, ctrl, nattempts, dt_source == SNAPX_DT_FROM_U, u, w, ub, wb, dt_log,
                      dtype, as_stream(stream), [&](const double* dt) {
                        return snapx_rk3_step_dt(cfg, dt, u, w, sws, sbytes, dtype, stream);
                      });
}

extern "C" size_t snapx_fused_attempts_workspace_bytes(const snapx_config* cfg) {
  const size_t fws = snapx_fused_workspace_bytes(cfg);
  if (!fws) return 0;
  const Grid g = make_grid(cfg);
  return 2 * align256(sizeof(float) * 5 * (size_t)g.ncell) + fws;
}

extern "C" int snapx_fused_attempts(const snapx_config* cfg, snapx_step_control* ctrl,
                                    int nattempts, float* u, double* dt_log, void* workspace,
                                    size_t workspace_bytes, void* stream) {
  const size_t need = snapx_fused_attempts_workspace_bytes(cfg);
  SNAPX_CHECK_ARG(need > 0, "snapx_fused_attempts: configuration not supported by the fused path");
  SNAPX_CHECK_ARG(ctrl && u && workspace && nattempts >= 0, "snapx_fused_attempts: bad arguments");
  SNAPX_CHECK_ARG(workspace_bytes >= need, "snapx_fused_attempts: workspace too small");
  const Grid g = make_grid(cfg);
  const size_t field = align256(sizeof(float) * 5 * (size_t)g.ncell);
  char* base = reinterpret_cast<char*>(workspace);
  void* w = base;  // primitives of the attempt's input, for the CFL step
  void* ub = base + field;
  void* fws = base + 2 * field;
  const size_t fbytes = workspace_bytes - 2 * field;
  return run_attempts(cfg, ctrl, nattempts, true, u, w, ub, nullptr, dt_log, SNAPX_F32,
                      as_stream(stream), [&](const double* dt) {
                        return snapx_fused_rk3_step_dt(cfg, dt, u, fws, fbytes, stream);
                      });
}
