// Shared device/host helpers of libsnapx (MI355X, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdint.h>

#include <string>

#include "snapx.h"

namespace snapx {

enum { IDN = 0, IVX = 1, IVY = 2, IVZ = 3, IPR = 4 };
enum { ILT = 0, IRT = 1 };

// Block extents derived from a snapx_config (src/coord/coordinate.hpp:58-60).
struct Grid {
  int n1, n2, n3;     // cells including ghosts (1 on inactive axes)
  int ng;             // nghost
  int nx1, nx2, nx3;  // interior cells
  int ncell;          // n1 * n2 * n3
  double dx1, dx2, dx3;
};

inline Grid make_grid(const snapx_config* c) {
  Grid g;
  g.ng = c->nghost;
  g.nx1 = c->nx1;
  g.nx2 = c->nx2;
  g.nx3 = c->nx3;
  g.n1 = c->nx1 > 1 ? c->nx1 + 2 * c->nghost : 1;
  g.n2 = c->nx2 > 1 ? c->nx2 + 2 * c->nghost : 1;
  g.n3 = c->nx3 > 1 ? c->nx3 + 2 * c->nghost : 1;
  g.ncell = g.n1 * g.n2 * g.n3;
  g.dx1 = (c->x1max - c->x1min) / c->nx1;
  g.dx2 = (c->x2max - c->x2min) / c->nx2;
  g.dx3 = (c->x3max - c->x3min) / c->nx3;
  return g;
}

// Thread-local error string behind snapx_last_error().
void set_error(const std::string& msg);

#define SNAPX_CHECK_ARG(cond, msg)        \
  do {                                    \
    if (!(cond)) {                        \
      ::snapx::set_error(msg);            \
      return SNAPX_EINVAL;                \
    }                                     \
  } while (0)

#define SNAPX_HIP(expr)                                                    \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      ::snapx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
      return SNAPX_EHIP;                                                   \
    }                                                                      \
  } while (0)

#define SNAPX_LAUNCH_CHECK()                                               \
  do {                                                                     \
    hipError_t e_ = hipGetLastError();                                     \
    if (e_ != hipSuccess) {                                                \
      ::snapx::set_error(std::string("kernel launch: ") +                  \
                         hipGetErrorString(e_));                           \
      return SNAPX_EHIP;                                                   \
    }                                                                      \
  } while (0)

template <typename T>
__device__ __forceinline__ T sq(T x) {
  return x * x;
}

// Grid-stride launch shape for n independent items.
inline dim3 blocks_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b > 65535 * 16) b = 65535 * 16;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// rocTX range over the launches enqueued in one scope (host side): a
// profiler run with marker tracing (rocprofv3 --marker-trace) attributes
// each kernel to the step, stage and kernel role that launched it.
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

// Validates a config for the kernels of this library.
int validate_config(const snapx_config* c);

// The fused path's explicit tendency (its K1) for another unit of the
// library: du <- dt times the flux divergence and the sources of the FP32
// state u, interior cells only, enqueued on st.  fused_rhs_ok tells, without
// setting an error, whether the fused path takes the block's options.
bool fused_rhs_ok(const snapx_config* c);
int fused_explicit_tendency(const snapx_config* c, double dt, const float* u, float* du,
                            hipStream_t st);

// Quotient and remainder of a small non-negative index by a positive
// divisor through the FP32 reciprocal and one correction step (an integer
// division by a runtime divisor costs some forty instructions).  Exact for
// 0 <= l < 2^22.
struct DivMod {
  int d;
  float inv;
  __device__ __forceinline__ explicit DivMod(int dd) : d(dd), inv(1.0f / (float)dd) {}
  __device__ __forceinline__ int div(int l, int& r) const {
    int q = (int)((float)l * inv);
    r = l - q * d;
    if (r < 0) r += d, q -= 1;
    if (r >= d) r -= d, q += 1;
    return q;
  }
};

}  // namespace snapx
