// Vertically implicit correction (VIC) on MI355X, FP64 arithmetic.
//
// Three kernels per call, like the reference GPU split
// (src/implicit/implicit_dispatch.cu:28-181):
//   assemble      one thread per (column, layer): block rows a, b, c
//   sweep         one thread per column: block Thomas forward sweep,
//                 back substitution, column reductions
//   redistribute  one thread per (column, layer): MS-VIC mass redistribution
// Columns are the interior (x2, x3) positions; x1 is the vertical.  Block
// rows are stored per column ([col][layer][N*N]) so a column's data is
// contiguous.  Storage type T of du / w may be float or double; all solve
// arithmetic is double.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "geom_impl.hpp"
#include "vic_impl.hpp"

namespace snapx {

constexpr int kVicThreads = 128;

struct VicArgs {
  int n1, n2, n3, ng;
  int nl, ncol, ncol2;  // layers, columns, interior columns along x2
  double gammad, area, vol, dt, grav;  // Cartesian face area and volume
  Geometry geo;
  const double* dtp;  // device time step; replaces dt when set
};

__device__ __forceinline__ double step_dt(const VicArgs& a) { return a.dtp ? *a.dtp : a.dt; }

__device__ __forceinline__ int col_cell(const VicArgs& a, int col) {
  int j = a.n2 > 1 ? a.ng + col % a.ncol2 : 0;
  int k = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
  return (k * a.n2 + j) * a.n1 + a.ng;  // first interior cell of the column
}

// face_area1 at the lower and upper x1 faces and cell_volume of layer i of
// a column (src/implicit/implicit_hydro.cpp:147-150): constants on Cartesian
// blocks, r-dependent on spherical-polar ones
struct LayerGeom {
  double area_i, area_ip1, vol;
};

__device__ __forceinline__ LayerGeom layer_geom(const VicArgs& a, int col, int i) {
  if (a.geo.coord == SNAPX_COORD_GNOMONIC_EQUIANGLE) {
    int j = a.n2 > 1 ? a.ng + col % a.ncol2 : 0;
    int k = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
    const GnoCell gc = gno_cell(a.geo, k, j, a.ng + i);
    return {gc.a1m, gc.a1p, gc.vol};
  }
  if (a.geo.coord != SNAPX_COORD_SPHERICAL_POLAR) return {a.area, a.area, a.vol};
  int j = a.n2 > 1 ? a.ng + col % a.ncol2 : 0;
  int k = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
  const SphCell gc = sph_cell(a.geo, k, j, a.ng + i);
  return {gc.a1m, gc.a1p, gc.vol};
}

template <typename T>
__device__ __forceinline__ void load_prim(const T* w, int ncell, int q, double* out) {
#pragma unroll
  for (int v = 0; v < 5; ++v) out[v] = (double)w[v * ncell + q];
}

// vic_assemble_partial_impl / vic_assemble_full_impl for one cell
template <typename T, int N>
__global__ void __launch_bounds__(kVicThreads)
    vic_assemble_kernel(const T* __restrict__ w, double* __restrict__ abc, VicArgs a) {
  int ncell = a.n1 * a.n2 * a.n3;
  int64_t total = (int64_t)a.ncol * a.nl;
  const int NN = N * N;
  const double dt = step_dt(a);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int col = (int)(t / a.nl);
    int i = (int)(t - (int64_t)col * a.nl);
    int q = col_cell(a, col) + i;
    double wm[5], w0[5], wp[5];
    load_prim(w, ncell, q - 1, wm);
    load_prim(w, ncell, q, w0);
    load_prim(w, ncell, q + 1, wp);
    double* base = abc + ((int64_t)col * a.nl + i) * 3 * NN;
    double A[NN], B[NN], C[NN];
    LayerGeom lg = layer_geom(a, col, i);
    vic_assemble_cell<double, N>(wm, w0, wp, a.gammad, lg.area_i, lg.area_ip1, lg.vol, dt,
                                 a.grav, i == 0, i == a.nl - 1, A, B, C);
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      base[e] = A[e];
      base[NN + e] = B[e];
      base[2 * NN + e] = C[e];
    }
  }
}

// ForwardSweep (src/implicit/forward_sweep_impl.h:20-109) +
// vic_backward_reduce (src/implicit/vic_redistribute_impl.h:26-62)
template <typename T, int N>
__global__ void __launch_bounds__(kVicThreads)
    vic_sweep_kernel(const T* __restrict__ du, const T* __restrict__ w, double* __restrict__ abc,
                     double* __restrict__ delta, double* __restrict__ colsum, VicArgs a) {
  int ncell = a.n1 * a.n2 * a.n3;
  const int NN = N * N;
  const double dt = step_dt(a);
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < a.ncol;
       col += gridDim.x * blockDim.x) {
    int q0 = col_cell(a, col);
    double* cabc = abc + (int64_t)col * a.nl * 3 * NN;
    double* cd = delta + (int64_t)col * a.nl * N;
    double aprev[NN], dprev[N];
    for (int i = 0; i < a.nl; ++i) {
      double rhs[N];
      int q = q0 + i;
      rhs[0] = (double)du[IDN * ncell + q];
      rhs[0] /= dt;
      rhs[1] = (double)du[IVX * ncell + q] / dt;
      if (N == 3) {
        rhs[2] = (double)du[IPR * ncell + q] / dt;
      } else {
        rhs[2] = (double)du[IVY * ncell + q] / dt;
        rhs[3] = (double)du[IVZ * ncell + q] / dt;
        rhs[4] = (double)du[IPR * ncell + q] / dt;
      }
      double* ai = cabc + (int64_t)i * 3 * NN;
      double A[NN], B[NN], C[NN];
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        A[e] = ai[e];
        B[e] = ai[NN + e];
        C[e] = ai[2 * NN + e];
      }
      if (i > 0) {
        double tmp[NN], v[N];
        matmul<double, N>(B, aprev, tmp);
#pragma unroll
        for (int e = 0; e < NN; ++e) A[e] -= tmp[e];
        matvec<double, N>(B, dprev, v);
#pragma unroll
        for (int e = 0; e < N; ++e) rhs[e] = rhs[e] - v[e];
      }
      double anew[NN], dnew[N];
      if (N == 3) {
        double inv[9];
        inverse3<double>(A, inv);
        matvec<double, 3>(inv, rhs, dnew);
        matmul<double, 3>(inv, C, anew);
      } else {
        int indx[5];
        ludcmp5<double>(A, indx);
        double solved[30];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
#pragma unroll
          for (int c = 0; c < 5; ++c) solved[r * 6 + c] = C[r * 5 + c];
          solved[r * 6 + 5] = rhs[r];
        }
        lubksb5<double, 6>(A, indx, solved);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
#pragma unroll
          for (int c = 0; c < 5; ++c) anew[r * 5 + c] = solved[r * 6 + c];
          dnew[r] = solved[r * 6 + 5];
        }
      }
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        ai[e] = anew[e];
        aprev[e] = anew[e];
      }
#pragma unroll
      for (int e = 0; e < N; ++e) {
        cd[i * N + e] = dnew[e];
        dprev[e] = dnew[e];
      }
    }
    // back substitution: delta_i -= a_i delta_{i+1}
    for (int i = a.nl - 2; i >= 0; --i) {
      double* ai = cabc + (int64_t)i * 3 * NN;
      double Ai[NN], v[N];
#pragma unroll
      for (int e = 0; e < NN; ++e) Ai[e] = ai[e];
      matvec<double, N>(Ai, dprev, v);
#pragma unroll
      for (int e = 0; e < N; ++e) {
        double d = cd[i * N + e] - v[e];
        cd[i * N + e] = d;
      }
#pragma unroll
      for (int e = 0; e < N; ++e) dprev[e] = cd[i * N + e];
    }
    // column reductions for the MS-VIC redistribution
    double sum_a2 = 0., b_dry = 0.;
    for (int i = 0; i < a.nl; ++i) {
      int q = q0 + i;
      double explicit_total = (double)du[IDN * ncell + q];
      double vol = layer_geom(a, col, i).vol;
      double a_i = (double)w[IDN * ncell + q] * vol;
      sum_a2 += a_i * a_i;
      b_dry += (explicit_total - cd[i * N]) * 1.0 * vol;
    }
    colsum[2 * col] = sum_a2;
    colsum[2 * col + 1] = b_dry;
  }
}

// vic_redistribute_cell (src/implicit/vic_redistribute_impl.h:66-106)
template <typename T, int N>
__global__ void __launch_bounds__(kVicThreads)
    vic_redistribute_kernel(T* __restrict__ du, const T* __restrict__ w,
                            const double* __restrict__ delta, const double* __restrict__ colsum,
                            VicArgs a) {
  int ncell = a.n1 * a.n2 * a.n3;
  int64_t total = (int64_t)a.ncol * a.nl;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int col = (int)(t / a.nl);
    int i = (int)(t - (int64_t)col * a.nl);
    int q = col_cell(a, col) + i;
    const double* d = delta + ((int64_t)col * a.nl + i) * N;
    double sum_a2 = colsum[2 * col];
    double rho = (double)w[IDN * ncell + q];
    double a_i = rho * layer_geom(a, col, i).vol;
    double du0 = (double)du[IDN * ncell + q];
    double diffusion_fix = d[0] - du0;
    double dry_structural = rho * a_i * colsum[2 * col + 1] / sum_a2;
    du[IDN * ncell + q] = (T)((du0 + diffusion_fix * 1.0) + dry_structural);
    du[IVX * ncell + q] = (T)d[1];
    if (N == 3) {
      du[IPR * ncell + q] = (T)d[2];
    } else {
      du[IVY * ncell + q] = (T)d[2];
      du[IVZ * ncell + q] = (T)d[3];
      du[IPR * ncell + q] = (T)d[4];
    }
  }
}

static VicArgs vic_args(const snapx_config* c, double dt, const double* dtp = nullptr) {
  Grid g = make_grid(c);
  VicArgs a;
  a.n1 = g.n1, a.n2 = g.n2, a.n3 = g.n3, a.ng = g.ng;
  a.nl = c->nx1;
  a.ncol2 = g.n2 > 1 ? c->nx2 : 1;
  int ncol3 = g.n3 > 1 ? c->nx3 : 1;
  a.ncol = a.ncol2 * ncol3;
  a.gammad = c->gammad;
  a.area = g.dx3 * g.dx2;
  a.vol = g.dx3 * (g.dx2 * g.dx1);
  a.geo = make_geometry(c);
  a.dt = dt;
  a.dtp = dtp;
  a.grav = c->grav1;
  return a;
}

static int vic_size(const snapx_config* c) { return ((c->implicit_scheme >> 3) & 1) ? 5 : 3; }

size_t implicit_workspace_bytes(const snapx_config* c) {
  if (!c->implicit_scheme) return 0;
  VicArgs a = vic_args(c, 1.0);
  int N = vic_size(c);
  size_t n = (size_t)a.ncol * a.nl;
  return sizeof(double) * (n * 3 * N * N + n * N + 2 * (size_t)a.ncol);
}

template <typename T, int N>
static int vic_launch(const snapx_config* c, double dt, void* du, const void* w, void* ws,
                      hipStream_t st, const double* dtp) {
  VicArgs a = vic_args(c, dt, dtp);
  size_t n = (size_t)a.ncol * a.nl;
  double* abc = reinterpret_cast<double*>(ws);
  double* delta = abc + n * 3 * N * N;
  double* colsum = delta + n * N;
  hipLaunchKernelGGL((vic_assemble_kernel<T, N>), blocks_for(n, kVicThreads), dim3(kVicThreads),
                     0, st, reinterpret_cast<const T*>(w), abc, a);
  SNAPX_LAUNCH_CHECK();
  hipLaunchKernelGGL((vic_sweep_kernel<T, N>), blocks_for(a.ncol, 64), dim3(64), 0, st,
                     reinterpret_cast<const T*>(du), reinterpret_cast<const T*>(w), abc, delta,
                     colsum, a);
  SNAPX_LAUNCH_CHECK();
  hipLaunchKernelGGL((vic_redistribute_kernel<T, N>), blocks_for(n, kVicThreads),
                     dim3(kVicThreads), 0, st, reinterpret_cast<T*>(du),
                     reinterpret_cast<const T*>(w), delta, colsum, a);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

int implicit_correction(const snapx_config* c, double dt, void* du, const void* w,
                        void* workspace, size_t bytes, int dtype, hipStream_t st,
                        const double* dtp) {
  if (!c->implicit_scheme) return 0;
  SNAPX_CHECK_ARG(c->implicit_scheme == 1 || c->implicit_scheme == 9,
                  "implicit correction: unsupported implicit scheme (use 0, 1 or 9)");
  SNAPX_CHECK_ARG(c->nx1 > 1, "implicit correction: x1 axis is inactive");
  SNAPX_CHECK_ARG(bytes >= implicit_workspace_bytes(c),
                  "implicit correction: workspace too small");
  int N = vic_size(c);
  if (dtype == SNAPX_F64)
    return N == 3 ? vic_launch<double, 3>(c, dt, du, w, workspace, st, dtp)
                  : vic_launch<double, 5>(c, dt, du, w, workspace, st, dtp);
  return N == 3 ? vic_launch<float, 3>(c, dt, du, w, workspace, st, dtp)
                : vic_launch<float, 5>(c, dt, du, w, workspace, st, dtp);
}

}  // namespace snapx

using namespace snapx;

extern "C" size_t snapx_implicit_workspace_bytes(const snapx_config* cfg) {
  if (validate_config(cfg)) return 0;
  return implicit_workspace_bytes(cfg);
}

extern "C" int snapx_implicit_correction(const snapx_config* cfg, double dt, void* du,
                                         const void* w, void* workspace, size_t workspace_bytes,
                                         int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(du && w, "snapx_implicit_correction: null pointer");
  SNAPX_CHECK_ARG(!cfg->implicit_scheme || workspace, "snapx_implicit_correction: no workspace");
  return implicit_correction(cfg, dt, du, w, workspace, workspace_bytes, dtype,
                             as_stream(stream), nullptr);
}
