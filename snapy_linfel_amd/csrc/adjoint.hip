// Hand-written discrete adjoint of one SSP-RK3 step of the FP64 plain path.
//
// Reverse of u_{s+1} = BC(a u0 + b u_s + c du(u_s)), stage by stage, with the
// forward stage inputs recomputed from the step input.  du(u_s) is
//   VIC( -dt D(F(recon(w))) + gravity(w) , w ),  w = revise_ghost(cons2prim(u_s)).
// Every local Jacobian (interpolant window, Riemann face, VIC block row,
// ghost refill, cons2prim) is the exact derivative of the forward device code,
// obtained by evaluating that same templated code on Dual<K> numbers; the
// reverse sweep applies them transposed:
//   boundaries   ghost adjoints fold into their source cells
//   VIC          transposed redistribution, transposed block-tridiagonal
//                solve (block Thomas on a^T, c^T, b^T), and the block-row
//                derivatives contracted with -lambda_r delta^T
//   divergence   lambda_F = (-dt A / V)(lambda_du[left] - lambda_du[right])
//   Riemann      J^T lambda_F per face (Dual<10>)
//   recon        gathered per cell over the faces whose window holds it
// All arithmetic is double.  Deterministic: no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "geom_impl.hpp"
#include "interp_impl.hpp"
#include "riemann_impl.hpp"
#include "vic_impl.hpp"
#include "pcr.hpp"
// dual.hpp last: its overloads must not hide the double ones at the template
// definitions above
#include "dual.hpp"
#include "riemann_adj.hpp"

namespace snapx {

constexpr int kT = 128;
constexpr int kFluxDirs = 1;  // input directions per Riemann-adjoint pass

struct AdjArgs {
  int n1, n2, n3, ng;
  int nx1, nx2, nx3;
  double dx1, dx2, dx3, dt, gammad, grav1, grav2, grav3, nh;
  int recon_type[2], recon_scale[2], recon_shock[2];
  int riemann[3];
  int bc[6];
  int scheme;
  int lanes, lshift;  // cell-parallel column kernels: lanes per column, log2
  // 3x3 columns of 33 .. 256 cells: two rows per lane, `half` lanes per
  // column (log2 hshift); 0 for one row per lane
  int half, hshift;
  // spherical-polar blocks: per-cell areas, volumes and source factors
  // (sph_cell), per-layer ghost distances (sph_dx1v)
  int sph;
  Geometry geo;
};

// (j, k) full indices of column col
__device__ __forceinline__ void col_jk(const AdjArgs& a, int col, int& j, int& k) {
  const int ncol2 = a.n2 > 1 ? a.nx2 : 1;
  j = a.n2 > 1 ? a.ng + col % ncol2 : 0;
  k = a.n3 > 1 ? a.ng + col / ncol2 : 0;
}

// spherical-polar layer i (interior-relative, ghosts included) of column
// col: x1 face areas below / above and volume (layer_geom of the column
// solve)
struct SphLayer {
  double am, ap, vol;
};
__device__ __forceinline__ SphLayer sph_layer(const AdjArgs& a, int col, int i) {
  int j, k;
  col_jk(a, col, j, k);
  const SphCell c = sph_cell(a.geo, k, j, a.ng + i);
  return {c.a1m, c.a1p, c.vol};
}

__device__ __forceinline__ void decode(const AdjArgs a, int q, int& i, int& j, int& k) {
  i = q % a.n1;
  j = (q / a.n1) % a.n2;
  k = q / (a.n1 * a.n2);
}

__device__ __forceinline__ bool interior(const AdjArgs a, int i, int j, int k) {
  return (a.n1 == 1 || (i >= a.ng && i < a.n1 - a.ng)) &&
         (a.n2 == 1 || (j >= a.ng && j < a.n2 - a.ng)) &&
         (a.n3 == 1 || (k >= a.ng && k < a.n3 - a.ng));
}

// ------------------------------------------------ boundary-function adjoint

// reverse of one axis of bc_kernel: one thread per (variable, pencil) folds
// the ghost adjoints into their source cells, then zeroes the ghosts
template <typename S>
__global__ void bc_adj_kernel(S* __restrict__ lam, AdjArgs a, int dim, int tin, int tout,
                              int flip) {
  int s, n, np;
  if (dim == 3) {
    s = 1, n = a.n1, np = a.n2 * a.n3;
  } else if (dim == 2) {
    s = a.n1, n = a.n2, np = a.n1 * a.n3;
  } else {
    s = a.n1 * a.n2, n = a.n3, np = a.n1 * a.n2;
  }
  int ncell = a.n1 * a.n2 * a.n3;
  int ng = a.ng;
  int vnorm = 4 - dim;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 5 * np; t += gridDim.x * blockDim.x) {
    int p = t % np, v = t / np;
    int base = dim == 3 ? p * a.n1 : (dim == 2 ? (p / a.n1) * a.n1 * a.n2 + p % a.n1 : p);
    S* x = lam + v * ncell + base;
    for (int side = 0; side < 2; ++side) {
      int type = side ? tout : tin;
      if (type < 0) continue;
      for (int g = 0; g < ng; ++g) {
        int dst, src;
        if (!side) {
          dst = g;
          src = type == SNAPX_BC_REFLECTING ? 2 * ng - 1 - g
                : type == SNAPX_BC_OUTFLOW  ? ng
                                            : n - 2 * ng + g;
        } else {
          dst = n - ng + g;
          src = type == SNAPX_BC_REFLECTING ? n - ng - 1 - g
                : type == SNAPX_BC_OUTFLOW  ? n - ng - 1
                                            : ng + g;
        }
        const bool neg = type == SNAPX_BC_REFLECTING && v == vnorm && flip;
        x[src * s] += neg ? -x[dst * s] : x[dst * s];
        x[dst * s] = S(0);
      }
    }
  }
}

// --------------------------------------------------------- stage combine

template <typename S>
__global__ void combine_adj_kernel(const S* __restrict__ lam, S* __restrict__ lam_du,
                                   S* __restrict__ lam_us, S* __restrict__ lam_u0, double wa,
                                   double wb, double wc, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const double l = lam[q];
    lam_du[q] = S(wc * l);
    lam_us[q] = S(wb * l);
    lam_u0[q] = S((double)lam_u0[q] + wa * l);
  }
}

// ------------------------------------------------------------------- VIC

__device__ __forceinline__ int col_q0(const AdjArgs a, int col) {
  int ncol2 = a.n2 > 1 ? a.nx2 : 1;
  int j = a.n2 > 1 ? a.ng + col % ncol2 : 0;
  int k = a.n3 > 1 ? a.ng + col / ncol2 : 0;
  return (k * a.n2 + j) * a.n1 + a.ng;
}

template <typename T, typename S>
__device__ __forceinline__ void load_w(const S* w, int ncell, int q, T* out) {
#pragma unroll
  for (int v = 0; v < 5; ++v) out[v] = T((double)w[v * ncell + q]);
}

// One thread per column: forward assembly + block Thomas (delta), MS-VIC
// redistribution adjoint, transposed block solve.  Writes
//   lam_w   (rho part of the redistribution)  += ...
//   lam_dx  (adjoint of the explicit tendency) = ...
//   dl      [col][layer][2N]: delta and lambda_r for the row-derivative kernel
template <int N, typename S>
__global__ void __launch_bounds__(64)
    vic_adj_column_kernel(const S* __restrict__ w, const S* __restrict__ dux,
                          const S* __restrict__ lam_du, S* __restrict__ lam_dx,
                          S* __restrict__ lam_w, double* __restrict__ abc,
                          double* __restrict__ pr, double* __restrict__ dl, AdjArgs a, int ncol) {
  const int NN = N * N;
  const int ncell = a.n1 * a.n2 * a.n3;
  const int nl = a.nx1;
  const double area = a.dx3 * a.dx2, vol = a.dx3 * (a.dx2 * a.dx1);
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < ncol;
       col += gridDim.x * blockDim.x) {
    int q0 = col_q0(a, col);
    double* cabc = abc + (int64_t)col * nl * 3 * NN;
    double* cpr = pr + (int64_t)col * nl * (NN + N);
    double* cdl = dl + (int64_t)col * nl * 2 * N;
    // ---- assembly of every block row ----
    for (int i = 0; i < nl; ++i) {
      double wm[5], w0[5], wp[5];
      load_w(w, ncell, q0 + i - 1, wm);
      load_w(w, ncell, q0 + i, w0);
      load_w(w, ncell, q0 + i + 1, wp);
      vic_assemble_cell<double, N>(wm, w0, wp, a.gammad, area, area, vol, a.dt, a.grav1, i == 0,
                                   i == nl - 1, cabc + i * 3 * NN, cabc + i * 3 * NN + NN,
                                   cabc + i * 3 * NN + 2 * NN);
    }
    // ---- forward sweep + back substitution -> delta ----
    double aprev[NN], dprev[N];
    for (int i = 0; i < nl; ++i) {
      int q = q0 + i;
      double rhs[N];
      rhs[0] = (double)dux[IDN * ncell + q] / a.dt;
      rhs[1] = (double)dux[IVX * ncell + q] / a.dt;
      if (N == 3) {
        rhs[2] = (double)dux[IPR * ncell + q] / a.dt;
      } else {
        rhs[2] = (double)dux[IVY * ncell + q] / a.dt;
        rhs[3] = (double)dux[IVZ * ncell + q] / a.dt;
        rhs[4] = (double)dux[IPR * ncell + q] / a.dt;
      }
      double A[NN], B[NN], C[NN];
      for (int e = 0; e < NN; ++e) {
        A[e] = cabc[i * 3 * NN + e];
        B[e] = cabc[i * 3 * NN + NN + e];
        C[e] = cabc[i * 3 * NN + 2 * NN + e];
      }
      if (i > 0) {
        double tmp[NN], v[N];
        matmul<double, N>(B, aprev, tmp);
        for (int e = 0; e < NN; ++e) A[e] -= tmp[e];
        matvec<double, N>(B, dprev, v);
        for (int e = 0; e < N; ++e) rhs[e] -= v[e];
      }
      if (N == 3) {
        double inv[9];
        inverse3<double>(A, inv);
        matvec<double, 3>(inv, rhs, dprev);
        matmul<double, 3>(inv, C, aprev);
      } else {
        int indx[5];
        ludcmp5<double>(A, indx);
        double solved[30];
        for (int r = 0; r < 5; ++r) {
          for (int c = 0; c < 5; ++c) solved[r * 6 + c] = C[r * 5 + c];
          solved[r * 6 + 5] = rhs[r];
        }
        lubksb5<double, 6>(A, indx, solved);
        for (int r = 0; r < 5; ++r) {
          for (int c = 0; c < 5; ++c) aprev[r * 5 + c] = solved[r * 6 + c];
          dprev[r] = solved[r * 6 + 5];
        }
      }
      for (int e = 0; e < NN; ++e) cpr[i * (NN + N) + e] = aprev[e];
      for (int e = 0; e < N; ++e) cpr[i * (NN + N) + NN + e] = dprev[e];
    }
    for (int e = 0; e < N; ++e) cdl[(nl - 1) * 2 * N + e] = dprev[e];
    for (int i = nl - 2; i >= 0; --i) {
      double v[N];
      matvec<double, N>(cpr + i * (NN + N), dprev, v);
      for (int e = 0; e < N; ++e) {
        dprev[e] = cpr[i * (NN + N) + NN + e] - v[e];
        cdl[i * 2 * N + e] = dprev[e];
      }
    }
    // ---- redistribution adjoint ----
    double Ssum = 0., Bd = 0., G = 0.;
    for (int i = 0; i < nl; ++i) {
      int q = q0 + i;
      double rho = w[IDN * ncell + q];
      double ai = rho * vol;
      Ssum += ai * ai;
      Bd += ((double)dux[IDN * ncell + q] - cdl[i * 2 * N]) * 1.0 * vol;
      G += (double)lam_du[IDN * ncell + q] * rho * ai;
    }
    G /= Ssum;
    double lamS = -G * Bd / Ssum;
    // lambda_delta into the second half of dl (temporarily), direct lam_dx
    for (int i = 0; i < nl; ++i) {
      int q = q0 + i;
      double rho = w[IDN * ncell + q];
      double ln[5];
      for (int v = 0; v < 5; ++v) ln[v] = lam_du[v * ncell + q];
      double ld[N];
      ld[0] = ln[IDN] - G * vol;
      ld[1] = ln[IVX];
      if (N == 3) {
        ld[2] = ln[IPR];
        lam_dx[IVY * ncell + q] = ln[IVY];
        lam_dx[IVZ * ncell + q] = ln[IVZ];
      } else {
        ld[2] = ln[IVY];
        ld[3] = ln[IVZ];
        ld[4] = ln[IPR];
      }
      lam_dx[IDN * ncell + q] = G * vol;
      lam_w[IDN * ncell + q] =
          (double)lam_w[IDN * ncell + q] + ln[IDN] * 2. * rho * vol * Bd / Ssum + lamS * 2. * rho * vol * vol;
      for (int e = 0; e < N; ++e) cdl[i * 2 * N + N + e] = ld[e];
    }
    // ---- transposed block solve  M^T y = lambda_delta ----
    // rows of M^T: diagonal a_i^T, lower c_{i-1}^T, upper b_{i+1}^T
    double yprev[N];
    for (int i = 0; i < nl; ++i) {
      double A[NN], Bt[NN], Ct[NN], rhs[N];
      const double* ai = cabc + i * 3 * NN;
      for (int r = 0; r < N; ++r)
        for (int c = 0; c < N; ++c) {
          A[r * N + c] = ai[c * N + r];
          Bt[r * N + c] = i > 0 ? cabc[(i - 1) * 3 * NN + 2 * NN + c * N + r] : 0.;
          Ct[r * N + c] = i < nl - 1 ? cabc[(i + 1) * 3 * NN + NN + c * N + r] : 0.;
        }
      for (int e = 0; e < N; ++e) rhs[e] = cdl[i * 2 * N + N + e];
      if (i > 0) {
        double tmp[NN], v[N];
        matmul<double, N>(Bt, aprev, tmp);
        for (int e = 0; e < NN; ++e) A[e] -= tmp[e];
        matvec<double, N>(Bt, yprev, v);
        for (int e = 0; e < N; ++e) rhs[e] -= v[e];
      }
      if (N == 3) {
        double inv[9];
        inverse3<double>(A, inv);
        matvec<double, 3>(inv, rhs, yprev);
        matmul<double, 3>(inv, Ct, aprev);
      } else {
        int indx[5];
        ludcmp5<double>(A, indx);
        double solved[30];
        for (int r = 0; r < 5; ++r) {
          for (int c = 0; c < 5; ++c) solved[r * 6 + c] = Ct[r * 5 + c];
          solved[r * 6 + 5] = rhs[r];
        }
        lubksb5<double, 6>(A, indx, solved);
        for (int r = 0; r < 5; ++r) {
          for (int c = 0; c < 5; ++c) aprev[r * 5 + c] = solved[r * 6 + c];
          yprev[r] = solved[r * 6 + 5];
        }
      }
      for (int e = 0; e < NN; ++e) cpr[i * (NN + N) + e] = aprev[e];
      for (int e = 0; e < N; ++e) cpr[i * (NN + N) + NN + e] = yprev[e];
    }
    for (int e = 0; e < N; ++e) cdl[(nl - 1) * 2 * N + N + e] = yprev[e];
    for (int i = nl - 2; i >= 0; --i) {
      double v[N];
      matvec<double, N>(cpr + i * (NN + N), yprev, v);
      for (int e = 0; e < N; ++e) {
        yprev[e] = cpr[i * (NN + N) + NN + e] - v[e];
        cdl[i * 2 * N + N + e] = yprev[e];
      }
    }
    // lambda of the explicit tendency through r = du / dt
    for (int i = 0; i < nl; ++i) {
      int q = q0 + i;
      const double* y = cdl + i * 2 * N + N;
      lam_dx[IDN * ncell + q] = (double)lam_dx[IDN * ncell + q] + y[0] / a.dt;
      lam_dx[IVX * ncell + q] = y[1] / a.dt;
      if (N == 3) {
        lam_dx[IPR * ncell + q] = y[2] / a.dt;
      } else {
        lam_dx[IVY * ncell + q] = y[2] / a.dt;
        lam_dx[IVZ * ncell + q] = y[3] / a.dt;
        lam_dx[IPR * ncell + q] = y[4] / a.dt;
      }
    }
  }
}

// ---------------------------------------- cell-parallel VIC (forward+adjoint)

// block rows a_i, b_i, c_i of every interior cell, abc [col][layer][3][NN]
template <int N, typename S, bool kSph>
__global__ void __launch_bounds__(kT)
    vic_asm_kernel(const S* __restrict__ w, double* __restrict__ abc, AdjArgs a, int ncol) {
  const int NN = N * N;
  const int ncell = a.n1 * a.n2 * a.n3;
  const int nl = a.nx1;
  const double area = a.dx3 * a.dx2, vol = a.dx3 * (a.dx2 * a.dx1);
  int64_t total = (int64_t)ncol * nl;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int col = (int)(t / nl), i = (int)(t - (int64_t)col * nl);
    int q = col_q0(a, col) + i;
    double wm[5], w0[5], wp[5];
    load_w(w, ncell, q - 1, wm);
    load_w(w, ncell, q, w0);
    load_w(w, ncell, q + 1, wp);
    double* row = abc + t * 3 * NN;
    if constexpr (kSph) {
      const SphLayer lg = sph_layer(a, col, i);
      vic_assemble_cell<double, N>(wm, w0, wp, a.gammad, lg.am, lg.ap, lg.vol, a.dt, a.grav1,
                                   i == 0, i == nl - 1, row, row + NN, row + 2 * NN);
      continue;
    }
    vic_assemble_cell<double, N>(wm, w0, wp, a.gammad, area, area, vol, a.dt, a.grav1, i == 0,
                                 i == nl - 1, row, row + NN, row + 2 * NN);
  }
}

// rhs layout of the column solve: (rho, v1, p) or (rho, v1, v2, v3, p)
template <int N, typename S>
__device__ __forceinline__ void vic_pick(const S* x, int ncell, int q, double* r) {
  r[0] = x[IDN * ncell + q];
  r[1] = x[IVX * ncell + q];
  if (N == 3) {
    r[2] = x[IPR * ncell + q];
  } else {
    r[2] = x[IVY * ncell + q];
    r[3] = x[IVZ * ncell + q];
    r[4] = x[IPR * ncell + q];
  }
}

// Lane geometry of the cell-parallel column kernels: lane t of the
// workgroup serves layer i of column col.
struct ColLane {
  int col, i, q;
  bool act;
};

__device__ __forceinline__ ColLane col_lane(const AdjArgs a, int ncol) {
  ColLane L;
  L.col = blockIdx.x * (blockDim.x >> a.lshift) + (threadIdx.x >> a.lshift);
  L.i = threadIdx.x & (a.lanes - 1);
  L.act = L.col < ncol && L.i < a.nx1;
  L.q = L.act ? col_q0(a, L.col) + L.i : 0;
  return L;
}

// row i of M (transpose = false) or of M^T (transpose = true) from abc; the
// folded end rows carry no coupling outside the column
template <int N>
__device__ __forceinline__ void vic_row_blocks(const double* __restrict__ abc, int64_t col, int i,
                                               int nl, bool transpose, double* A, double* B,
                                               double* C) {
  const int NN = N * N;
  const double* row = abc + (col * nl + i) * 3 * NN;
  if (!transpose) {
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      A[e] = row[e];
      B[e] = i > 0 ? row[NN + e] : 0.;
      C[e] = i < nl - 1 ? row[2 * NN + e] : 0.;
    }
    return;
  }
  // M^T: diagonal a_i^T, lower c_{i-1}^T, upper b_{i+1}^T
  const double* lo = row - 3 * NN + 2 * NN;
  const double* up = row + 3 * NN + NN;
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      A[r * N + c] = row[c * N + r];
      B[r * N + c] = i > 0 ? lo[c * N + r] : 0.;
      C[r * N + c] = i < nl - 1 ? up[c * N + r] : 0.;
    }
}

// Forward column solve + MS-VIC redistribution of one stage (the recompute
// of the stage states): du (interior) from the explicit tendency dux; delta
// is kept in the first half of dl [col][layer][2N] for the reverse sweep.
template <int N, typename S, bool kSph>
__global__ void __launch_bounds__(PcrShape<N>::kThreads)
    vic_fwd_solve_kernel(const S* __restrict__ w, const S* __restrict__ dux,
                         const double* __restrict__ abc, S* __restrict__ du,
                         double* __restrict__ dl, AdjArgs a, int ncol) {
  extern __shared__ double sh[];
  const int NN = N * N;
  const int ncell = a.n1 * a.n2 * a.n3, nl = a.nx1;
  ColLane L = col_lane(a, ncol);
  const double vol = !kSph ? a.dx3 * (a.dx2 * a.dx1) : (L.act ? sph_layer(a, L.col, L.i).vol : 1.);
  double A[NN], B[NN], C[NN], r[N], d[N];
  if (L.act) {
    vic_row_blocks<N>(abc, L.col, L.i, nl, false, A, B, C);
    vic_pick<N>(dux, ncell, L.q, r);
#pragma unroll
    for (int e = 0; e < N; ++e) r[e] /= a.dt;
  }
  pcr_solve<N>(A, B, C, r, d, sh, blockDim.x, threadIdx.x, L.i, nl, L.act, a.lanes == 64);
  double rho = L.act ? w[IDN * ncell + L.q] : 0.;
  double xr = L.act ? dux[IDN * ncell + L.q] : 0.;
  double ai = rho * vol;
  double sums[2] = {ai * ai, L.act ? (xr - d[0]) * 1.0 * vol : 0.};
  column_sums<2>(sums, sh, threadIdx.x, a.lanes, a.lshift);
  if (!L.act) return;
  double* cdl = dl + ((int64_t)L.col * nl + L.i) * 2 * N;
#pragma unroll
  for (int e = 0; e < N; ++e) cdl[e] = d[e];
  du[IDN * ncell + L.q] = (xr + (d[0] - xr)) + rho * ai * sums[1] / sums[0];
  du[IVX * ncell + L.q] = d[1];
  if (N == 3) {
    du[IPR * ncell + L.q] = d[2];
  } else {
    du[IVY * ncell + L.q] = d[2];
    du[IVZ * ncell + L.q] = d[3];
    du[IPR * ncell + L.q] = d[4];
  }
}

// Adjoint of the column solve + redistribution, with delta = M^-1 (dux / dt)
// from the forward recompute (first half of dl): the redistribution adjoint
// (column sums S, Bd, G) and y = M^-T lambda_delta by PCR.  Writes y to the
// second half of dl, lam_dx (interior) and the rho part of lam_w; lam_dx
// enters holding lam_du.
template <int N, typename S, bool kSph>
__global__ void __launch_bounds__(PcrShape<N>::kThreads)
    vic_adj_solve_kernel(const S* __restrict__ w, const S* __restrict__ dux,
                         const S* __restrict__ lam_du, const double* __restrict__ abc,
                         S* __restrict__ lam_dx, S* __restrict__ lam_w,
                         double* __restrict__ dl, AdjArgs a, int ncol) {
  extern __shared__ double sh[];
  const int NN = N * N;
  const int ncell = a.n1 * a.n2 * a.n3, nl = a.nx1;
  ColLane L = col_lane(a, ncol);
  const double vol = !kSph ? a.dx3 * (a.dx2 * a.dx1) : (L.act ? sph_layer(a, L.col, L.i).vol : 1.);
  double* cdl = dl + ((int64_t)L.col * nl + L.i) * 2 * N;
  double d0 = L.act ? cdl[0] : 0.;
  double rho = L.act ? w[IDN * ncell + L.q] : 0.;
  double lr = L.act ? lam_du[IDN * ncell + L.q] : 0.;
  double ai = rho * vol;
  double sums[3] = {ai * ai, L.act ? ((double)dux[IDN * ncell + L.q] - d0) * 1.0 * vol : 0.,
                    lr * rho * ai};
  column_sums<3>(sums, sh, threadIdx.x, a.lanes, a.lshift);
  const double Ssum = sums[0], Bd = sums[1], G = sums[2] / Ssum;
  const double lamS = -G * Bd / Ssum;
  double A[NN], B[NN], C[NN], ld[N];
  if (L.act) {
    vic_pick<N>(lam_du, ncell, L.q, ld);
    ld[0] = lr - G * vol;
    lam_w[IDN * ncell + L.q] += lr * 2. * rho * vol * Bd / Ssum + lamS * 2. * rho * vol * vol;
    vic_row_blocks<N>(abc, L.col, L.i, nl, true, A, B, C);
  }
  double y[N];
  pcr_solve<N>(A, B, C, ld, y, sh, blockDim.x, threadIdx.x, L.i, nl, L.act, a.lanes == 64);
  if (!L.act) return;
#pragma unroll
  for (int e = 0; e < N; ++e) cdl[N + e] = y[e];
  lam_dx[IDN * ncell + L.q] = G * vol + y[0] / a.dt;
  lam_dx[IVX * ncell + L.q] = y[1] / a.dt;
  if (N == 3) {
    lam_dx[IPR * ncell + L.q] = y[2] / a.dt;
  } else {
    lam_dx[IVY * ncell + L.q] = y[2] / a.dt;
    lam_dx[IVZ * ncell + L.q] = y[3] / a.dt;
    lam_dx[IPR * ncell + L.q] = y[4] / a.dt;
  }
}

// Solves M x = rhs (transpose false) or M^T x = rhs (transpose true) for the
// rows 2m and 2m + 1 of column col (3 x 3 blocks of the tape) on lane m of
// the column's `half` lanes: one cyclic-reduction level inside the lane
// removes the even rows (the odd row eliminates the even row below from its
// own registers and the one above from the next lane through LDS), PCR
// solves the odd rows, and the even rows follow by back substitution.
// Inactive rows are identity rows with a zero right-hand side.  Every lane
// of the workgroup calls it.
__device__ __forceinline__ void column_solve2(const double* __restrict__ abc, int64_t col, int m,
                                              int nl, int half, bool transpose, bool act0,
                                              bool act1, const double* r0, const double* r1,
                                              double* x0, double* x1, double* __restrict__ sh,
                                              int T, int t, bool wave) {
  constexpr int N = 3, NN = 9;
  double Pb0[NN], Pc0[NN], pr0[N];  // the even row, normalised
  {
    double A[NN], B[NN], C[NN], rr[N], inv[NN];
    if (act0) {
      vic_row_blocks<N>(abc, col, 2 * m, nl, transpose, A, B, C);
    } else {
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        A[e] = (e % (N + 1)) == 0 ? 1. : 0.;
        B[e] = C[e] = 0.;
      }
    }
#pragma unroll
    for (int e = 0; e < N; ++e) rr[e] = act0 ? r0[e] : 0.;
    inverse3<double, false>(A, inv);
    matmul<double, N>(inv, B, Pb0);
    matmul<double, N>(inv, C, Pc0);
    matvec<double, N>(inv, rr, pr0);
  }
#pragma unroll
  for (int e = 0; e < NN; ++e) {
    sh[e * T + t] = Pb0[e];
    sh[(NN + e) * T + t] = Pc0[e];
  }
#pragma unroll
  for (int e = 0; e < N; ++e) sh[(2 * NN + e) * T + t] = pr0[e];
  // the odd row
  double A1[NN], B1[NN], C1[NN], q1[N];
  if (act1) {
    vic_row_blocks<N>(abc, col, 2 * m + 1, nl, transpose, A1, B1, C1);
  } else {
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      A1[e] = (e % (N + 1)) == 0 ? 1. : 0.;
      B1[e] = C1[e] = 0.;
    }
  }
#pragma unroll
  for (int e = 0; e < N; ++e) q1[e] = act1 ? r1[e] : 0.;
  column_barrier(wave);
  {
    double nB[NN], nC[NN], y[NN];
    // the even row below (own registers) couples the odd row to row 2m - 1
    matmul<double, N>(B1, Pb0, y);
#pragma unroll
    for (int e = 0; e < NN; ++e) nB[e] = -y[e];
    matmul<double, N>(B1, Pc0, y);
#pragma unroll
    for (int e = 0; e < NN; ++e) A1[e] -= y[e];
    matvec<double, N>(B1, pr0, y);
#pragma unroll
    for (int e = 0; e < N; ++e) q1[e] -= y[e];
    // the even row above (next lane) couples it to row 2m + 3
#pragma unroll
    for (int e = 0; e < NN; ++e) nC[e] = 0.;
    if (m + 1 < half) {
      double Pb[NN], Pc[NN], pr[N];
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        Pb[e] = sh[e * T + t + 1];
        Pc[e] = sh[(NN + e) * T + t + 1];
      }
#pragma unroll
      for (int e = 0; e < N; ++e) pr[e] = sh[(2 * NN + e) * T + t + 1];
      matmul<double, N>(C1, Pc, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) nC[e] = -y[e];
      matmul<double, N>(C1, Pb, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) A1[e] -= y[e];
      matvec<double, N>(C1, pr, y);
#pragma unroll
      for (int e = 0; e < N; ++e) q1[e] -= y[e];
    }
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      B1[e] = nB[e];
      C1[e] = nC[e];
    }
  }
  column_barrier(wave);
  pcr_solve<N>(A1, B1, C1, q1, x1, sh, T, t, m, nl / 2, act1, wave);
  // the even rows by back substitution
#pragma unroll
  for (int e = 0; e < N; ++e) sh[e * T + t] = x1[e];
  column_barrier(wave);
  double xb[N];
#pragma unroll
  for (int e = 0; e < N; ++e) xb[e] = m > 0 ? sh[e * T + t - 1] : 0.;
#pragma unroll
  for (int rr = 0; rr < N; ++rr) {
    double v = pr0[rr];
#pragma unroll
    for (int c = 0; c < N; ++c) v -= Pb0[rr * N + c] * xb[c] + Pc0[rr * N + c] * x1[c];
    x0[rr] = v;
  }
  column_barrier(wave);
}

// Lane geometry of the two-rows-per-lane column kernels: lane m of column
// col serves layers 2m and 2m + 1.
struct ColLane2 {
  int col, m, qb;
  bool act0, act1;
};

__device__ __forceinline__ ColLane2 col_lane2(const AdjArgs& a, int ncol) {
  ColLane2 L;
  L.col = blockIdx.x * (blockDim.x >> a.hshift) + (threadIdx.x >> a.hshift);
  L.m = threadIdx.x & (a.half - 1);
  const bool colok = L.col < ncol;
  L.act0 = colok && 2 * L.m < a.nx1;
  L.act1 = colok && 2 * L.m + 1 < a.nx1;
  L.qb = colok ? col_q0(a, L.col) : 0;
  return L;
}

// vic_fwd_solve_kernel with two rows per lane (3x3 blocks)
template <typename S, bool kSph>
__global__ void __launch_bounds__(PcrShape<3>::kThreads, 2)
    vic_fwd_solve2_kernel(const S* __restrict__ w, const S* __restrict__ dux,
                          const double* __restrict__ abc, S* __restrict__ du,
                          double* __restrict__ dl, AdjArgs a, int ncol) {
  extern __shared__ double sh[];
  constexpr int N = 3;
  const int ncell = a.n1 * a.n2 * a.n3, nl = a.nx1;
  const ColLane2 L = col_lane2(a, ncol);
  const bool act[2] = {L.act0, L.act1};
  double r[2][N], x[2][N], vol[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = 2 * L.m + h;
    vol[h] = !kSph ? a.dx3 * (a.dx2 * a.dx1) : (act[h] ? sph_layer(a, L.col, i).vol : 1.);
#pragma unroll
    for (int e = 0; e < N; ++e) r[h][e] = 0.;
    if (act[h]) {
      vic_pick<N>(dux, ncell, L.qb + i, r[h]);
#pragma unroll
      for (int e = 0; e < N; ++e) r[h][e] /= a.dt;
    }
  }
  column_solve2(abc, L.col, L.m, nl, a.half, false, L.act0, L.act1, r[0], r[1], x[0], x[1], sh,
                blockDim.x, threadIdx.x, a.half <= 64);
  double rho[2], xr[2], ai[2];
  double sums[2] = {0., 0.};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int q = L.qb + 2 * L.m + h;
    rho[h] = act[h] ? w[IDN * ncell + q] : 0.;
    xr[h] = act[h] ? dux[IDN * ncell + q] : 0.;
    ai[h] = rho[h] * vol[h];
    sums[0] += ai[h] * ai[h];
    sums[1] += act[h] ? (xr[h] - x[h][0]) * 1.0 * vol[h] : 0.;
  }
  column_sums<2>(sums, sh, threadIdx.x, a.half, a.hshift);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (!act[h]) continue;
    const int i = 2 * L.m + h, q = L.qb + i;
    double* cdl = dl + ((int64_t)L.col * nl + i) * 2 * N;
#pragma unroll
    for (int e = 0; e < N; ++e) cdl[e] = x[h][e];
    du[IDN * ncell + q] = (xr[h] + (x[h][0] - xr[h])) + rho[h] * ai[h] * sums[1] / sums[0];
    du[IVX * ncell + q] = x[h][1];
    du[IPR * ncell + q] = x[h][2];
  }
}

// vic_adj_solve_kernel with two rows per lane (3x3 blocks)
template <typename S, bool kSph>
__global__ void __launch_bounds__(PcrShape<3>::kThreads, 2)
    vic_adj_solve2_kernel(const S* __restrict__ w, const S* __restrict__ dux,
                          const S* __restrict__ lam_du, const double* __restrict__ abc,
                          S* __restrict__ lam_dx, S* __restrict__ lam_w,
                          double* __restrict__ dl, AdjArgs a, int ncol) {
  extern __shared__ double sh[];
  constexpr int N = 3;
  const int ncell = a.n1 * a.n2 * a.n3, nl = a.nx1;
  const ColLane2 L = col_lane2(a, ncol);
  const bool act[2] = {L.act0, L.act1};
  double vol[2], rho[2], lr[2];
  double sums[3] = {0., 0., 0.};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = 2 * L.m + h, q = L.qb + i;
    vol[h] = !kSph ? a.dx3 * (a.dx2 * a.dx1) : (act[h] ? sph_layer(a, L.col, i).vol : 1.);
    const double d0 = act[h] ? dl[((int64_t)L.col * nl + i) * 2 * N] : 0.;
    rho[h] = act[h] ? w[IDN * ncell + q] : 0.;
    lr[h] = act[h] ? lam_du[IDN * ncell + q] : 0.;
    const double ai = rho[h] * vol[h];
    sums[0] += ai * ai;
    sums[1] += act[h] ? ((double)dux[IDN * ncell + q] - d0) * 1.0 * vol[h] : 0.;
    sums[2] += lr[h] * rho[h] * ai;
  }
  column_sums<3>(sums, sh, threadIdx.x, a.half, a.hshift);
  const double Ssum = sums[0], Bd = sums[1], G = sums[2] / Ssum;
  const double lamS = -G * Bd / Ssum;
  double ld[2][N], y[2][N];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int e = 0; e < N; ++e) ld[h][e] = 0.;
    if (!act[h]) continue;
    const int q = L.qb + 2 * L.m + h;
    vic_pick<N>(lam_du, ncell, q, ld[h]);
    ld[h][0] = lr[h] - G * vol[h];
    lam_w[IDN * ncell + q] += lr[h] * 2. * rho[h] * vol[h] * Bd / Ssum +
                              lamS * 2. * rho[h] * vol[h] * vol[h];
  }
  column_solve2(abc, L.col, L.m, nl, a.half, true, L.act0, L.act1, ld[0], ld[1], y[0], y[1], sh,
                blockDim.x, threadIdx.x, a.half <= 64);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (!act[h]) continue;
    const int i = 2 * L.m + h, q = L.qb + i;
    double* cdl = dl + ((int64_t)L.col * nl + i) * 2 * N;
#pragma unroll
    for (int e = 0; e < N; ++e) cdl[N + e] = y[h][e];
    lam_dx[IDN * ncell + q] = G * vol[h] + y[h][0] / a.dt;
    lam_dx[IVX * ncell + q] = y[h][1] / a.dt;
    lam_dx[IPR * ncell + q] = y[h][2] / a.dt;
  }
}

// delta and y of layer k extended past the column ends: the folded end rows
// act as delta_{-1} = Bnd delta_0, delta_nl = Bnd delta_{nl-1}, with no row
// (y = 0) outside [0, nl)
template <int N>
__device__ __forceinline__ void ext_delta_y(const double* cdl, int k, int nl, double* d,
                                            double* y) {
  int kk = k < 0 ? 0 : (k >= nl ? nl - 1 : k);
  bool out = kk != k;
#pragma unroll
  for (int e = 0; e < N; ++e) {
    d[e] = cdl[kk * 2 * N + e] * ((out && e == 1) ? -1. : 1.);
    y[e] = out ? 0. : cdl[kk * 2 * N + N + e];
  }
}

// Roe-averaged (r, u, v, w, p) of an interface, the arithmetic of
// abs_jacobian's averaging step
template <typename T>
__device__ __forceinline__ void roe_state(T gm1, const T* wl, const T* wr, T* z) {
  T sqrtdl = sqrt(wl[IDN]);
  T sqrtdr = sqrt(wr[IDN]);
  T isdlpdr = T(1.0) / (sqrtdl + sqrtdr);
  z[0] = sqrtdl * sqrtdr;
  z[1] = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  z[2] = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  z[3] = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  T el = wl[IPR] / gm1 + T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  T er = wr[IPR] / gm1 + T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));
  T hbar = ((el + wl[IPR]) / sqrtdl + (er + wr[IPR]) / sqrtdr) * isdlpdr;
  z[4] = (hbar - T(0.5) * ((sq(z[1]) + sq(z[2])) + sq(z[3]))) * gm1 / (gm1 + T(1.)) * z[0];
}

// P^T |A| Q with |A| = R |Lambda| R^-1 at the Roe state z (the matrices of
// abs_jacobian, contracted without forming them); P, Q are 5-vectors
template <typename T>
__device__ __forceinline__ T abs_contract(T gm1, const T* z, const double* P, const double* Q) {
  T r = z[0], u = z[1], v = z[2], w = z[3], p = z[4];
  T cs = sqrt(p * (gm1 + T(1.)) / r);
  T ke = T(0.5) * ((u * u + v * v) + w * w);
  T hp = (gm1 + T(1.)) / gm1 * p / r;
  T h = hp + ke;
  T lam[5] = {fabs(u - cs), fabs(u), fabs(u + cs), fabs(u), fabs(u)};
  T base = (T(P[0]) + T(P[2]) * v) + T(P[3]) * w;
  T PR[5] = {base + T(P[1]) * (u - cs) + T(P[4]) * (h - u * cs),
             base + T(P[1]) * u + T(P[4]) * ke,
             base + T(P[1]) * (u + cs) + T(P[4]) * (h + u * cs),
             T(P[2]) + T(P[4]) * v, T(P[3]) + T(P[4]) * w};
  T two_cs_hp = T(2.) * cs * hp, two_hp = T(2.) * hp;
  T tang = (-v * T(Q[2]) - w * T(Q[3]) + T(Q[4])) / two_hp;
  T RQ[5] = {((cs * ke + u * hp) * T(Q[0]) + (-hp - cs * u) * T(Q[1])) / two_cs_hp + tang,
             ((hp - ke) * T(Q[0]) + u * T(Q[1]) + v * T(Q[2]) + w * T(Q[3]) - T(Q[4])) / hp,
             ((cs * ke - u * hp) * T(Q[0]) + (hp - cs * u) * T(Q[1])) / two_cs_hp + tang,
             -v * T(Q[0]) + T(Q[2]), -w * T(Q[0]) + T(Q[3])};
  T phi = T(0.);
#pragma unroll
  for (int k = 0; k < 5; ++k) phi += PR[k] * lam[k] * RQ[k];
  return phi;
}

// Face f (between layers f-1 and f, f in [0, nl]) of the row-derivative
// contraction: phi_f = (y_{f-1} - y_f)^T |A|_f (delta_{f-1} - delta_f);
// G[col][f][0..4] = d phi / d w_{f-1}, G[col][f][5..9] = d phi / d w_f.
// The Roe state is differentiated with Dual<10>, phi(z) with Dual<5>.
template <int N, typename S, bool kSph>
__global__ void __launch_bounds__(kT)
    vic_adj_face_kernel(const S* __restrict__ w, const double* __restrict__ dl,
                        double* __restrict__ G, AdjArgs a, int ncol) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int nl = a.nx1;
  const double gm1 = a.gammad - 1.;
  int64_t total = (int64_t)ncol * (nl + 1);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int col = (int)(t / (nl + 1)), f = (int)(t - (int64_t)col * (nl + 1));
    const double* cdl = dl + (int64_t)col * nl * 2 * N;
    double dm[N], ym[N], dp[N], yp[N];
    ext_delta_y<N>(cdl, f - 1, nl, dm, ym);
    ext_delta_y<N>(cdl, f, nl, dp, yp);
    double P[5] = {0., 0., 0., 0., 0.}, Q[5] = {0., 0., 0., 0., 0.};
    if constexpr (kSph) {
      // rows f - 1 and f carry |A|_f with the face's area and their own
      // 0.5 / V: P = a_f (y_{f-1} / 2V_{f-1} - y_f / 2V_f)
      const SphLayer lo = sph_layer(a, col, f - 1), hi = sph_layer(a, col, f);
      const double cm = hi.am * (0.5 / lo.vol), cp = hi.am * (0.5 / hi.vol);
#pragma unroll
      for (int e = 0; e < N; ++e) {
        P[vic_row(N, e)] = cm * ym[e] - cp * yp[e];
        Q[vic_row(N, e)] = dm[e] - dp[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < N; ++e) {
        P[vic_row(N, e)] = ym[e] - yp[e];
        Q[vic_row(N, e)] = dm[e] - dp[e];
      }
    }
    int q = col_q0(a, col) + f;  // the cell above the face
    Dual<10> wl[5], wr[5], z[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      wl[v] = Dual<10>::seed(w[v * ncell + q - 1], v);
      wr[v] = Dual<10>::seed(w[v * ncell + q], 5 + v);
    }
    roe_state<Dual<10>>(Dual<10>(gm1), wl, wr, z);
    Dual<5> zs[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) zs[k] = Dual<5>::seed(z[k].v, k);
    Dual<5> phi = abs_contract<Dual<5>>(Dual<5>(gm1), zs, P, Q);
    double* g = G + t * 10;
#pragma unroll
    for (int m = 0; m < 10; ++m) {
      double s = 0.;
#pragma unroll
      for (int k = 0; k < 5; ++k) s += phi.d[k] * z[k].d[m];
      g[m] = s;
    }
  }
}

// Cell j in [-1, nl] of a column: lam_w[j] -= (area / 2V) d/dw_j of
//   phi_j + phi_{j+1} + (y_{j-1} - y_{j+1})^T J(w_j) delta_j
// (the faces below / above and the flux Jacobian of the cell, which enters
// row j+1 as its lower and row j-1 as its upper coupling).
template <int N, typename S, bool kSph>
__global__ void __launch_bounds__(kT)
    vic_adj_cell_kernel(const S* __restrict__ w, const double* __restrict__ dl,
                        const double* __restrict__ G, S* __restrict__ lam_w, AdjArgs a,
                        int ncol) {
  const int NN = N * N;
  const int ncell = a.n1 * a.n2 * a.n3;
  const int nl = a.nx1;
  const double ah = a.dx3 * a.dx2 * (0.5 / (a.dx3 * (a.dx2 * a.dx1)));
  int64_t total = (int64_t)ncol * (nl + 2);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int col = (int)(t / (nl + 2)), j = (int)(t - (int64_t)col * (nl + 2)) - 1;
    const double* cdl = dl + (int64_t)col * nl * 2 * N;
    const double* gc = G + (int64_t)col * (nl + 1) * 10;
    double acc[5] = {0., 0., 0., 0., 0.};
    if (j >= 0)
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += gc[j * 10 + 5 + v];
    if (j + 1 <= nl)
#pragma unroll
      for (int v = 0; v < 5; ++v) acc[v] += gc[(j + 1) * 10 + v];
    double dj[N], yj[N], dlo[N], ylo[N], dhi[N], yhi[N];
    ext_delta_y<N>(cdl, j, nl, dj, yj);
    ext_delta_y<N>(cdl, j - 1, nl, dlo, ylo);
    ext_delta_y<N>(cdl, j + 1, nl, dhi, yhi);
    int q = col_q0(a, col) + j;
    Dual<5> wj[5], J[NN];
#pragma unroll
    for (int v = 0; v < 5; ++v) wj[v] = Dual<5>::seed(w[v * ncell + q], v);
    flux_jacobian<Dual<5>, N>(Dual<5>(a.gammad - 1.), wj, J);
    Dual<5> psi = Dual<5>(0.);
    if constexpr (kSph) {
      // J(w_j) enters row j + 1 (lower coupling, area a_{j+1}), row j - 1
      // (upper coupling, area a_j) and, with a_{j+1} - a_j, row j itself,
      // each with its own 0.5 / V; the face terms of G carry their factors
      const SphLayer lo = sph_layer(a, col, j - 1), me = sph_layer(a, col, j),
                     hi = sph_layer(a, col, j + 1);
      const double clo = me.am * (0.5 / lo.vol), chi = me.ap * (0.5 / hi.vol),
                   cme = (me.ap - me.am) * (0.5 / me.vol);
#pragma unroll
      for (int r = 0; r < N; ++r) {
        const double coef = clo * ylo[r] - chi * yhi[r] + cme * yj[r];
#pragma unroll
        for (int c = 0; c < N; ++c) psi += J[r * N + c] * Dual<5>(coef * dj[c]);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) lam_w[v * ncell + q] -= acc[v] + psi.d[v];
      continue;
    }
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = 0; c < N; ++c) psi += J[r * N + c] * Dual<5>((ylo[r] - yhi[r]) * dj[c]);
#pragma unroll
    for (int v = 0; v < 5; ++v) lam_w[v * ncell + q] -= ah * (acc[v] + psi.d[v]);
  }
}

// ------------------------------------------------------- explicit part

// Slot v of the face frame of axis d holds variable vmap[v]: density,
// normal velocity, the two tangential velocities in cyclic order, pressure.
// Kernels that work in this frame index their register arrays by constants
// only (a runtime slot index would put the arrays in scratch memory) and
// move data in and out of the frame through memory addresses.
__device__ __forceinline__ void face_frame(int d, int* vmap) {
  vmap[0] = IDN;
  vmap[1] = IVX + d;
  vmap[2] = IVX + (d + 1) % 3;
  vmap[3] = IVX + (d + 2) % 3;
  vmap[4] = IPR;
}

// interface states of face q along the axis with stride s (face f along the
// axis) in the face frame; the same arithmetic as recon_kernel for the faces
// whose adjoint is nonzero (f in [ng, n - ng]: never a dummy face)
template <typename T, typename S>
__device__ __forceinline__ void face_states(const S* w, int ncell, int q, int s, int f,
                                            const AdjArgs a, int d, const int* vmap, T* wl,
                                            T* wr) {
  int r = d == 0 ? 0 : 1;
  // selects, not a runtime index into the argument struct (which would
  // place it in scratch memory)
  int type = r ? a.recon_type[1] : a.recon_type[0];
  int scale = r ? a.recon_scale[1] : a.recon_scale[0];
  int shock = r ? a.recon_shock[1] : a.recon_shock[0];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    int tv = (!shock && v >= IVX && v <= IPR) ? interp2_type(type) : type;
    T win[7];
#pragma unroll
    for (int m = -3; m <= 3; ++m) {
      int c = f + m;
      int n = d == 0 ? a.n1 : (d == 1 ? a.n2 : a.n3);
      c = c < 0 ? 0 : (c >= n ? n - 1 : c);
      win[m + 3] = T((double)w[vmap[v] * ncell + q + (c - f) * s]);
    }
    // window centred on cell f (index 3) for the right state and cell f-1
    // (index 2) for the left state
    wr[v] = interp_face<T>(tv, 0, scale, win, 1, 3);
    wl[v] = interp_face<T>(tv, 1, scale, win, 1, 2);
  }
}

__device__ __forceinline__ int axis_len(const AdjArgs a, int d) {
  return d == 0 ? a.n1 : (d == 1 ? a.n2 : a.n3);
}
__device__ __forceinline__ int axis_stride(const AdjArgs a, int d) {
  return d == 0 ? 1 : (d == 1 ? a.n1 : a.n1 * a.n2);
}

// Faces of axis d whose flux reaches an interior cell (a face f in
// [ng, n - ng], never a dummy face), and lambda_F there in the face frame:
// lambda_F = (-dt A / V)(lambda_dx[left cell] - lambda_dx[right cell]).
// Spherical-polar blocks (SphericalPolarImpl::forward,
// src/coord/spherical_polar.cpp:212-279): each cell's own area over its own
// volume, plus the sources that read the fluxes (x1 faces: src2_i A1 F of the
// transverse momenta; x2 faces: src1_i src2_j A2 F of v3) and, for x1 faces,
// lambda of the face pressure in lfp.
template <typename S, bool kSph>
__device__ __forceinline__ bool face_lambda(const AdjArgs a, int d, const S* lam_dx,
                                            int ncell, int q, int s, const int* vmap,
                                            int& f, double* lf, double* lfp = nullptr) {
  const DivMod dn1(a.n1), dn2(a.n2);
  int i, j;
  int k = dn2.div(dn1.div(q, i), j);
  const int n = axis_len(a, d);
  f = d == 0 ? i : (d == 1 ? j : k);
#pragma unroll
  for (int v = 0; v < 5; ++v) lf[v] = 0.;
  if (lfp) *lfp = 0.;
  bool any = false;
  if constexpr (kSph) {
    // F enters the left cell (its upper face) with +Ap / V and the right
    // cell (its lower face) with -Am / V; lambda_d = -dt lambda_dx
    if (f >= 1) {
      int il = i - (d == 0), jl = j - (d == 1), kl = k - (d == 2);
      if (interior(a, il, jl, kl)) {
        const SphCell g = sph_cell(a.geo, kl, jl, il);
        const double ap = d == 0 ? g.a1p : (d == 1 ? g.a2p : g.a3);
        const double c = -a.dt * ap / g.vol;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          const int var = vmap[v];
          const double l = lam_dx[var * ncell + q - s];
          double x = c * l;
          if (d == 0 && (var == IVY || var == IVZ)) x += -a.dt * g.src2_i * g.a1p * l;
          if (d == 1 && var == IVZ) x += -a.dt * g.src1_i * g.src2_j * g.a2p * l;
          lf[v] += x;
        }
        if (d == 0 && lfp)
          *lfp += -a.dt * (double)lam_dx[IVX * ncell + q - s] * (1. / g.dx1 - g.a1p / g.vol);
        any = true;
      }
    }
    if (interior(a, i, j, k)) {
      const SphCell g = sph_cell(a.geo, k, j, i);
      const double am = d == 0 ? g.a1m : (d == 1 ? g.a2m : g.a3);
      const double c = -a.dt * am / g.vol;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const int var = vmap[v];
        const double l = lam_dx[var * ncell + q];
        double x = -c * l;
        if (d == 0 && (var == IVY || var == IVZ)) x += -a.dt * g.src2_i * g.a1m * l;
        if (d == 1 && var == IVZ) x += -a.dt * g.src1_i * g.src2_j * g.a2m * l;
        lf[v] += x;
      }
      if (d == 0 && lfp)
        *lfp += -a.dt * (double)lam_dx[IVX * ncell + q] * (g.a1m / g.vol - 1. / g.dx1);
      any = true;
    }
    return any && f >= a.ng && f <= n - a.ng;
  }
  double area = d == 0 ? a.dx3 * a.dx2 : (d == 1 ? a.dx3 * a.dx1 : a.dx2 * a.dx1);
  double coef = -a.dt * area / (a.dx3 * (a.dx2 * a.dx1));
  if (f >= 1) {
    int il = i - (d == 0), jl = j - (d == 1), kl = k - (d == 2);
    if (interior(a, il, jl, kl)) {
#pragma unroll
      for (int v = 0; v < 5; ++v) lf[v] += coef * lam_dx[vmap[v] * ncell + q - s];
      any = true;
    }
  }
  if (interior(a, i, j, k)) {
#pragma unroll
    for (int v = 0; v < 5; ++v) lf[v] -= coef * lam_dx[vmap[v] * ncell + q];
    any = true;
  }
  return any && f >= a.ng && f <= n - a.ng;
}

// Interface states (face frame, x1 LR revision applied) of the faces that
// carry an adjoint, into buf [2][5][ncell] (slot order; left, then right).
template <typename S, bool kSph>
__global__ void __launch_bounds__(kT)
    face_state_kernel(const S* __restrict__ w, const S* __restrict__ lam_dx,
                      S* __restrict__ buf, AdjArgs a, int d) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int s = axis_stride(a, d), n = axis_len(a, d);
  int vmap[5];
  face_frame(d, vmap);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int f;
    double lf[5];
    if (!face_lambda<S, kSph>(a, d, lam_dx, ncell, q, s, vmap, f, lf)) continue;
    double wl[5], wr[5];
    face_states<double, S>(w, ncell, q, s, f, a, d, vmap, wl, wr);
    if (d == 0 && a.grav1 != 0.) {
      if (f == a.ng) wl[IDN] = wr[IDN], wl[IPR] = wr[IPR];
      if (f == n - a.ng) wr[IDN] = wl[IDN], wr[IPR] = wl[IPR];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      buf[v * ncell + q] = wl[v];
      buf[(5 + v) * ncell + q] = wr[v];
    }
  }
}

// Riemann adjoint (Dual<kFluxDirs> passes) and x1 LR-revision adjoint of
// every face of axis d: lam_wlr [2][5][ncell] (ILT, IRT; variable order)
// replaces the face states face_state_kernel left there (same index).
// One instantiation per solver: each holds only its own solver's registers.
template <int SOLVER, typename S, bool kSph>
__global__ void __launch_bounds__(kT)
    flux_adj_kernel(const S* __restrict__ lam_dx, S* __restrict__ lam_wlr, AdjArgs a, int d) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int s = axis_stride(a, d), n = axis_len(a, d);
  const int solver = SOLVER;
  int vmap[5];
  face_frame(d, vmap);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int f;
    double lf[5], lfp = 0.;
    bool act = face_lambda<S, kSph>(a, d, lam_dx, ncell, q, s, vmap, f, lf,
                                    (kSph && d == 0) ? &lfp : nullptr);
    double lwl[5] = {0., 0., 0., 0., 0.}, lwr[5] = {0., 0., 0., 0., 0.};
    if (act) {
      double wl[5], wr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        wl[v] = lam_wlr[v * ncell + q];
        wr[v] = lam_wlr[(5 + v) * ncell + q];
      }
      // J^T lambda_F: Roe and LMARS by their hand-written reverse sweeps;
      // HLLC by forward-mode duals, kFluxDirs input directions per pass (the
      // solver on wider duals does not fit the register file)
      constexpr bool kReverse = SOLVER != SNAPX_RIEMANN_HLLC;
      if (SOLVER == SNAPX_RIEMANN_ROE) roe_flux_adj(wl, wr, a.gammad, lf, lwl, lwr);
      if (SOLVER == SNAPX_RIEMANN_LMARS) lmars_flux_adj(wl, wr, a.gammad, lf, lwl, lwr);
#pragma unroll 1
      for (int p0 = 0; p0 < (kReverse ? 0 : 10); p0 += kFluxDirs) {
        Dual<kFluxDirs> sl[5], sr[5], flx[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          sl[v] = Dual<kFluxDirs>(wl[v]);
          sr[v] = Dual<kFluxDirs>(wr[v]);
#pragma unroll
          for (int j = 0; j < kFluxDirs; ++j) {
            sl[v].d[j] = p0 + j == v ? 1. : 0.;
            sr[v].d[j] = p0 + j == 5 + v ? 1. : 0.;
          }
        }
        // x1 frame: the solver's velocity slots are constants
        riemann_flux<Dual<kFluxDirs>>(solver, sl, sr, Dual<kFluxDirs>(a.gammad), 3, flx,
                                      (Dual<kFluxDirs>*)nullptr);
#pragma unroll
        for (int j = 0; j < kFluxDirs; ++j) {
          double acc = 0.;
#pragma unroll
          for (int m = 0; m < 5; ++m) acc += lf[m] * flx[m].d[j];
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            if (p0 + j == v) lwl[v] += acc;
            if (p0 + j == 5 + v) lwr[v] += acc;
          }
        }
      }
      // spherical-polar x1 faces: the adjoint of the face pressure the solver
      // returns for the radial source, by Dual<1> per input direction through
      // the same solver code
      if (kSph && d == 0 && lfp != 0.) {
#pragma unroll 1
        for (int p0 = 0; p0 < 10; ++p0) {
          Dual<1> sl[5], sr[5], flx[5], pf;
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            sl[v] = Dual<1>(wl[v]);
            sr[v] = Dual<1>(wr[v]);
            sl[v].d[0] = p0 == v ? 1. : 0.;
            sr[v].d[0] = p0 == 5 + v ? 1. : 0.;
          }
          riemann_flux<Dual<1>>(solver, sl, sr, Dual<1>(a.gammad), 3, flx, &pf);
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            if (p0 == v) lwl[v] += lfp * pf.d[0];
            if (p0 == 5 + v) lwr[v] += lfp * pf.d[0];
          }
        }
      }
      // LR revision adjoint: the copied states take the adjoint of the copies
      if (d == 0 && a.grav1 != 0.) {
        if (f == a.ng) {
          lwr[IDN] += lwl[IDN], lwl[IDN] = 0.;
          lwr[IPR] += lwl[IPR], lwl[IPR] = 0.;
        }
        if (f == n - a.ng) {
          lwl[IDN] += lwr[IDN], lwr[IDN] = 0.;
          lwl[IPR] += lwr[IPR], lwr[IPR] = 0.;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      lam_wlr[vmap[v] * ncell + q] = lwl[v];
      lam_wlr[(5 + vmap[v]) * ncell + q] = lwr[v];
    }
  }
}

// lambda_w[c] += sum over the faces whose interpolation window holds cell c
// (faces f = c - m, m in [-3, 3]).  The 13 cells c-6..c+6 that those windows
// span are loaded once; linear interpolants (DC, CP3, CP5) contribute their
// constant weights, the others are differentiated with Dual<1> seeded at
// cell c.
template <typename S>
__global__ void __launch_bounds__(kT)
    recon_adj_kernel(const S* __restrict__ w, const S* __restrict__ lam_wlr,
                     S* __restrict__ lam_w, AdjArgs a, int d) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int s = axis_stride(a, d), n = axis_len(a, d);
  int r = d == 0 ? 0 : 1;
  // selects, not a runtime index into the argument struct (which would
  // place it in scratch memory)
  int type = r ? a.recon_type[1] : a.recon_type[0];
  int scale = r ? a.recon_scale[1] : a.recon_scale[0];
  int shock = r ? a.recon_shock[1] : a.recon_shock[0];
  const DivMod dn1(a.n1), dn2(a.n2), dnc(ncell);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 5 * ncell;
       t += gridDim.x * blockDim.x) {
    int q;
    int v = dnc.div(t, q);
    int i, j;
    int k = dn2.div(dn1.div(q, i), j);
    int c = d == 0 ? i : (d == 1 ? j : k);
    int tv = (!shock && v >= IVX && v <= IPR) ? interp2_type(type) : type;
    // WENO5 without scaling: weno_face_kernel + weno_gather_kernel
    if (tv == SNAPX_INTERP_WENO5 && !scale) continue;
    bool linear = tv == SNAPX_INTERP_DC || tv == SNAPX_INTERP_CP3 || tv == SNAPX_INTERP_CP5;
    const S* wv = w + (int64_t)v * ncell + q;
    double cells[13];
    if (!linear) {
#pragma unroll
      for (int o = -6; o <= 6; ++o) {
        int cc = c + o;
        cc = cc < 0 ? 0 : (cc >= n ? n - 1 : cc);
        cells[o + 6] = wv[(cc - c) * s];
      }
    }
    // constant weights of a linear interpolant on its 7-cell window
    double clo[7], chi[7];
    if (linear) {
      Dual<7> win[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) win[m] = Dual<7>::seed(0.0, m);
      Dual<7> lo = interp_face<Dual<7>>(tv, 0, 0, win, 1, 3);
      Dual<7> hi = interp_face<Dual<7>>(tv, 1, 0, win, 1, 2);
#pragma unroll
      for (int m = 0; m < 7; ++m) clo[m] = lo.d[m], chi[m] = hi.d[m];
    }
    double acc = 0.;
#pragma unroll
    for (int m = -3; m <= 3; ++m) {
      int f = c - m;  // cell c sits at window position m + 3 of face f
      if (f < a.ng || f > n - a.ng) continue;
      double lr = lam_wlr[(5 + v) * ncell + q - m * s];
      double ll = lam_wlr[v * ncell + q - m * s];
      if (lr == 0. && ll == 0.) continue;
      if (linear) {
        acc += lr * clo[m + 3] + ll * chi[m + 3];
        continue;
      }
      // window of face f: cells f-3..f+3 = cells[3 - m .. 9 - m]; the
      // clamped copies of an edge cell are the same cell
      Dual<1> win[7];
#pragma unroll
      for (int p = 0; p < 7; ++p) {
        int cc = f - 3 + p;
        int ccl = cc < 0 ? 0 : (cc >= n ? n - 1 : cc);
        win[p] = ccl == c ? Dual<1>::seed(cells[3 - m + p], 0) : Dual<1>(cells[3 - m + p]);
      }
      if (lr != 0.) acc += lr * interp_face<Dual<1>>(tv, 0, scale, win, 1, 3).d[0];
      if (ll != 0.) acc += ll * interp_face<Dual<1>>(tv, 1, scale, win, 1, 2).d[0];
    }
    lam_w[t] += acc;
  }
}

// recon_adj_kernel for directions whose interpolants are all linear (DC, CP3,
// CP5) or unscaled WENO5 (served by weno_face_kernel / weno_gather_kernel):
// lambda_w[c] += sum over the faces f = c - m of the interpolant's constant
// weights, which come from evaluating it on unit windows (folded by the
// compiler) instead of on dual numbers per cell.
template <typename S>
__global__ void __launch_bounds__(kT)
    recon_adj_linear_kernel(const S* __restrict__ lam_wlr, S* __restrict__ lam_w, AdjArgs a,
                            int d) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int s = axis_stride(a, d), n = axis_len(a, d);
  const int r = d == 0 ? 0 : 1;
  const int type = r ? a.recon_type[1] : a.recon_type[0];
  const int shock = r ? a.recon_shock[1] : a.recon_shock[0];
  const int type2 = shock ? type : interp2_type(type);
  const DivMod dn1(a.n1), dn2(a.n2), dnc(ncell);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 5 * ncell;
       t += gridDim.x * blockDim.x) {
    int q;
    const int v = dnc.div(t, q);
    const int tv = v == IDN ? type : type2;
    if (tv == SNAPX_INTERP_WENO5) continue;
    int i, j;
    const int k = dn2.div(dn1.div(q, i), j);
    const int c = d == 0 ? i : (d == 1 ? j : k);
    double clo[7], chi[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      double e[7] = {0., 0., 0., 0., 0., 0., 0.};
      e[m] = 1.;
      clo[m] = interp_face<double>(tv, 0, 0, e, 1, 3);
      chi[m] = interp_face<double>(tv, 1, 0, e, 1, 2);
    }
    double acc = 0.;
#pragma unroll
    for (int m = -3; m <= 3; ++m) {
      const int f = c - m;  // cell c sits at window position m + 3 of face f
      if (f < a.ng || f > n - a.ng) continue;
      const double lr = lam_wlr[(5 + v) * ncell + q - m * s];
      const double ll = lam_wlr[v * ncell + q - m * s];
      acc += lr * clo[m + 3] + ll * chi[m + 3];
    }
    lam_w[t] += acc;
  }
}

// Gradient of the WENO5 lower-face value (weno5_lo, WENO5-JS) with respect
// to its window f0..f4, in closed form:
//   v = sum_k w_k p_k,  w_k = a_k / sum a,  a_k = d_k / (b_k + eps)^2,
//   dv/df = sum_k w_k [dp_k/df - 2 (p_k - v) / (b_k + eps) db_k/df],
//   b_k = 13/12 L1_k^2 + 1/4 L2_k^2 with L1_k, L2_k linear in f.
// The upper-face value is the lower-face value of the mirrored window.
__device__ __forceinline__ void weno5_lo_grad(double f0, double f1, double f2, double f3,
                                              double f4, double* g) {
  const double c13 = 13. / 12., eps = 1e-6;
  const double p[3] = {(-1. / 6.) * f0 + (5. / 6.) * f1 + (1. / 3.) * f2,
                       (1. / 3.) * f1 + (5. / 6.) * f2 + (-1. / 6.) * f3,
                       (11. / 6.) * f2 + (-7. / 6.) * f3 + (1. / 3.) * f4};
  const double L1[3] = {f0 - 2. * f1 + f2, f1 - 2. * f2 + f3, f2 - 2. * f3 + f4};
  const double L2[3] = {f0 - 4. * f1 + 3. * f2, -f1 + f3, 3. * f2 - 4. * f3 + f4};
  const double dk[3] = {.3, .6, .1};
  double a[3], be[3], S = 0.;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    be[k] = c13 * L1[k] * L1[k] + .25 * L2[k] * L2[k] + eps;
    a[k] = dk[k] / (be[k] * be[k]);
    S += a[k];
  }
  const double iS = 1. / S;
  const double v = (a[0] * p[0] + a[1] * p[1] + a[2] * p[2]) * iS;
  // d p_k / d f_m, d L1_k / d f_m, d L2_k / d f_m
  const double dp[3][5] = {{-1. / 6., 5. / 6., 1. / 3., 0., 0.},
                           {0., 1. / 3., 5. / 6., -1. / 6., 0.},
                           {0., 0., 11. / 6., -7. / 6., 1. / 3.}};
  const double d1[3][5] = {{1., -2., 1., 0., 0.}, {0., 1., -2., 1., 0.}, {0., 0., 1., -2., 1.}};
  const double d2[3][5] = {{1., -4., 3., 0., 0.}, {0., -1., 0., 1., 0.}, {0., 0., 3., -4., 1.}};
  double wk[3], hk[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    wk[k] = a[k] * iS;
    hk[k] = -2. * wk[k] * (p[k] - v) / be[k];
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    double acc = 0.;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      acc += wk[k] * dp[k][m] + hk[k] * (2. * c13 * L1[k] * d1[k][m] + .5 * L2[k] * d2[k][m]);
    g[m] = acc;
  }
}

// Face f of axis d, variable v (WENO5, no scaling): the contributions of its
// two interface states to the six cells f-3..f+2 of their windows,
//   G6[k][q] = ll d(hi of cell f-1)/dw_{f-3+k} + lr d(lo of cell f)/dw_{f-3+k},
// zero for faces that carry no adjoint.
template <typename S>
__global__ void __launch_bounds__(kT)
    weno_face_kernel(const S* __restrict__ w, const S* __restrict__ lam_wlr,
                     S* __restrict__ G6, AdjArgs a, int d, int v) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int s = axis_stride(a, d), n = axis_len(a, d);
  const DivMod dn1(a.n1), dn2(a.n2);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i, j;
    int k = dn2.div(dn1.div(q, i), j);
    int f = d == 0 ? i : (d == 1 ? j : k);
    double g6[6] = {0., 0., 0., 0., 0., 0.};
    double ll = lam_wlr[v * ncell + q], lr = lam_wlr[(5 + v) * ncell + q];
    if (f >= a.ng && f <= n - a.ng && (ll != 0. || lr != 0.)) {
      const S* wv = w + (int64_t)v * ncell + q;
      double c6[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) c6[m] = wv[(m - 3) * s];  // cells f-3 .. f+2
      double gl[5], gr[5];
      // left state: upper face of cell f-1, window f-3..f+1, mirrored
      weno5_lo_grad(c6[4], c6[3], c6[2], c6[1], c6[0], gl);
      // right state: lower face of cell f, window f-2..f+2
      weno5_lo_grad(c6[1], c6[2], c6[3], c6[4], c6[5], gr);
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        g6[m] += ll * gl[4 - m];
        g6[m + 1] += lr * gr[m];
      }
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) G6[(int64_t)m * ncell + q] = g6[m];
  }
}

// lam_w[v][c] += sum over faces f = c+3-k (k = 0..5) carrying an adjoint of
// G6[k][f]
template <typename S>
__global__ void __launch_bounds__(kT)
    weno_gather_kernel(const S* __restrict__ G6, S* __restrict__ lam_w, AdjArgs a, int d,
                       int v) {
  const int ncell = a.n1 * a.n2 * a.n3;
  const int s = axis_stride(a, d), n = axis_len(a, d);
  const DivMod dn1(a.n1), dn2(a.n2);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i, j;
    int k = dn2.div(dn1.div(q, i), j);
    int c = d == 0 ? i : (d == 1 ? j : k);
    double acc = 0.;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      int f = c + 3 - m;
      if (f >= a.ng && f <= n - a.ng) acc += G6[(int64_t)m * ncell + q + (3 - m) * s];
    }
    lam_w[(int64_t)v * ncell + q] += acc;
  }
}

// constant gravity (src/forcing/const_gravity.cpp:26-45) adjoint
template <typename S>
__global__ void grav_adj_kernel(const S* __restrict__ w, const S* __restrict__ lam_dx,
                                S* __restrict__ lam_w, AdjArgs a) {
  const int ncell = a.n1 * a.n2 * a.n3;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    double rho = w[IDN * ncell + q];
    double l1 = lam_dx[IVX * ncell + q], l2 = lam_dx[IVY * ncell + q];
    double l3 = lam_dx[IVZ * ncell + q], le = lam_dx[IPR * ncell + q];
    double lr = 0.;
    if (a.grav1 != 0.) {
      lr += a.dt * a.grav1 * a.nh * (l1 + w[IVX * ncell + q] * le);
      lam_w[IVX * ncell + q] += a.dt * rho * a.grav1 * a.nh * le;
    }
    if (a.grav2 != 0.) {
      lr += a.dt * a.grav2 * (l2 + w[IVY * ncell + q] * le);
      lam_w[IVY * ncell + q] += a.dt * rho * a.grav2 * le;
    }
    if (a.grav3 != 0.) {
      lr += a.dt * a.grav3 * (l3 + w[IVZ * ncell + q] * le);
      lam_w[IVZ * ncell + q] += a.dt * rho * a.grav3 * le;
    }
    lam_w[IDN * ncell + q] += lr;
  }
}

// the cell sources of spherical-polar blocks that read the cell's own
// primitives (spherical_polar.cpp:212-279: -src1_i rho (v2^2 + v3^2) on v1,
// -src1_i src1_j (rho v3^2 + p) on v2 and, without an x2 axis,
// src1_i src1_j rho v3 v2 on v3), by Dual<5>; lambda_d = -dt lambda_dx of
// the interior cells
template <typename S>
__global__ void sph_src_adj_kernel(const S* __restrict__ w, const S* __restrict__ lam_dx,
                                   S* __restrict__ lam_w, AdjArgs a) {
  const int ncell = a.n1 * a.n2 * a.n3;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i, j, k;
    decode(a, q, i, j, k);
    if (!interior(a, i, j, k)) continue;
    const SphCell g = sph_cell(a.geo, k, j, i);
    Dual<5> W[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) W[v] = Dual<5>::seed((double)w[v * ncell + q], v);
    const Dual<5> s1i(g.src1_i), s1ij(g.src1_i * g.src1_j);
    const Dual<5> sx = Dual<5>(0.) - s1i * W[IDN] * (W[IVY] * W[IVY] + W[IVZ] * W[IVZ]);
    const Dual<5> sy = Dual<5>(0.) - s1ij * (W[IDN] * (W[IVZ] * W[IVZ]) + W[IPR]);
    const Dual<5> sz = a.n2 > 1 ? Dual<5>(0.) : s1ij * (W[IDN] * W[IVZ] * W[IVY]);
    const double lx = -a.dt * (double)lam_dx[IVX * ncell + q];
    const double ly = -a.dt * (double)lam_dx[IVY * ncell + q];
    const double lz = -a.dt * (double)lam_dx[IVZ * ncell + q];
#pragma unroll
    for (int m = 0; m < 5; ++m)
      lam_w[m * ncell + q] += lx * sx.d[m] + ly * sy.d[m] + lz * sz.d[m];
  }
}

// isentropic ghost refill (src/hydro/hydro.cpp:231-269) adjoint, one thread
// per column; wpre holds the primitives before the refill (only rho, p of
// the boundary cells are read)
template <typename S, bool kSph>
__global__ void ghost_adj_kernel(const S* __restrict__ w, S* __restrict__ lam_w, AdjArgs a) {
  const int ncell = a.n1 * a.n2 * a.n3;
  int ncol = a.n2 * a.n3;
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < ncol;
       col += gridDim.x * blockDim.x) {
    for (int side = 0; side < 2; ++side) {
      int is = side ? a.n1 - a.ng - 1 : a.ng;
      int base = col * a.n1;
      Dual<2> rho_b = Dual<2>::seed(w[IDN * ncell + base + is], 0);
      Dual<2> pre_b = Dual<2>::seed(w[IPR * ncell + base + is], 1);
      Dual<2> gamma = Dual<2>(a.gammad), gm = gamma - Dual<2>(1.), ae = gm / gamma;
      Dual<2> inv_g = Dual<2>(1.) / gamma, inv_a = Dual<2>(1.) / ae;
      Dual<2> grav = Dual<2>(-a.grav1);
      Dual<2> K = pre_b / pow(rho_b, gamma);
      Dual<2> prev = pre_b;
      double lr = 0., lp = 0.;
      for (int n = 0; n < a.ng; ++n) {
        // the centre distance of the ghost layer (dx1v of the coordinate)
        const Dual<2> dz = Dual<2>(!kSph ? a.dx1 : sph_dx1v(a.geo, side ? is + n : is - n - 1));
        Dual<2> h = side ? pow(prev, ae) - ae * grav * dz / pow(K, inv_g)
                         : pow(prev, ae) + ae * grav * dz / pow(K, inv_g);
        Dual<2> pg = pow(h, inv_a);
        Dual<2> rg = pow(pg / K, inv_g);
        int g = side ? is + n + 1 : is - n - 1;
        double lrg = lam_w[IDN * ncell + base + g], lpg = lam_w[IPR * ncell + base + g];
        lr += lrg * rg.d[0] + lpg * pg.d[0];
        lp += lrg * rg.d[1] + lpg * pg.d[1];
        lam_w[IDN * ncell + base + g] = 0.;
        lam_w[IPR * ncell + base + g] = 0.;
        prev = pg;
      }
      lam_w[IDN * ncell + base + is] += lr;
      lam_w[IPR * ncell + base + is] += lp;
    }
  }
}

// cons2prim (src/eos/ideal_gas_impl.h:15-35) adjoint: lam_u += J^T lam_w
template <typename S>
__global__ void c2p_adj_kernel(const S* __restrict__ u, const S* __restrict__ lam_w,
                               S* __restrict__ lam_u, AdjArgs a) {
  const int ncell = a.n1 * a.n2 * a.n3;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    Dual<5> U[5], W[5];
    for (int v = 0; v < 5; ++v) U[v] = Dual<5>::seed(u[v * ncell + q], v);
    W[IDN] = U[IDN];
    W[IVX] = U[IVX] / W[IDN];
    W[IVY] = U[IVY] / W[IDN];
    W[IVZ] = U[IVZ] / W[IDN];
    Dual<5> ke = Dual<5>(0.5) * ((W[IVX] * U[IVX] + W[IVY] * U[IVY]) + W[IVZ] * U[IVZ]);
    W[IPR] = (Dual<5>(a.gammad) - Dual<5>(1.)) * (U[IPR] - ke);
    for (int v = 0; v < 5; ++v) {
      double acc = 0.;
      for (int m = 0; m < 5; ++m) acc += lam_w[m * ncell + q] * W[m].d[v];
      lam_u[v * ncell + q] += acc;
    }
  }
}

template <typename S>
__global__ void zero_kernel(S* __restrict__ x, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    x[q] = S(0);
}

template <typename S>
__global__ void axpy_kernel(S* __restrict__ y, const S* __restrict__ x, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    y[q] += x[q];
}

static AdjArgs adj_args(const snapx_config* c, double dt) {
  Grid g = make_grid(c);
  AdjArgs a;
  a.n1 = g.n1, a.n2 = g.n2, a.n3 = g.n3, a.ng = g.ng;
  a.nx1 = c->nx1, a.nx2 = c->nx2, a.nx3 = c->nx3;
  a.dx1 = g.dx1, a.dx2 = g.dx2, a.dx3 = g.dx3;
  a.dt = dt, a.gammad = c->gammad;
  a.grav1 = c->grav1, a.grav2 = c->grav2, a.grav3 = c->grav3, a.nh = c->non_hydrostatic;
  for (int r = 0; r < 2; ++r) {
    a.recon_type[r] = c->recon_type[r];
    a.recon_scale[r] = c->recon_scale[r];
    a.recon_shock[r] = c->recon_shock[r];
  }
  for (int d = 0; d < 3; ++d) a.riemann[d] = c->riemann[d];
  for (int b = 0; b < 6; ++b) a.bc[b] = c->bc[b];
  a.scheme = c->implicit_scheme;
  a.lanes = 64, a.lshift = 6;
  while (a.lanes < c->nx1) a.lanes <<= 1, a.lshift += 1;
  const bool n3 = !((c->implicit_scheme >> 3) & 1);
  a.half = (n3 && c->nx1 > 32 && a.lanes <= PcrShape<3>::kThreads) ? a.lanes / 2 : 0;
  a.hshift = a.half ? a.lshift - 1 : 0;
  a.sph = c->coord == SNAPX_COORD_SPHERICAL_POLAR;
  a.geo = make_geometry(c);
  return a;
}

static size_t al(size_t x) { return (x + 255) & ~size_t(255); }

template <typename S>
constexpr int dtype_of() {
  return sizeof(S) == sizeof(double) ? SNAPX_F64 : SNAPX_F32;
}

// What the reverse sweep of one step reads from its forward recompute: per
// stage the input state, its primitives (revised ghosts), the explicit
// tendency (storage type S), and the VIC blocks and (delta, rows) of the
// column solve (FP64 for either storage type).
template <typename S>
struct AdjTape {
  S *us[3], *w[3], *dux[3];
  double *abc[3], *dl[3];
  size_t bytes;
};

template <typename S>
static AdjTape<S> tape_layout(const snapx_config* c, char* base) {
  Grid g = make_grid(c);
  size_t f = al(sizeof(S) * 5 * (size_t)g.ncell);
  int N = ((c->implicit_scheme >> 3) & 1) ? 5 : 3;
  size_t ncol = (size_t)(g.n2 > 1 ? c->nx2 : 1) * (g.n3 > 1 ? c->nx3 : 1);
  size_t cells = ncol * c->nx1;
  AdjTape<S> t;
  size_t off = 0;
  auto take = [&](size_t b) {
    char* p = base + off;
    off += al(b);
    return p;
  };
  for (int k = 0; k < 3; ++k) {
    t.us[k] = reinterpret_cast<S*>(take(f));
    t.w[k] = reinterpret_cast<S*>(take(f));
    t.dux[k] = reinterpret_cast<S*>(take(f));
  }
  for (int k = 0; k < 3; ++k) {
    t.abc[k] = reinterpret_cast<double*>(take(sizeof(double) * cells * 3 * N * N));
    t.dl[k] = reinterpret_cast<double*>(take(sizeof(double) * cells * 2 * N));
  }
  t.bytes = off;
  return t;
}

template <typename S>
struct AdjSpace {
  S *us[3], *w[3], *dux[3], *lam, *lam_us, *lam_u0, *lam_du, *lam_dx, *lam_w, *lam_wlr, *G6;
  double *abc[3], *pr, *dl[3], *G;
  char* hydro_ws;
  size_t hydro_bytes, bytes;
};

// The work space of the recompute and the sweep (everything but the tape);
// the tape fields of the result point into `tape`.
template <typename S>
static AdjSpace<S> adj_layout(const snapx_config* c, char* base, char* tape) {
  Grid g = make_grid(c);
  size_t f = al(sizeof(S) * 5 * (size_t)g.ncell);
  int N = ((c->implicit_scheme >> 3) & 1) ? 5 : 3;
  size_t ncol = (size_t)(g.n2 > 1 ? c->nx2 : 1) * (g.n3 > 1 ? c->nx3 : 1);
  size_t cells = ncol * c->nx1;
  AdjSpace<S> s;
  AdjTape<S> t = tape_layout<S>(c, tape);
  for (int k = 0; k < 3; ++k) {
    s.us[k] = t.us[k], s.w[k] = t.w[k], s.dux[k] = t.dux[k];
    s.abc[k] = t.abc[k], s.dl[k] = t.dl[k];
  }
  size_t off = 0;
  auto take = [&](size_t b) {
    char* p = base + off;
    off += al(b);
    return p;
  };
  auto field = [&](size_t b) { return reinterpret_cast<S*>(take(b)); };
  s.lam = field(f);
  s.lam_us = field(f);
  s.lam_u0 = field(f);
  s.lam_du = field(f);
  s.lam_dx = field(f);
  s.lam_w = field(f);
  s.lam_wlr = field(2 * f);
  s.pr = reinterpret_cast<double*>(take(sizeof(double) * cells * (N * N + N)));
  s.G = reinterpret_cast<double*>(take(sizeof(double) * ncol * (c->nx1 + 1) * 10));
  s.G6 = field(6 * f / 5);
  s.hydro_bytes = snapx_hydro_workspace_bytes(c, dtype_of<S>());
  s.hydro_ws = take(s.hydro_bytes);
  s.bytes = off;
  return s;
}

template <typename S>
static int bc_adj(const snapx_config* c, const AdjArgs a, S* lam, hipStream_t st) {
  int dims[3] = {3, 2, 1};
  int ns[3] = {a.n1, a.n2, a.n3};
  for (int ax = 2; ax >= 0; --ax) {  // reverse of the forward x1, x2, x3 order
    if (ns[ax] == 1) continue;
    int tin = c->bc[2 * ax], tout = c->bc[2 * ax + 1];
    if (tin < 0 && tout < 0) continue;
    int np = a.n1 * a.n2 * a.n3 / ns[ax];
    hipLaunchKernelGGL(bc_adj_kernel<S>, blocks_for(5 * np, kT), dim3(kT), 0, st, lam, a,
                       dims[ax], tin, tout, 1);
    SNAPX_LAUNCH_CHECK();
  }
  return 0;
}

static bool adj_pcr(const AdjArgs a, int N) {
  return a.lanes <= (N == 3 ? PcrShape<3>::kThreads : PcrShape<5>::kThreads);
}

template <int N, typename S, bool kSph>
static int launch_vic_fwd(const AdjArgs& a, const AdjSpace<S>& s, int k, S* du, int ncol,
                          hipStream_t st) {
  const int T = PcrShape<N>::kThreads;
  hipLaunchKernelGGL((vic_asm_kernel<N, S, kSph>), blocks_for((int64_t)ncol * a.nx1, kT), dim3(kT), 0,
                     st, s.w[k], s.abc[k], a, ncol);
  SNAPX_LAUNCH_CHECK();
  if (N == 3 && a.half) {
    const int T3 = PcrShape<3>::kThreads, cpw = T3 >> a.hshift;
    hipLaunchKernelGGL((vic_fwd_solve2_kernel<S, kSph>), dim3((ncol + cpw - 1) / cpw), dim3(T3),
                       sizeof(double) * T3 * PcrShape<3>::kPub, st, s.w[k], s.dux[k], s.abc[k], du,
                       s.dl[k], a, ncol);
    SNAPX_LAUNCH_CHECK();
    return 0;
  }
  int cpw = T >> a.lshift;
  hipLaunchKernelGGL((vic_fwd_solve_kernel<N, S, kSph>), dim3((ncol + cpw - 1) / cpw), dim3(T),
                     sizeof(double) * T * PcrShape<N>::kPub, st, s.w[k], s.dux[k], s.abc[k], du,
                     s.dl[k], a, ncol);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

template <int N, typename S, bool kSph>
static int launch_vic_adj(const AdjArgs& a, const AdjSpace<S>& s, int k, int ncol,
                          hipStream_t st) {
  const int T = PcrShape<N>::kThreads;
  if (N == 3 && a.half) {
    // two rows per lane; blocks and delta come from the forward recompute
    const int T3 = PcrShape<3>::kThreads, cpw = T3 >> a.hshift;
    hipLaunchKernelGGL((vic_adj_solve2_kernel<S, kSph>), dim3((ncol + cpw - 1) / cpw), dim3(T3),
                       sizeof(double) * T3 * PcrShape<3>::kPub, st, s.w[k], s.dux[k], s.lam_du,
                       s.abc[k], s.lam_dx, s.lam_w, s.dl[k], a, ncol);
  } else if (adj_pcr(a, N)) {
    // blocks and delta of this stage come from the forward recompute
    int cpw = T >> a.lshift;
    hipLaunchKernelGGL((vic_adj_solve_kernel<N, S, kSph>), dim3((ncol + cpw - 1) / cpw), dim3(T),
                       sizeof(double) * T * PcrShape<N>::kPub, st, s.w[k], s.dux[k], s.lam_du,
                       s.abc[k], s.lam_dx, s.lam_w, s.dl[k], a, ncol);
  } else {
    hipLaunchKernelGGL((vic_adj_column_kernel<N, S>), blocks_for(ncol, 64), dim3(64), 0, st,
                       s.w[k], s.dux[k], s.lam_du, s.lam_dx, s.lam_w, s.abc[k], s.pr, s.dl[k], a,
                       ncol);
  }
  SNAPX_LAUNCH_CHECK();
  hipLaunchKernelGGL((vic_adj_face_kernel<N, S, kSph>), blocks_for((int64_t)ncol * (a.nx1 + 1), kT),
                     dim3(kT), 0, st, s.w[k], s.dl[k], s.G, a, ncol);
  SNAPX_LAUNCH_CHECK();
  hipLaunchKernelGGL((vic_adj_cell_kernel<N, S, kSph>), blocks_for((int64_t)ncol * (a.nx1 + 2), kT),
                     dim3(kT), 0, st, s.w[k], s.dl[k], s.G, s.lam_w, a, ncol);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

static int adjoint_check(const snapx_config* cfg) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(!cfg->limiter, "adjoint: the EOS limiter is not differentiable (turn it off)");
  SNAPX_CHECK_ARG(cfg->non_hydrostatic == 1.0, "adjoint: non-hydrostatic < 1 is not supported");
  SNAPX_CHECK_ARG(cfg->coord == SNAPX_COORD_CARTESIAN || cfg->coord == SNAPX_COORD_SPHERICAL_POLAR,
                  "adjoint: Cartesian and spherical-polar geometry only");
  if (cfg->coord == SNAPX_COORD_SPHERICAL_POLAR && cfg->implicit_scheme) {
    const int N = ((cfg->implicit_scheme >> 3) & 1) ? 5 : 3;
    SNAPX_CHECK_ARG(adj_pcr(adj_args(cfg, 1.0), N),
                    "adjoint: spherical-polar columns of at most 256 (3x3) / 128 (5x5) cells");
  }
  for (int b = 0; b < 6; ++b)
    SNAPX_CHECK_ARG(cfg->bc[b] != SNAPX_BC_CUSTOM, "adjoint: custom boundaries are not supported");
  for (int d = 0; d < 3; ++d)
    SNAPX_CHECK_ARG(!cfg->disable_flux[d], "adjoint: disable-flux is not supported");
  return 0;
}

static const double kStageW[3][3] = {
    {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};

// Forward recompute of one step from u_in into the tape: stage inputs, their
// primitives (revised ghosts), explicit tendencies, VIC blocks and delta.
// With u_out, also the step's output (third stage combined, boundaries).
// The explicit part runs in the storage type: FP64 on the plain path, FP32
// through the fused path's K1 (the arithmetic of the fused forward that the
// FP32 tapes replay) where the fused path takes the block, else on the plain
// path; the column solve is this file's FP64 one for either type.
template <typename S, bool kSph>
static int adjoint_record(const snapx_config* cfg, double dt, const S* u_in, S* u_out,
                          const AdjSpace<S>& s, hipStream_t st) {
  Grid g = make_grid(cfg);
  AdjArgs a = adj_args(cfg, dt);
  int64_t n = 5 * (int64_t)g.ncell;
  size_t fb = sizeof(S) * n;
  const int dtype = dtype_of<S>();
  snapx_config c0 = *cfg;
  c0.implicit_scheme = 0;
  const auto& W = kStageW;
  const int N = ((cfg->implicit_scheme >> 3) & 1) ? 5 : 3;
  const int ncol = (g.n2 > 1 ? cfg->nx2 : 1) * (g.n3 > 1 ? cfg->nx3 : 1);
  const bool fused_rhs = sizeof(S) == sizeof(float) && fused_rhs_ok(cfg);
  TraceRange range("snapx adjoint record");
  SNAPX_HIP(hipMemcpyAsync(s.us[0], u_in, fb, hipMemcpyDeviceToDevice, st));
  for (int k = 0; k < 3; ++k) {
    if (fused_rhs) {
      // primitives with the revised x1 ghosts, as the plain forward leaves
      // them, and the tendency of the interior from K1; the ghost entries
      // (the plain path's body forces of the ghost cells, which the
      // boundary functions overwrite) are zero
      if (int e = snapx_ideal_gas_cons2prim(cfg, s.us[k], s.w[k], dtype, st)) return e;
      if (g.n1 > 1 && cfg->grav1 != 0.)
        if (int e = snapx_revise_x1_ghost(cfg, s.w[k], dtype, st)) return e;
      hipLaunchKernelGGL(zero_kernel<S>, blocks_for(n, kT), dim3(kT), 0, st, s.dux[k], n);
      SNAPX_LAUNCH_CHECK();
      if (int e = fused_explicit_tendency(cfg, dt, reinterpret_cast<const float*>(s.us[k]),
                                          reinterpret_cast<float*>(s.dux[k]), st))
        return e;
    } else if (int e = snapx_hydro_forward(&c0, dt, s.us[k], s.w[k], s.dux[k], nullptr, nullptr,
                                           nullptr, s.hydro_ws, s.hydro_bytes, dtype, st)) {
      return e;
    }
    // du of the stage: the column solve on a copy of the explicit tendency
    // (without u_out, the third stage's solve matters only for the blocks and
    // delta it leaves for the reverse sweep)
    S* du = s.lam_du;
    SNAPX_HIP(hipMemcpyAsync(du, s.dux[k], fb, hipMemcpyDeviceToDevice, st));
    if (cfg->implicit_scheme) {
      if (adj_pcr(a, N)) {
        if (int e = N == 3 ? launch_vic_fwd<3, S, kSph>(a, s, k, du, ncol, st)
                           : launch_vic_fwd<5, S, kSph>(a, s, k, du, ncol, st))
          return e;
      } else if (k < 2 || u_out) {
        if (int e = snapx_hydro_forward(cfg, dt, s.us[k], s.lam_w, du, nullptr, nullptr,
                                        nullptr, s.hydro_ws, s.hydro_bytes, dtype, st))
          return e;
      }
    }
    S* next = k < 2 ? s.us[k + 1] : u_out;
    if (!next) break;
    if (int e = snapx_rk_combine(next, s.us[0], s.us[k], du, W[k][0], W[k][1], W[k][2], n,
                                 dtype, st))
      return e;
    if (int e = snapx_apply_bc(cfg, next, 5, SNAPX_CONSERVED, dtype, st)) return e;
  }
  return 0;
}

// Reverse sweep of one step over its tape: lambda (dL/du_out) becomes
// dL/du_in.
template <typename S, bool kSph>
static int adjoint_sweep(const snapx_config* cfg, double dt, S* lambda, const AdjSpace<S>& s,
                         hipStream_t st) {
  Grid g = make_grid(cfg);
  AdjArgs a = adj_args(cfg, dt);
  int64_t n = 5 * (int64_t)g.ncell;
  size_t fb = sizeof(S) * n;
  const auto& W = kStageW;
  dim3 bn = blocks_for(n, kT), bc = blocks_for(g.ncell, kT);
  const int N = ((cfg->implicit_scheme >> 3) & 1) ? 5 : 3;
  const int ncol = (g.n2 > 1 ? cfg->nx2 : 1) * (g.n3 > 1 ? cfg->nx3 : 1);
  const int ns[3] = {g.n1, g.n2, g.n3};
  TraceRange range("snapx adjoint sweep");
  SNAPX_HIP(hipMemcpyAsync(s.lam, lambda, fb, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(zero_kernel<S>, bn, dim3(kT), 0, st, s.lam_u0, n);
  SNAPX_LAUNCH_CHECK();
  for (int stg = 2; stg >= 0; --stg) {
    if (int e = bc_adj<S>(cfg, a, s.lam, st)) return e;
    hipLaunchKernelGGL(combine_adj_kernel<S>, bn, dim3(kT), 0, st, s.lam, s.lam_du, s.lam_us,
                       s.lam_u0, W[stg][0], W[stg][1], W[stg][2], n);
    SNAPX_LAUNCH_CHECK();
    hipLaunchKernelGGL(zero_kernel<S>, bn, dim3(kT), 0, st, s.lam_w, n);
    hipLaunchKernelGGL(zero_kernel<S>, bn, dim3(kT), 0, st, s.lam_dx, n);
    SNAPX_LAUNCH_CHECK();
    // the column solve rewrites the interior of du only; ghost entries (the
    // gravity forcing) pass through
    hipLaunchKernelGGL(axpy_kernel<S>, bn, dim3(kT), 0, st, s.lam_dx, s.lam_du, n);
    SNAPX_LAUNCH_CHECK();
    if (cfg->implicit_scheme)
      if (int e = N == 3 ? launch_vic_adj<3, S, kSph>(a, s, stg, ncol, st)
                         : launch_vic_adj<5, S, kSph>(a, s, stg, ncol, st))
        return e;
    const S* w = s.w[stg];
    hipLaunchKernelGGL(grav_adj_kernel<S>, bc, dim3(kT), 0, st, w, s.lam_dx, s.lam_w, a);
    SNAPX_LAUNCH_CHECK();
    if (kSph) {
      hipLaunchKernelGGL(sph_src_adj_kernel<S>, bc, dim3(kT), 0, st, w, s.lam_dx, s.lam_w, a);
      SNAPX_LAUNCH_CHECK();
    }
    for (int d = 0; d < 3; ++d) {
      if (ns[d] == 1) continue;
      hipLaunchKernelGGL((face_state_kernel<S, kSph>), bc, dim3(kT), 0, st, w, s.lam_dx, s.lam_wlr, a,
                         d);
      SNAPX_LAUNCH_CHECK();
      int solver = cfg->riemann[d];
      if (solver == SNAPX_RIEMANN_LMARS)
        hipLaunchKernelGGL((flux_adj_kernel<SNAPX_RIEMANN_LMARS, S, kSph>), bc, dim3(kT), 0, st,
                           s.lam_dx, s.lam_wlr, a, d);
      else if (solver == SNAPX_RIEMANN_HLLC)
        hipLaunchKernelGGL((flux_adj_kernel<SNAPX_RIEMANN_HLLC, S, kSph>), bc, dim3(kT), 0, st,
                           s.lam_dx, s.lam_wlr, a, d);
      else
        hipLaunchKernelGGL((flux_adj_kernel<SNAPX_RIEMANN_ROE, S, kSph>), bc, dim3(kT), 0, st,
                           s.lam_dx, s.lam_wlr, a, d);
      SNAPX_LAUNCH_CHECK();
      int r = d == 0 ? 0 : 1;
      {
        // every interpolant of this direction linear or unscaled WENO5?
        const int ty = cfg->recon_type[r];
        const int ty2 = cfg->recon_shock[r] ? ty : interp2_type(ty);
        auto lean = [&](int tv) {
          return tv == SNAPX_INTERP_DC || tv == SNAPX_INTERP_CP3 || tv == SNAPX_INTERP_CP5 ||
                 (tv == SNAPX_INTERP_WENO5 && !cfg->recon_scale[r]);
        };
        if (lean(ty) && lean(ty2))
          hipLaunchKernelGGL(recon_adj_linear_kernel<S>, bn, dim3(kT), 0, st, s.lam_wlr,
                             s.lam_w, a, d);
        else
          hipLaunchKernelGGL(recon_adj_kernel<S>, bn, dim3(kT), 0, st, w, s.lam_wlr, s.lam_w,
                             a, d);
        SNAPX_LAUNCH_CHECK();
      }
      if (cfg->recon_type[r] == SNAPX_INTERP_WENO5 && !cfg->recon_scale[r]) {
        for (int v = 0; v < 5; ++v) {
          bool weno = cfg->recon_shock[r] || v == IDN;
          if (!weno) continue;
          hipLaunchKernelGGL(weno_face_kernel<S>, bc, dim3(kT), 0, st, w, s.lam_wlr, s.G6, a,
                             d, v);
          SNAPX_LAUNCH_CHECK();
          hipLaunchKernelGGL(weno_gather_kernel<S>, bc, dim3(kT), 0, st, s.G6, s.lam_w, a, d,
                             v);
          SNAPX_LAUNCH_CHECK();
        }
      }
    }
    if (g.n1 > 1 && cfg->grav1 != 0.) {
      hipLaunchKernelGGL((ghost_adj_kernel<S, kSph>), blocks_for(g.n2 * g.n3, kT), dim3(kT), 0, st, w,
                         s.lam_w, a);
      SNAPX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(c2p_adj_kernel<S>, bc, dim3(kT), 0, st, s.us[stg], s.lam_w, s.lam_us, a);
    SNAPX_LAUNCH_CHECK();
    // lam_us is now dL/du_stg; the next iteration folds its ghosts (u_stg,
    // stg > 0, left a boundary pass); u_0 = u_in keeps its ghost adjoints
    SNAPX_HIP(hipMemcpyAsync(s.lam, s.lam_us, fb, hipMemcpyDeviceToDevice, st));
  }
  hipLaunchKernelGGL(axpy_kernel<S>, bn, dim3(kT), 0, st, s.lam, s.lam_u0, n);
  SNAPX_LAUNCH_CHECK();
  SNAPX_HIP(hipMemcpyAsync(lambda, s.lam, fb, hipMemcpyDeviceToDevice, st));
  return 0;
}

template <typename S>
static size_t tape_bytes(const snapx_config* cfg) {
  if (adjoint_check(cfg)) return 0;
  return tape_layout<S>(cfg, nullptr).bytes;
}

template <typename S>
static size_t work_bytes(const snapx_config* cfg) {
  if (adjoint_check(cfg)) return 0;
  return adj_layout<S>(cfg, nullptr, nullptr).bytes;
}

// one step's work space followed by its tape
template <typename S>
static size_t workspace_bytes(const snapx_config* cfg) {
  const size_t w = work_bytes<S>(cfg);
  return w ? w + tape_bytes<S>(cfg) : 0;
}

template <typename S>
static int rk3_step(const snapx_config* cfg, double dt, const S* u_in, S* lambda,
                    void* workspace, size_t workspace_bytes_, void* stream, const char* who) {
  if (int e = adjoint_check(cfg)) return e;
  const std::string w(who);
  SNAPX_CHECK_ARG(u_in && lambda && workspace, w + ": null pointer");
  SNAPX_CHECK_ARG(workspace_bytes_ >= workspace_bytes<S>(cfg), w + ": workspace too small");
  char* base = reinterpret_cast<char*>(workspace);
  AdjSpace<S> s = adj_layout<S>(cfg, base, base + work_bytes<S>(cfg));
  hipStream_t st = as_stream(stream);
  if (cfg->coord == SNAPX_COORD_SPHERICAL_POLAR) {
    if (int e = adjoint_record<S, true>(cfg, dt, u_in, nullptr, s, st)) return e;
    return adjoint_sweep<S, true>(cfg, dt, lambda, s, st);
  }
  if (int e = adjoint_record<S, false>(cfg, dt, u_in, nullptr, s, st)) return e;
  return adjoint_sweep<S, false>(cfg, dt, lambda, s, st);
}

template <typename S>
static int record(const snapx_config* cfg, double dt, const S* u_in, S* u_out, void* tape,
                  size_t tape_bytes_, void* work, size_t work_bytes_, void* stream,
                  const char* who) {
  if (int e = adjoint_check(cfg)) return e;
  const std::string w(who);
  SNAPX_CHECK_ARG(u_in && tape && work, w + ": null pointer");
  SNAPX_CHECK_ARG(u_out != u_in, w + ": u_out must not alias u_in");
  SNAPX_CHECK_ARG(tape_bytes_ >= tape_bytes<S>(cfg), w + ": tape too small");
  SNAPX_CHECK_ARG(work_bytes_ >= work_bytes<S>(cfg), w + ": work space too small");
  AdjSpace<S> s =
      adj_layout<S>(cfg, reinterpret_cast<char*>(work), reinterpret_cast<char*>(tape));
  return cfg->coord == SNAPX_COORD_SPHERICAL_POLAR
             ? adjoint_record<S, true>(cfg, dt, u_in, u_out, s, as_stream(stream))
             : adjoint_record<S, false>(cfg, dt, u_in, u_out, s, as_stream(stream));
}

template <typename S>
static int sweep(const snapx_config* cfg, double dt, void* tape, size_t tape_bytes_, S* lambda,
                 void* work, size_t work_bytes_, void* stream, const char* who) {
  if (int e = adjoint_check(cfg)) return e;
  const std::string w(who);
  SNAPX_CHECK_ARG(lambda && tape && work, w + ": null pointer");
  SNAPX_CHECK_ARG(tape_bytes_ >= tape_bytes<S>(cfg), w + ": tape too small");
  SNAPX_CHECK_ARG(work_bytes_ >= work_bytes<S>(cfg), w + ": work space too small");
  AdjSpace<S> s =
      adj_layout<S>(cfg, reinterpret_cast<char*>(work), reinterpret_cast<char*>(tape));
  return cfg->coord == SNAPX_COORD_SPHERICAL_POLAR
             ? adjoint_sweep<S, true>(cfg, dt, lambda, s, as_stream(stream))
             : adjoint_sweep<S, false>(cfg, dt, lambda, s, as_stream(stream));
}

}  // namespace snapx

using namespace snapx;

extern "C" size_t snapx_adjoint_tape_bytes(const snapx_config* cfg) {
  return tape_bytes<double>(cfg);
}

extern "C" size_t snapx_adjoint_work_bytes(const snapx_config* cfg) {
  return work_bytes<double>(cfg);
}

extern "C" size_t snapx_adjoint_workspace_bytes(const snapx_config* cfg) {
  return workspace_bytes<double>(cfg);
}

extern "C" int snapx_adjoint_rk3_step(const snapx_config* cfg, double dt, const double* u_in,
                                      double* lambda, void* workspace, size_t workspace_bytes,
                                      void* stream) {
  return rk3_step<double>(cfg, dt, u_in, lambda, workspace, workspace_bytes, stream,
                          "snapx_adjoint_rk3_step");
}

extern "C" int snapx_adjoint_record(const snapx_config* cfg, double dt, const double* u_in,
                                    double* u_out, void* tape, size_t tape_bytes, void* work,
                                    size_t work_bytes, void* stream) {
  return record<double>(cfg, dt, u_in, u_out, tape, tape_bytes, work, work_bytes, stream,
                        "snapx_adjoint_record");
}

extern "C" int snapx_adjoint_sweep(const snapx_config* cfg, double dt, void* tape,
                                   size_t tape_bytes, double* lambda, void* work,
                                   size_t work_bytes, void* stream) {
  return sweep<double>(cfg, dt, tape, tape_bytes, lambda, work, work_bytes, stream,
                       "snapx_adjoint_sweep");
}

extern "C" size_t snapx_adjoint_f32_tape_bytes(const snapx_config* cfg) {
  return tape_bytes<float>(cfg);
}

extern "C" size_t snapx_adjoint_f32_work_bytes(const snapx_config* cfg) {
  return work_bytes<float>(cfg);
}

extern "C" size_t snapx_adjoint_f32_workspace_bytes(const snapx_config* cfg) {
  return workspace_bytes<float>(cfg);
}

extern "C" int snapx_adjoint_f32_rk3_step(const snapx_config* cfg, double dt, const float* u_in,
                                          float* lambda, void* workspace,
                                          size_t workspace_bytes, void* stream) {
  return rk3_step<float>(cfg, dt, u_in, lambda, workspace, workspace_bytes, stream,
                         "snapx_adjoint_f32_rk3_step");
}

extern "C" int snapx_adjoint_f32_record(const snapx_config* cfg, double dt, const float* u_in,
                                        float* u_out, void* tape, size_t tape_bytes, void* work,
                                        size_t work_bytes, void* stream) {
  return record<float>(cfg, dt, u_in, u_out, tape, tape_bytes, work, work_bytes, stream,
                       "snapx_adjoint_f32_record");
}

extern "C" int snapx_adjoint_f32_sweep(const snapx_config* cfg, double dt, void* tape,
                                       size_t tape_bytes, float* lambda, void* work,
                                       size_t work_bytes, void* stream) {
  return sweep<float>(cfg, dt, tape, tape_bytes, lambda, work, work_bytes, stream,
                      "snapx_adjoint_f32_sweep");
}
