// Parallel cyclic reduction for the block-tridiagonal column systems.
//
// Lane t of a workgroup holds row i of one column's system
//     B_i x_{i-1} + A_i x_i + C_i x_{i+1} = r_i,   0 <= i < nl,
// with N x N blocks (B_0 = C_{nl-1} = 0); whole columns sit in consecutive
// lanes.  At level s every row publishes its normalised row
// (A^-1 B, A^-1 C, A^-1 r) to LDS and eliminates its couplings to rows i -+ s
// with those of its neighbours, so that after ceil(log2 nl) levels every row
// stands alone.  LDS is laid out as structure of arrays (element e of lane t
// at sh[e * T + t]) so that neighbouring lanes read neighbouring words.
#pragma once

#include <hip/hip_runtime.h>

#include "vic_impl.hpp"

namespace snapx {

template <int N>
struct PcrShape {
  static constexpr int kThreads = N == 3 ? 256 : 128;
  static constexpr int kPub = 2 * N * N + N;  // doubles published per lane
};

// Solves A X = R for the N x M right-hand side R (row-major, overwritten),
// partial pivoting with fully unrolled row selection (no dynamic indexing).
template <int N, int M>
__device__ __forceinline__ void solve_rows(double* A, double* R) {
#pragma unroll
  for (int k = 0; k < N; ++k) {
    // pivot row p >= k of largest |A[., k]|
    int p = k;
    double big = fabs(A[k * N + k]);
#pragma unroll
    for (int r = k + 1; r < N; ++r) {
      double v = fabs(A[r * N + k]);
      if (v > big) {
        big = v;
        p = r;
      }
    }
#pragma unroll
    for (int r = k + 1; r < N; ++r) {
      if (p == r) {
#pragma unroll
        for (int c = 0; c < N; ++c) {
          double t = A[k * N + c];
          A[k * N + c] = A[r * N + c];
          A[r * N + c] = t;
        }
#pragma unroll
        for (int c = 0; c < M; ++c) {
          double t = R[k * M + c];
          R[k * M + c] = R[r * M + c];
          R[r * M + c] = t;
        }
      }
    }
    double inv = 1.0 / A[k * N + k];
#pragma unroll
    for (int r = 0; r < N; ++r) {
      if (r == k) continue;
      double f = A[r * N + k] * inv;
#pragma unroll
      for (int c = k; c < N; ++c) A[r * N + c] -= f * A[k * N + c];
#pragma unroll
      for (int c = 0; c < M; ++c) R[r * M + c] -= f * R[k * M + c];
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double inv = 1.0 / A[k * N + k];
#pragma unroll
    for (int c = 0; c < M; ++c) R[k * M + c] *= inv;
  }
}

// normalised row of lane t: Pb = A^-1 B, Pc = A^-1 C, pr = A^-1 r, written
// to LDS (structure of arrays, stride T)
template <int N, bool Fast>
__device__ __forceinline__ void pcr_publish(const double* A, const double* B, const double* C,
                                            const double* r, double* sh, int T, int t) {
  const int NN = N * N;
  if (N == 3) {
    double inv[9], x[9];
    inverse3<double, Fast>(A, inv);
    matmul<double, 3>(inv, B, x);
#pragma unroll
    for (int e = 0; e < 9; ++e) sh[e * T + t] = x[e];
    matmul<double, 3>(inv, C, x);
#pragma unroll
    for (int e = 0; e < 9; ++e) sh[(9 + e) * T + t] = x[e];
    matvec<double, 3>(inv, r, x);
#pragma unroll
    for (int e = 0; e < 3; ++e) sh[(18 + e) * T + t] = x[e];
  } else {
    const int M = 2 * N + 1;
    double a[NN], R[N * M];
#pragma unroll
    for (int e = 0; e < NN; ++e) a[e] = A[e];
#pragma unroll
    for (int row = 0; row < N; ++row) {
#pragma unroll
      for (int c = 0; c < N; ++c) {
        R[row * M + c] = B[row * N + c];
        R[row * M + N + c] = C[row * N + c];
      }
      R[row * M + 2 * N] = r[row];
    }
    solve_rows<N, M>(a, R);
#pragma unroll
    for (int row = 0; row < N; ++row) {
#pragma unroll
      for (int c = 0; c < N; ++c) {
        sh[(row * N + c) * T + t] = R[row * M + c];
        sh[(NN + row * N + c) * T + t] = R[row * M + N + c];
      }
      sh[(2 * NN + row) * T + t] = R[row * M + 2 * N];
    }
  }
}

// Barrier among the lanes of one column.  When every column lies within one
// wavefront (64 lanes) a workgroup barrier is not needed: a wavefront executes
// its LDS instructions in order, so the compiler only must not move LDS
// accesses across this point.
__device__ __forceinline__ void column_barrier(bool wave) {
  if (wave) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Sum over the aligned groups of `width` lanes (a power of two, at most 64)
// of a wavefront.
__device__ __forceinline__ double wave_sum(double x, int width = 64) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    if (m < width) x += __shfl_xor(x, m);
  return x;
}

// Solves the system held by the workgroup (all threads call it; act is false
// for lanes without a row, whose published data is never read).  On return
// x holds the solution of row i; A, B, C, r are destroyed.  wave: every
// column lies within one wavefront (see column_barrier).  Fast: 3x3 inverses
// through rcp_fast (N = 3 only; the N = 5 pivoted solve divides).
template <int N, bool Fast = false>
__device__ __forceinline__ void pcr_solve(double* A, double* B, double* C, double* r, double* x,
                                          double* __restrict__ sh, int T, int t, int i, int nl,
                                          bool act, bool wave = false) {
  constexpr int NN = N * N;
  if (!act) {
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      A[e] = (e % (N + 1)) == 0 ? 1. : 0.;
      B[e] = C[e] = 0.;
    }
#pragma unroll
    for (int e = 0; e < N; ++e) r[e] = 0.;
  }
  for (int sft = 1; sft < nl; sft <<= 1) {
    if constexpr (N == 3) {
      if (2 * sft >= nl) {
        // Last level: nl <= 2 sft, so a row couples to row i - sft or to
        // row i + sft, never to both, and the coupling of that row back to
        // it is its only one (its B is zero below sft and its C from sft
        // on).  Row j < sft publishes A^-1 C_j for row j + sft, row j >= sft
        // publishes A^-1 B_j for row j - sft: one product instead of two,
        // and one elimination instead of two with the new couplings (which
        // are zero).  N = 5 keeps the general level: its registers are at
        // the limit.
        const bool lo = i < sft;
        double X[NN], inv[NN], P[NN], p[N];
#pragma unroll
        for (int e = 0; e < NN; ++e) X[e] = lo ? C[e] : B[e];
        inverse3<double, Fast>(A, inv);
        matmul<double, 3>(inv, X, P);
        matvec<double, 3>(inv, r, p);
#pragma unroll
        for (int e = 0; e < NN; ++e) sh[e * T + t] = P[e];
#pragma unroll
        for (int e = 0; e < N; ++e) sh[(NN + e) * T + t] = p[e];
        column_barrier(wave);
        const bool hm = act && i - sft >= 0;
        const bool hp = act && i + sft < nl;
        if (hm || hp) {
          const int tn = hm ? t - sft : t + sft;
          double y[NN];
#pragma unroll
          for (int e = 0; e < NN; ++e) {
            P[e] = sh[e * T + tn];
            X[e] = hm ? B[e] : C[e];
          }
#pragma unroll
          for (int e = 0; e < N; ++e) p[e] = sh[(NN + e) * T + tn];
          matmul<double, N>(X, P, y);
#pragma unroll
          for (int e = 0; e < NN; ++e) A[e] -= y[e];
          matvec<double, N>(X, p, y);
#pragma unroll
          for (int e = 0; e < N; ++e) r[e] -= y[e];
        }
        column_barrier(wave);
        break;
      }
    }
    pcr_publish<N, Fast>(A, B, C, r, sh, T, t);
    column_barrier(wave);
    const bool hm = act && i - sft >= 0;
    const bool hp = act && i + sft < nl;
    double nB[NN], nC[NN];
#pragma unroll
    for (int e = 0; e < NN; ++e) nB[e] = nC[e] = 0.;
    if (hm) {
      const int tm = t - sft;
      double Pb[NN], Pc[NN], pr[N], y[NN];
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        Pb[e] = sh[e * T + tm];
        Pc[e] = sh[(NN + e) * T + tm];
      }
#pragma unroll
      for (int e = 0; e < N; ++e) pr[e] = sh[(2 * NN + e) * T + tm];
      matmul<double, N>(B, Pb, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) nB[e] = -y[e];
      matmul<double, N>(B, Pc, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) A[e] -= y[e];
      matvec<double, N>(B, pr, y);
#pragma unroll
      for (int e = 0; e < N; ++e) r[e] -= y[e];
    }
    if (hp) {
      const int tp = t + sft;
      double Pb[NN], Pc[NN], pr[N], y[NN];
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        Pb[e] = sh[e * T + tp];
        Pc[e] = sh[(NN + e) * T + tp];
      }
#pragma unroll
      for (int e = 0; e < N; ++e) pr[e] = sh[(2 * NN + e) * T + tp];
      matmul<double, N>(C, Pc, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) nC[e] = -y[e];
      matmul<double, N>(C, Pb, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) A[e] -= y[e];
      matvec<double, N>(C, pr, y);
#pragma unroll
      for (int e = 0; e < N; ++e) r[e] -= y[e];
    }
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      B[e] = nB[e];
      C[e] = nC[e];
    }
    column_barrier(wave);
  }
  if (N == 3) {
    double inv[9];
    inverse3<double, Fast>(A, inv);
    matvec<double, 3>(inv, r, x);
  } else {
    solve_rows<N, 1>(A, r);
#pragma unroll
    for (int e = 0; e < N; ++e) x[e] = r[e];
  }
}

// Sum of x over the lanes of each column (columns of L = 2^lshift >= 32
// lanes: aligned parts of a wavefront or whole wavefronts): wave reductions,
// then, for columns taller than a wavefront, the per-wavefront partials
// through sh[] in a fixed order.  Every thread calls it; sh must not be in
// use by other lanes (a barrier precedes the partial writes).  Deterministic.
template <int K>
__device__ __forceinline__ void column_sums(double* x, double* __restrict__ sh, int t,
                                            int lanes, int lshift) {
  const int width = lanes < 64 ? lanes : 64;
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] = wave_sum(x[k], width);
  if (lanes > 64) {
    const int wv = t >> 6, nwv = lanes >> 6;
    __syncthreads();
    if ((t & 63) == 0)
#pragma unroll
      for (int k = 0; k < K; ++k) sh[K * wv + k] = x[k];
    __syncthreads();
    const int w0 = (t >> lshift) * nwv;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double acc = 0.;
      for (int m = 0; m < nwv; ++m) acc += sh[K * (w0 + m) + k];
      x[k] = acc;
    }
    __syncthreads();
  }
}

}  // namespace snapx
