// Device building blocks of the vertically implicit correction (VIC).
//
// All matrices are row-major, N x N with N = 3 (vic-partial: rows/columns
// IDN, IVX, IPR of the 5x5 system) or N = 5 (vic-full).  The arithmetic
// follows src/implicit/flux_decomposition_impl.h:16-128,
// src/implicit/vic_assemble_partial_impl.h:28-111,
// src/implicit/vic_assemble_full_impl.h:15-106,
// src/implicit/forward_sweep_impl.h:20-109 and src/math/{ludcmp,lubksb}.h,
// with dir = 0 (the vertical x1 axis).
#pragma once

#include "common.hpp"

namespace snapx {

// rows / columns of the 5x5 system kept by the partial (3x3) scheme
__device__ __forceinline__ int vic_row(int N, int r) { return N == 5 ? r : (r == 2 ? IPR : r); }

// flux Jacobian dF/dq at a cell (flux_decomposition_impl.h:107-128), entry
// (r, c) of the 5x5 matrix
template <typename T>
__device__ __forceinline__ T flux_jac(int r, int c, T gm1, T rho, T v1, T v2, T v3, T pres) {
  T s2 = (v1 * v1 + v2 * v2) + v3 * v3;
  switch (r * 5 + c) {
    case 0 * 5 + 1: return T(1.);
    case 1 * 5 + 0: return gm1 * s2 / T(2) - v1 * v1;
    case 1 * 5 + 1: return (T(2.) - gm1) * v1;
    case 1 * 5 + 2: return -gm1 * v2;
    case 1 * 5 + 3: return -gm1 * v3;
    case 1 * 5 + 4: return gm1;
    case 2 * 5 + 0: return -v1 * v2;
    case 2 * 5 + 1: return v2;
    case 2 * 5 + 2: return v1;
    case 3 * 5 + 0: return -v1 * v3;
    case 3 * 5 + 1: return v3;
    case 3 * 5 + 3: return v1;
    case 4 * 5 + 0:
      return ((gm1 - T(1)) * s2 / T(2) - (gm1 + T(1)) / gm1 * pres / rho) * v1;
    case 4 * 5 + 1:
      return ((gm1 + T(1)) / gm1 * pres / rho + s2 / T(2)) - gm1 * v1 * v1;
    case 4 * 5 + 2: return -gm1 * v2 * v1;
    case 4 * 5 + 3: return -gm1 * v3 * v1;
    case 4 * 5 + 4: return (gm1 + T(1)) * v1;
    default: return T(0.);
  }
}

template <typename T, int N>
__device__ __forceinline__ void flux_jacobian(T gm1, const T* w, T* J) {
  for (int r = 0; r < N; ++r)
    for (int c = 0; c < N; ++c)
      J[r * N + c] = flux_jac(vic_row(N, r), vic_row(N, c), gm1, w[IDN], w[IVX], w[IVY], w[IVZ],
                              w[IPR]);
}

// Roe average at an interface (flux_decomposition_impl.h:39-70) followed by
// |A| = R |Lambda| R^-1 (flux_decomposition_impl.h:72-105)
template <typename T, int N>
__device__ __forceinline__ void abs_jacobian(T gm1, const T* wl, const T* wr, T* A) {
  T sqrtdl = sqrt(wl[IDN]);
  T sqrtdr = sqrt(wr[IDN]);
  T isdlpdr = T(1.0) / (sqrtdl + sqrtdr);
  T r = sqrtdl * sqrtdr;
  T u = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  T v = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  T w = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  T el = wl[IPR] / gm1 + T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  T er = wr[IPR] / gm1 + T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));
  T hbar = ((el + wl[IPR]) / sqrtdl + (er + wr[IPR]) / sqrtdr) * isdlpdr;
  T p = (hbar - T(0.5) * ((sq(u) + sq(v)) + sq(w))) * gm1 / (gm1 + T(1.)) * r;
  T cs = sqrt(p * (gm1 + T(1.)) / r);

  T ke = T(0.5) * ((u * u + v * v) + w * w);
  T hp = (gm1 + T(1.)) / gm1 * p / r;
  T h = hp + ke;
  T lam[5] = {fabs(u - cs), fabs(u), fabs(u + cs), fabs(u), fabs(u)};
  // R (rows) and R^-1 (rows), flux_decomposition_impl.h:91-104
  T R[25] = {T(1.), T(1.), T(1.), T(0.), T(0.),
             u - cs, u, u + cs, T(0.), T(0.),
             v, v, v, T(1.), T(0.),
             w, w, w, T(0.), T(1.),
             h - u * cs, ke, h + u * cs, v, w};
  T two_cs_hp = T(2.) * cs * hp;
  T two_hp = T(2.) * hp;
  T Ri[25] = {(cs * ke + u * hp) / two_cs_hp, (-hp - cs * u) / two_cs_hp, -v / two_hp,
              -w / two_hp, T(1.) / two_hp,
              (hp - ke) / hp, u / hp, v / hp, w / hp, T(-1.) / hp,
              (cs * ke - u * hp) / two_cs_hp, (hp - cs * u) / two_cs_hp, -v / two_hp,
              -w / two_hp, T(1.) / two_hp,
              -v, T(0.), T(1.), T(0.), T(0.),
              -w, T(0.), T(0.), T(1.), T(0.)};
  for (int i = 0; i < N; ++i) {
    int ri = vic_row(N, i);
    for (int j = 0; j < N; ++j) {
      int cj = vic_row(N, j);
      T s = T(0.);
      for (int k = 0; k < 5; ++k) s += (R[ri * 5 + k] * lam[k]) * Ri[k * 5 + cj];
      A[i * N + j] = s;
    }
  }
}

// 1/x within an ulp or two: the hardware reciprocal estimate refined by two
// Newton steps, in place of the correctly rounded division sequence.  For
// the fused path only (normal, finite x); the FP64 plain path keeps IEEE
// division.
__device__ __forceinline__ double rcp_fast(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// abs_jacobian with the shared divisors inverted once, by rcp_fast; the same
// matrix up to rounding.  Used by the fused path, whose state is FP32; the
// FP64 plain path keeps the reference's arithmetic.
template <typename T, int N>
__device__ __forceinline__ void abs_jacobian_fast(T gm1, const T* wl, const T* wr, T* A) {
  T sqrtdl = sqrt(wl[IDN]);
  T sqrtdr = sqrt(wr[IDN]);
  T isdlpdr = rcp_fast(sqrtdl + sqrtdr);
  T r = sqrtdl * sqrtdr;
  T u = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  T v = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  T w = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  T igm1 = rcp_fast(gm1);
  T el = wl[IPR] * igm1 + T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  T er = wr[IPR] * igm1 + T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));
  // (e + p) / sqrt(rho_l) = (e + p) sqrt(rho_r) / r
  T ir = rcp_fast(r);
  T hbar = ((el + wl[IPR]) * sqrtdr + (er + wr[IPR]) * sqrtdl) * isdlpdr * ir;
  T p = (hbar - T(0.5) * ((sq(u) + sq(v)) + sq(w))) * gm1 * rcp_fast(gm1 + T(1.)) * r;
  T cs = sqrt(p * (gm1 + T(1.)) * ir);
  T ke = T(0.5) * ((u * u + v * v) + w * w);
  T hp = (gm1 + T(1.)) * igm1 * p * ir;
  T h = hp + ke;
  T lam[5] = {fabs(u - cs), fabs(u), fabs(u + cs), fabs(u), fabs(u)};
  T R[25] = {T(1.), T(1.), T(1.), T(0.), T(0.),
             u - cs, u, u + cs, T(0.), T(0.),
             v, v, v, T(1.), T(0.),
             w, w, w, T(0.), T(1.),
             h - u * cs, ke, h + u * cs, v, w};
  T ihp = rcp_fast(hp);
  T i2hp = T(0.5) * ihp;
  T i2cshp = i2hp * rcp_fast(cs);
  T Ri[25] = {(cs * ke + u * hp) * i2cshp, (-hp - cs * u) * i2cshp, -v * i2hp, -w * i2hp, i2hp,
              (hp - ke) * ihp, u * ihp, v * ihp, w * ihp, -ihp,
              (cs * ke - u * hp) * i2cshp, (hp - cs * u) * i2cshp, -v * i2hp, -w * i2hp, i2hp,
              -v, T(0.), T(1.), T(0.), T(0.),
              -w, T(0.), T(0.), T(1.), T(0.)};
#pragma unroll
  for (int i = 0; i < N; ++i) {
    int ri = vic_row(N, i);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      int cj = vic_row(N, j);
      T s = T(0.);
#pragma unroll
      for (int k = 0; k < 5; ++k) s += (R[ri * 5 + k] * lam[k]) * Ri[k * 5 + cj];
      A[i * N + j] = s;
    }
  }
}

template <typename T, bool Fast = false>
__device__ __forceinline__ void inverse3(const T* m, T* inv) {
#define M_(r, c) m[(r) * 3 + (c)]
#define COF_(i, j)                                                         \
  (M_(((i) + 1) % 3, ((j) + 1) % 3) * M_(((i) + 2) % 3, ((j) + 2) % 3) - \
   M_(((i) + 1) % 3, ((j) + 2) % 3) * M_(((i) + 2) % 3, ((j) + 1) % 3))
  T c00 = COF_(0, 0), c10 = COF_(1, 0), c20 = COF_(2, 0);
  T det = (c00 * M_(0, 0) + c10 * M_(1, 0)) + c20 * M_(2, 0);
  T invdet = Fast ? rcp_fast(det) : T(1.0) / det;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) inv[i * 3 + j] = COF_(j, i) * invdet;
#undef COF_
#undef M_
}

template <typename T, int N>
__device__ __forceinline__ void matmul(const T* a, const T* b, T* c) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      T s = T(0.);
#pragma unroll
      for (int k = 0; k < N; ++k) s += a[i * N + k] * b[k * N + j];
      c[i * N + j] = s;
    }
}

template <typename T, int N>
__device__ __forceinline__ void matvec(const T* a, const T* x, T* y) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    T s = T(0.);
#pragma unroll
    for (int k = 0; k < N; ++k) s += a[i * N + k] * x[k];
    y[i] = s;
  }
}

// LU with partial pivoting (src/math/ludcmp.h:25-76), row-major 5x5
template <typename T>
__device__ __forceinline__ void ludcmp5(T* a, int* indx) {
  const int N = 5;
  T vv[5];
  for (int i = 0; i < N; i++) {
    T big = T(0.0);
    for (int j = 0; j < N; j++) {
      T temp = fabs(a[i * N + j]);
      if (temp > big) big = temp;
    }
    vv[i] = T(1.0) / big;
  }
  for (int j = 0; j < N; j++) {
    for (int i = 0; i < j; i++) {
      T sum = a[i * N + j];
      for (int k = 0; k < i; k++) sum -= a[i * N + k] * a[k * N + j];
      a[i * N + j] = sum;
    }
    T big = T(0.0);
    int imax = j;
    for (int i = j; i < N; i++) {
      T sum = a[i * N + j];
      for (int k = 0; k < j; k++) sum -= a[i * N + k] * a[k * N + j];
      a[i * N + j] = sum;
      T dum = vv[i] * fabs(sum);
      if (dum >= big) {
        big = dum;
        imax = i;
      }
    }
    if (j != imax) {
      for (int k = 0; k < N; k++) {
        T dum = a[imax * N + k];
        a[imax * N + k] = a[j * N + k];
        a[j * N + k] = dum;
      }
      vv[imax] = vv[j];
    }
    indx[j] = imax;
    if (j != N - 1) {
      T dum = T(1.0) / a[j * N + j];
      for (int i = j + 1; i < N; i++) a[i * N + j] *= dum;
    }
  }
}

// back substitution with a matrix right-hand side b (5 x M, row-major)
// (src/math/lubksb.h:49-71)
template <typename T, int M>
__device__ __forceinline__ void lubksb5(const T* a, const int* indx, T* b) {
  const int N = 5;
  for (int i = 0; i < N; ++i) {
    int ip = indx[i];
    if (ip != i)
      for (int m = 0; m < M; ++m) {
        T t = b[ip * M + m];
        b[ip * M + m] = b[i * M + m];
        b[i * M + m] = t;
      }
    for (int j = 0; j < i; ++j)
      for (int m = 0; m < M; ++m) b[i * M + m] -= a[i * N + j] * b[j * M + m];
  }
  for (int i = N - 1; i >= 0; --i) {
    for (int j = i + 1; j < N; ++j)
      for (int m = 0; m < M; ++m) b[i * M + m] -= a[i * N + j] * b[j * M + m];
    for (int m = 0; m < M; ++m) b[i * M + m] /= a[i * N + i];
  }
}

// Coefficients a_i, b_i, c_i of layer i of a column from the primitive states
// of cells i-1 (wm), i (w0), i+1 (wp).  first / last apply the reflecting
// column ends (Bnd = diag(1, -1, 1[, 1, 1])).
template <typename T, int N>
__device__ __forceinline__ void vic_assemble_cell(const T* wm, const T* w0, const T* wp, T gammad,
                                                  T area_i, T area_ip1, T vol, T dt, T grav,
                                                  bool first, bool last, T* a, T* b, T* c) {
  T Jp[N * N], Jc[N * N], Jn[N * N], Am[N * N], Ap[N * N];
  T gm1 = gammad - T(1.);
  flux_jacobian<T, N>(gm1, wm, Jp);
  flux_jacobian<T, N>(gm1, w0, Jc);
  flux_jacobian<T, N>(gm1, wp, Jn);
  T gm1r = T(0.5) * (gammad + gammad) - T(1.);
  abs_jacobian<T, N>(gm1r, wm, w0, Am);
  abs_jacobian<T, N>(gm1r, w0, wp, Ap);
  T half_inv_vol = T(0.5) / vol;
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int q = 0; q < N; ++q) {
      int o = r * N + q;
      T phi = T(0.);
      if ((r == 1 && q == 0) || (vic_row(N, r) == IPR && q == 1)) phi = grav;
      a[o] = (Am[o] * area_i + Ap[o] * area_ip1 + (area_ip1 - area_i) * Jc[o]) * half_inv_vol -
             phi;
      if (r == q) a[o] += T(1.) / dt;
      b[o] = -(Am[o] + Jp[o]) * area_i * half_inv_vol;
      c[o] = -(Ap[o] - Jn[o]) * area_ip1 * half_inv_vol;
    }
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int q = 0; q < N; ++q) {
      T bnd = (q == 1) ? T(-1.) : T(1.);
      if (first) a[r * N + q] += b[r * N + q] * bnd;
      if (last) a[r * N + q] += c[r * N + q] * bnd;
    }
}

}  // namespace snapx
