// Forward-mode dual numbers for the local Jacobians of the adjoint.
//
// Dual<K> carries a value and K directional derivatives.  The templated
// device kernels of the forward path (interpolants, Riemann solvers, VIC
// block assembly) evaluated on Dual<K> with unit seeds give the exact
// Jacobian of the arithmetic they perform (branches follow the values, so the
// result is the derivative of the branch taken).  The adjoint kernels apply
// these Jacobians transposed.
#pragma once

#include <hip/hip_runtime.h>

namespace snapx {

template <int K>
struct Dual {
  double v;
  double d[K];

  __device__ __forceinline__ Dual() : v(0.) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.;
  }
  __device__ __forceinline__ Dual(double x) : v(x) {  // NOLINT: implicit from constants
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.;
  }
  __device__ __forceinline__ static Dual seed(double x, int k) {
    Dual r(x);
    r.d[k] = 1.;
    return r;
  }
  __device__ __forceinline__ Dual& operator+=(const Dual& b) {
    v += b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] += b.d[k];
    return *this;
  }
  __device__ __forceinline__ Dual& operator-=(const Dual& b) {
    v -= b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] -= b.d[k];
    return *this;
  }
  __device__ __forceinline__ Dual& operator*=(const Dual& b) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = d[k] * b.v + v * b.d[k];
    v *= b.v;
    return *this;
  }
  __device__ __forceinline__ Dual& operator/=(const Dual& b) {
    double inv = 1. / b.v;
    double q = v * inv;
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = (d[k] - q * b.d[k]) * inv;
    v = q;
    return *this;
  }
};

template <int K>
__device__ __forceinline__ Dual<K> operator+(Dual<K> a, const Dual<K>& b) {
  return a += b;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(Dual<K> a, const Dual<K>& b) {
  return a -= b;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(Dual<K> a, const Dual<K>& b) {
  return a *= b;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(Dual<K> a, const Dual<K>& b) {
  return a /= b;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a) {
  return a;
}

#define SNAPX_DUAL_CMP(op)                                                   \
  template <int K>                                                           \
  __device__ __forceinline__ bool operator op(const Dual<K>& a, const Dual<K>& b) { \
    return a.v op b.v;                                                       \
  }
SNAPX_DUAL_CMP(<)
SNAPX_DUAL_CMP(>)
SNAPX_DUAL_CMP(<=)
SNAPX_DUAL_CMP(>=)
SNAPX_DUAL_CMP(==)
SNAPX_DUAL_CMP(!=)
#undef SNAPX_DUAL_CMP

template <int K>
__device__ __forceinline__ Dual<K> sqrt(const Dual<K>& a) {
  Dual<K> r;
  r.v = ::sqrt(a.v);
  double g = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = g * a.d[k];
  return r;
}

template <int K>
__device__ __forceinline__ Dual<K> fabs(const Dual<K>& a) {
  return a.v < 0. ? -a : a;
}

template <int K>
__device__ __forceinline__ Dual<K> pow(const Dual<K>& a, const Dual<K>& e) {
  // a > 0 on every call site (pressures, densities)
  Dual<K> r;
  r.v = ::pow(a.v, e.v);
  double da = e.v * ::pow(a.v, e.v - 1.);
  double de = r.v * ::log(a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = da * a.d[k] + de * e.d[k];
  return r;
}

template <int K>
__device__ __forceinline__ bool isnan(const Dual<K>& a) {
  return ::isnan(a.v);
}

}  // namespace snapx
