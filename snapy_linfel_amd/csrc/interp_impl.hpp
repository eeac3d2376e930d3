// Device interpolants of the Reconstruct module.
//
// Each function returns the value of one variable at the lower face (lo) or
// upper face (hi) of cell c of a pencil p[i * s].  The arithmetic follows the
// table-driven CPU kernels of the reference (src/recon/interp_impl.h:22-122
// with the coefficient tables of src/recon/{weno5,weno3,cp5,cp3}.cpp, where
// the "cp" tables are the column-flipped "cm" tables); terms with a zero table
// coefficient are dropped, which leaves every partial sum unchanged.
#pragma once

#include "common.hpp"

namespace snapx {

// numerator (lo) and denominator (hi) of the WENO5 weighted average
template <typename T>
struct Weno5Out {
  T lo, hi;
};

// WENO5-JS on the window c-2..c+2 (src/recon/interp_impl.h:73-122).
template <typename T>
__device__ __forceinline__ Weno5Out<T> weno5_lo_nd(T f0, T f1, T f2, T f3, T f4) {
  const T c13 = T(13.) / T(12.);
  T p0 = (T(-1. / 6.) * f0 + T(5. / 6.) * f1) + T(1. / 3.) * f2;
  T p1 = (T(1. / 3.) * f1 + T(5. / 6.) * f2) + T(-1. / 6.) * f3;
  T p2 = (T(11. / 6.) * f2 + T(-7. / 6.) * f3) + T(1. / 3.) * f4;
  T b0 = c13 * sq((f0 + T(-2.) * f1) + f2) + T(.25) * sq((f0 + T(-4.) * f1) + T(3.) * f2);
  T b1 = c13 * sq((f1 + T(-2.) * f2) + f3) + T(.25) * sq(-f1 + f3);
  T b2 = c13 * sq((f2 + T(-2.) * f3) + f4) + T(.25) * sq((T(3.) * f2 + T(-4.) * f3) + f4);
  T a0 = T(.3) / sq(b0 + T(1e-6));
  T a1 = T(.6) / sq(b1 + T(1e-6));
  T a2 = T(.1) / sq(b2 + T(1e-6));
  return {(a0 * p0 + a1 * p1) + a2 * p2, (a0 + a1) + a2};
}

template <typename T>
__device__ __forceinline__ Weno5Out<T> weno5_hi_nd(T f0, T f1, T f2, T f3, T f4) {
  const T c13 = T(13.) / T(12.);
  T p0 = (T(1. / 3.) * f2 + T(5. / 6.) * f3) + T(-1. / 6.) * f4;
  T p1 = (T(-1. / 6.) * f1 + T(5. / 6.) * f2) + T(1. / 3.) * f3;
  T p2 = (T(1. / 3.) * f0 + T(-7. / 6.) * f1) + T(11. / 6.) * f2;
  T b0 = c13 * sq((f2 + T(-2.) * f3) + f4) + T(.25) * sq((T(3.) * f2 + T(-4.) * f3) + f4);
  T b1 = c13 * sq((f1 + T(-2.) * f2) + f3) + T(.25) * sq(f1 - f3);
  T b2 = c13 * sq((f0 + T(-2.) * f1) + f2) + T(.25) * sq((f0 + T(-4.) * f1) + T(3.) * f2);
  T a0 = T(.3) / sq(b0 + T(1e-6));
  T a1 = T(.6) / sq(b1 + T(1e-6));
  T a2 = T(.1) / sq(b2 + T(1e-6));
  return {(a0 * p0 + a1 * p1) + a2 * p2, (a0 + a1) + a2};
}

template <typename T>
__device__ __forceinline__ T weno5_lo(T f0, T f1, T f2, T f3, T f4) {
  Weno5Out<T> r = weno5_lo_nd(f0, f1, f2, f3, f4);
  return r.lo / r.hi;
}

template <typename T>
__device__ __forceinline__ T weno5_hi(T f0, T f1, T f2, T f3, T f4) {
  Weno5Out<T> r = weno5_hi_nd(f0, f1, f2, f3, f4);
  return r.lo / r.hi;
}

// WENO3 on the window c-1..c+1 (src/recon/interp_impl.h:34-70).
template <typename T>
__device__ __forceinline__ T weno3_lo(T f0, T f1, T f2) {
  T p0 = T(1. / 2.) * f0 + T(1. / 2.) * f1;
  T p1 = T(3. / 2.) * f1 + T(-1. / 2.) * f2;
  T b0 = sq(f0 - f1);
  T b1 = sq(f1 - f2);
  T a0 = T(2.0 / 3.0) / sq(b0 + T(1e-6));
  T a1 = T(1.0 / 3.0) / sq(b1 + T(1e-6));
  return (a0 * p0 + a1 * p1) / (a0 + a1);
}

template <typename T>
__device__ __forceinline__ T weno3_hi(T f0, T f1, T f2) {
  T p0 = T(1. / 2.) * f1 + T(1. / 2.) * f2;
  T p1 = T(-1. / 2.) * f0 + T(3. / 2.) * f1;
  T b0 = sq(-f1 + f2);
  T b1 = sq(-f0 + f1);
  T a0 = T(2.0 / 3.0) / sq(b0 + T(1e-6));
  T a1 = T(1.0 / 3.0) / sq(b1 + T(1e-6));
  return (a0 * p0 + a1 * p1) / (a0 + a1);
}

// CP5 / CP3 linear interpolants (src/recon/interp_impl.h:22-31).
template <typename T>
__device__ __forceinline__ T cp5_lo(T f0, T f1, T f2, T f3, T f4) {
  return (((T(-1. / 20.) * f0 + T(9. / 20.) * f1) + T(47. / 60.) * f2) + T(-13. / 60.) * f3) +
         T(1. / 30.) * f4;
}

template <typename T>
__device__ __forceinline__ T cp5_hi(T f0, T f1, T f2, T f3, T f4) {
  return (((T(1. / 30.) * f0 + T(-13. / 60.) * f1) + T(47. / 60.) * f2) + T(9. / 20.) * f3) +
         T(-1. / 20.) * f4;
}

template <typename T>
__device__ __forceinline__ T cp3_lo(T f0, T f1, T f2) {
  return (T(1. / 3.) * f0 + T(5. / 6.) * f1) + T(-1. / 6.) * f2;
}

template <typename T>
__device__ __forceinline__ T cp3_hi(T f0, T f1, T f2) {
  return (T(-1. / 6.) * f0 + T(5. / 6.) * f1) + T(1. / 3.) * f2;
}

// Interpolant dispatch on a pencil p[i*s].  hi = 0: lower face of cell c;
// hi = 1: upper face of cell c.  scale follows the "scale" option
// (normalise the window by its mean |value|; 0 for an all-zero window).
template <typename T>
__device__ __forceinline__ T interp_face(int type, int hi, int scale, const T* p, int s, int c) {
  switch (type) {
    case SNAPX_INTERP_DC:
      return p[c * s];
    case SNAPX_INTERP_PLM: {
      // PLMInterpImpl::forward (src/recon/plm.cpp:14-40)
      T wm = p[(c - 1) * s], w0 = p[c * s], wp = p[(c + 1) * s];
      T dwl = w0 - wm, dwr = wp - w0;
      T dw2 = dwl * dwr;
      T dwm = T(2.) * dw2 / ((dwl + dwr) + T(1.17549435082228750797e-38));
      dwm *= (dw2 >= T(0)) ? T(1) : T(0);
      return hi ? w0 + T(0.5) * dwm : w0 - T(0.5) * dwm;
    }
    case SNAPX_INTERP_CP3: {
      T f0 = p[(c - 1) * s], f1 = p[c * s], f2 = p[(c + 1) * s];
      return hi ? cp3_hi(f0, f1, f2) : cp3_lo(f0, f1, f2);
    }
    case SNAPX_INTERP_CP5: {
      T f0 = p[(c - 2) * s], f1 = p[(c - 1) * s], f2 = p[c * s], f3 = p[(c + 1) * s],
        f4 = p[(c + 2) * s];
      return hi ? cp5_hi(f0, f1, f2, f3, f4) : cp5_lo(f0, f1, f2, f3, f4);
    }
    case SNAPX_INTERP_WENO3: {
      T f0 = p[(c - 1) * s], f1 = p[c * s], f2 = p[(c + 1) * s];
      T vs = T(1.0);
      if (scale) {
        vs = ((fabs(f0) + fabs(f1)) + fabs(f2)) / T(3.0);
        if (vs == T(0.0)) return T(0.0);
        f0 /= vs;
        f1 /= vs;
        f2 /= vs;
      }
      T r = hi ? weno3_hi(f0, f1, f2) : weno3_lo(f0, f1, f2);
      return r * vs;
    }
    default: {  // WENO5
      T f0 = p[(c - 2) * s], f1 = p[(c - 1) * s], f2 = p[c * s], f3 = p[(c + 1) * s],
        f4 = p[(c + 2) * s];
      T vs = T(1.0);
      if (scale) {
        vs = ((((fabs(f0) + fabs(f1)) + fabs(f2)) + fabs(f3)) + fabs(f4)) / T(5.0);
        if (vs == T(0.0)) return T(0.0);
        f0 /= vs;
        f1 /= vs;
        f2 /= vs;
        f3 /= vs;
        f4 /= vs;
      }
      Weno5Out<T> r = hi ? weno5_hi_nd(f0, f1, f2, f3, f4) : weno5_lo_nd(f0, f1, f2, f3, f4);
      return vs * r.lo / r.hi;
    }
  }
}

// number of cells in the stencil (src/recon/interpolation.hpp stencils())
__host__ __device__ inline int interp_stencils(int type) {
  return type == SNAPX_INTERP_DC ? 1
         : (type == SNAPX_INTERP_PLM || type == SNAPX_INTERP_CP3 || type == SNAPX_INTERP_WENO3)
             ? 3
             : 5;
}

// interpolant used for velocity and pressure when shock == false
// (src/recon/interpolation.cpp:21-36)
__host__ __device__ inline int interp2_type(int type) {
  return type == SNAPX_INTERP_WENO5 ? SNAPX_INTERP_CP5
         : type == SNAPX_INTERP_WENO3 ? SNAPX_INTERP_CP3
                                      : type;
}

}  // namespace snapx
