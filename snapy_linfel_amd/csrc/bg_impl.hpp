// Hydrostatic background of the FP32 perturbation path of the fused step.
//
// The state is stored as u' = u - ubar(x1) in FP32, about a background at
// rest that varies with x1 only: ubar = (rho_b, 0, 0, 0, p_b / (gamma - 1)),
// given in FP64 per x1 cell level (ghost levels included).  The fused kernels
// then work on perturbations wherever the total state would cancel: the
// pressure jumps and the x1 pressure gradient against gravity, the density
// and energy jumps of the Roe dissipation.  Each x1 face carries a reference
// (rho_f, p_f); a face state is the reference plus a perturbation, and the
// background parts of the reconstructions enter as FP64-computed offsets
// (background sub-stencil value minus the reference).  The background flux
// of the face, (0, p_f, 0, 0, 0), and the background's gravity combine into
// one FP64 residual per level.  Along x2 and x3 the background is uniform,
// so the reconstruction of the perturbation is the perturbation of the
// reconstruction exactly (WENO5 is invariant under adding a constant to its
// window; CP5 coefficients sum to one) and the cell's own level is the
// reference.
//
// Tables (built on the host in FP64 by bg_build_tables): FP64 rho_b, p_b per
// level, then FP32 arrays per level l (L = n1) and per x1 face f (F = n1 + 1,
// face f is the lower face of cell f).
#pragma once

#include <cmath>
#include <vector>

#include "common.hpp"

namespace snapx {

constexpr int kBgLevelRows = 10;
constexpr int kBgFaceRows = 10;

struct BgLayout {
  int L, F;
  // FP32 offsets (floats) of the per-level and per-face arrays
  int rho_c, p_c, gdr, gdp, res;  // [L]
  // ghost refill of the physical x1 ghosts (rows at the ghost levels):
  // cn = n (+-1) a g dz, qk = qbar^(gamma / (gamma - 1)),
  // qk2 = qbar^(1 / (gamma - 1)) with qbar = 1 + cn rho_b / p_b of the
  // background's boundary cell, and the refilled background minus the
  // background at the ghost level, dpg (pressure) and drg (density)
  int cn, qk, qk2, dpg, drg;      // [L]
  int rho_f, p_f;                 // [F]
  int drl, drr;                   // WENO5 density offsets, [3][F] each
  int epl, epr;                   // CP5 pressure offsets, [F] each
  int nfloat;
  __host__ __device__ explicit BgLayout(int n1) : L(n1), F(n1 + 1) {
    rho_c = 0, p_c = L, gdr = 2 * L, gdp = 3 * L, res = 4 * L;
    cn = 5 * L, qk = 6 * L, qk2 = 7 * L, dpg = 8 * L, drg = 9 * L;
    rho_f = kBgLevelRows * L, p_f = rho_f + F;
    drl = p_f + F, drr = drl + 3 * F;
    epl = drr + 3 * F, epr = epl + F;
    nfloat = epr + F;
  }
};

// bytes of the device tables: FP64 rho_b, p_b, then the FP32 arrays
inline size_t bg_bytes(int n1) {
  BgLayout B(n1);
  return sizeof(double) * 2 * (size_t)n1 + sizeof(float) * (size_t)B.nfloat;
}

// WENO5 sub-stencil values of the lower (lo, window c-2..c+2) and upper (hi)
// face of the window centre, in the order of weno5_lo_nd / weno5_hi_nd
template <typename T>
__host__ __device__ inline void weno5_sub_lo(const T* f, T* p) {
  p[0] = (T(-1. / 6.) * f[0] + T(5. / 6.) * f[1]) + T(1. / 3.) * f[2];
  p[1] = (T(1. / 3.) * f[1] + T(5. / 6.) * f[2]) + T(-1. / 6.) * f[3];
  p[2] = (T(11. / 6.) * f[2] + T(-7. / 6.) * f[3]) + T(1. / 3.) * f[4];
}
template <typename T>
__host__ __device__ inline void weno5_sub_hi(const T* f, T* p) {
  p[0] = (T(1. / 3.) * f[2] + T(5. / 6.) * f[3]) + T(-1. / 6.) * f[4];
  p[1] = (T(-1. / 6.) * f[1] + T(5. / 6.) * f[2]) + T(1. / 3.) * f[3];
  p[2] = (T(1. / 3.) * f[0] + T(-7. / 6.) * f[1]) + T(11. / 6.) * f[2];
}

// WENO5 face value of a perturbation: the nonlinear weights from the total
// values t[0..4] of the window (smoothness indicators and linear weights of
// weno5_lo_nd / weno5_hi_nd), the sub-stencil values from the perturbations
// f[0..4] plus the background offsets d[0..2]
template <bool kHi>
__device__ __forceinline__ float weno5_off(const float* t, const float* f, const float* d) {
  const float c13 = 13.f / 12.f;
  float b0, b1, b2, p[3];
  if (kHi) {
    b0 = c13 * sq((t[2] - 2.f * t[3]) + t[4]) + .25f * sq((3.f * t[2] - 4.f * t[3]) + t[4]);
    b1 = c13 * sq((t[1] - 2.f * t[2]) + t[3]) + .25f * sq(t[1] - t[3]);
    b2 = c13 * sq((t[0] - 2.f * t[1]) + t[2]) + .25f * sq((t[0] - 4.f * t[1]) + 3.f * t[2]);
    weno5_sub_hi(f, p);
  } else {
    b0 = c13 * sq((t[0] - 2.f * t[1]) + t[2]) + .25f * sq((t[0] - 4.f * t[1]) + 3.f * t[2]);
    b1 = c13 * sq((t[1] - 2.f * t[2]) + t[3]) + .25f * sq(t[3] - t[1]);
    b2 = c13 * sq((t[2] - 2.f * t[3]) + t[4]) + .25f * sq((3.f * t[2] - 4.f * t[3]) + t[4]);
    weno5_sub_lo(f, p);
  }
  const float a0 = .3f / sq(b0 + 1e-6f);
  const float a1 = .6f / sq(b1 + 1e-6f);
  const float a2 = .1f / sq(b2 + 1e-6f);
  return ((a0 * (p[0] + d[0]) + a1 * (p[1] + d[1])) + a2 * (p[2] + d[2])) / ((a0 + a1) + a2);
}

// Isentropic refill of ghost layer n (closed form of IsentropicRefill):
// with q = 1 + cn rho_b / p_b, p_g = p_b q^k and rho_g = rho_b q^k2,
// k = gamma / (gamma - 1), k2 = 1 / (gamma - 1).  In perturbation form, for
// a boundary cell (rho_bar + rho', p_bar + p'): q - qbar = cn (rho' p_bar -
// rho_bar p') / (p p_bar) has no cancellation, q^k - qbar^k = qbar^k
// expm1(k log1p((q - qbar) / qbar)), and
//   p_g - p_bar_g = p' q^k + p_bar (q^k - qbar^k) + (p_bar qbar^k - p_bar_g)
// (the last term from the tables, dpg), likewise the density.  FP32
// throughout, accurate to the perturbations' own rounding.
__device__ __forceinline__ void bg_refill_pert(float rbb, float pbb, float rp, float pp, float cn,
                                               float qk, float qk2, float dpg, float drg,
                                               float gammad, float& rg, float& pg) {
  const float p = pbb + pp;
  const float dq = cn * (rp * pbb - rbb * pp) / (p * pbb);
  const float qbar = 1.f + cn * rbb / pbb;
  const float L = log1pf(dq / qbar);
  const float gm1 = gammad - 1.f;
  const float ek = expm1f(gammad / gm1 * L), ek2 = expm1f(L / gm1);
  pg = pp * (qk * (1.f + ek)) + pbb * (qk * ek) + dpg;
  rg = rp * (qk2 * (1.f + ek2)) + rbb * (qk2 * ek2) + drg;
}

// host: x1 source level of ghost level l through the x1 boundary functions
inline int bg_source_level(int l, int ng, int nx1, int bin, int bout) {
  int x = l - ng;
  if (x < 0) {
    if (bin == SNAPX_BC_REFLECTING) return ng + (-1 - x);
    if (bin == SNAPX_BC_PERIODIC) return ng + x + nx1;
    return ng;
  }
  if (x >= nx1) {
    if (bout == SNAPX_BC_REFLECTING) return ng + (2 * nx1 - 1 - x);
    if (bout == SNAPX_BC_PERIODIC) return ng + x - nx1;
    return ng + nx1 - 1;
  }
  return l;
}

// host: the device image of the tables (FP64 part, FP32 part) from the
// background rho_b[n1], p_b[n1] of a Cartesian block (x1 width dx1, gravity g1)
inline void bg_build_tables(const snapx_config* c, const double* rb, const double* pb,
                            std::vector<double>& d64, std::vector<float>& f32) {
  const int ng = c->nghost, nx1 = c->nx1, n1 = nx1 + 2 * ng;
  const double dx1 = (c->x1max - c->x1min) / nx1;
  BgLayout B(n1);
  d64.assign(2 * (size_t)n1, 0.);
  f32.assign((size_t)B.nfloat, 0.f);
  std::vector<double> pf(B.F, 0.);
  for (int l = 0; l < n1; ++l) {
    d64[l] = rb[l];
    d64[n1 + l] = pb[l];
    f32[B.rho_c + l] = (float)rb[l];
    f32[B.p_c + l] = (float)pb[l];
    const int s = bg_source_level(l, ng, nx1, c->bc[0], c->bc[1]);
    f32[B.gdr + l] = (float)(rb[s] - rb[l]);
    f32[B.gdp + l] = (float)(pb[s] - pb[l]);
    // physical x1 ghosts: refill of the background's boundary cell, layer
    // n = 1, 2, 3 away from it (without gravity a copy of the boundary cell,
    // as the column solve takes it)
    const bool lower = l < ng, upper = l >= ng + nx1;
    if (lower || upper) {
      const int lb = lower ? ng : ng + nx1 - 1;
      const int n = lower ? ng - l : l - (ng + nx1 - 1);
      const double gam = c->gammad, gm = gam - 1.;
      const double cn = n * (lower ? 1. : -1.) * (gm / gam) * (-c->grav1 * dx1);
      const double q = 1. + cn * rb[lb] / pb[lb];
      const double qk = std::pow(q, gam / gm), qk2 = std::pow(q, 1. / gm);
      f32[B.cn + l] = (float)cn;
      f32[B.qk + l] = (float)qk;
      f32[B.qk2 + l] = (float)qk2;
      f32[B.dpg + l] = (float)(pb[lb] * qk - pb[l]);
      f32[B.drg + l] = (float)(rb[lb] * qk2 - rb[l]);
    }
  }
  for (int f = ng; f <= ng + nx1; ++f) {
    // windows of the cells below (f - 1, upper face) and above (f, lower face)
    const double* wl = rb + f - 3;
    const double* wr = rb + f - 2;
    const double rf = 0.5 * (rb[f - 1] + rb[f]);
    const double* ql = pb + f - 3;
    const double* qr = pb + f - 2;
    const double cpl = (((1. / 30.) * ql[0] + (-13. / 60.) * ql[1]) + (47. / 60.) * ql[2] +
                        (9. / 20.) * ql[3]) + (-1. / 20.) * ql[4];
    const double cpr = (((-1. / 20.) * qr[0] + (9. / 20.) * qr[1]) + (47. / 60.) * qr[2] +
                        (-13. / 60.) * qr[3]) + (1. / 30.) * qr[4];
    pf[f] = 0.5 * (cpl + cpr);
    f32[B.rho_f + f] = (float)rf;
    f32[B.p_f + f] = (float)pf[f];
    double sl[3], sr[3];
    weno5_sub_hi<double>(wl, sl);
    weno5_sub_lo<double>(wr, sr);
    for (int k = 0; k < 3; ++k) {
      f32[B.drl + k * B.F + f] = (float)(sl[k] - rf);
      f32[B.drr + k * B.F + f] = (float)(sr[k] - rf);
    }
    f32[B.epl + f] = (float)(cpl - pf[f]);
    f32[B.epr + f] = (float)(cpr - pf[f]);
  }
  // the background's x1 flux (0, p_f, 0) and gravity: momentum tendency per
  // unit time of each interior level
  for (int l = ng; l < ng + nx1; ++l)
    f32[B.res + l] = (float)(-(pf[l + 1] - pf[l]) / dx1 + rb[l] * c->grav1);
}

}  // namespace snapx
