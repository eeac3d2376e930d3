// Riemann solvers of the FP32 fused path in perturbation form: the states are
// carried as a common reference at the face (the hydrostatic background at
// rest: density rho_f, pressure p_f) plus perturbations, and the flux is
// returned without the background flux (0, p_f n, 0).
//
// wl / wr are the total primitive states in the face frame (normal velocity
// in slot IVX); rl, rr and pl, pr are the density and pressure perturbations
// of the two states against the face reference.  Total values feed only the
// nonlinear coefficients (sound speeds, Roe averages, upwind selections);
// every difference of pressure, energy or density across the face, and the
// pressure in the normal momentum flux, comes from the perturbations.  In
// exact arithmetic each solver equals its total-state form (riemann_impl.hpp:
// LMARS src/riemann/lmars_impl.h:16-78, HLLC src/riemann/hllc_impl.h:16-130,
// Roe src/riemann/roe_impl.h:23-211) minus (0, p_f, 0, 0, 0); in FP32 it
// avoids the cancellation of the hydrostatic pressure (about 1e5 Pa, stored to
// about 0.01 Pa) against the perturbations that drive the flow.
#pragma once

#include "common.hpp"

namespace snapx {

// LMARS: face pressure pbar - p_f and interface velocity from the pressure
// perturbations; the upwind state carries the total enthalpy.
template <typename T>
__device__ __forceinline__ void lmars_pert(const T* wl, const T* wr, T rl, T rr, T pl, T pr,
                                           T gammad, T* flx) {
  (void)rl;
  (void)rr;
  const T gm1 = gammad - T(1.);
  const T il = T(1.) / wl[IDN], ir = T(1.) / wr[IDN];
  T hl = (wl[IPR] / gm1) * il;
  T hr = (wr[IPR] / gm1) * ir;
  hl += T(0.5) * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ])) + wl[IPR] * il;
  hr += T(0.5) * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ])) + wr[IPR] * ir;
  const T rhobar = T(0.5) * (wl[IDN] + wr[IDN]);
  const T cbar = sqrt(T(0.5) * gammad * (wl[IPR] + wr[IPR]) / rhobar);
  const T rc = rhobar * cbar;
  const T pbar = T(0.5) * (pl + pr) + T(0.5) * rc * (wl[IVX] - wr[IVX]);
  const T ubar = T(0.5) * (wl[IVX] + wr[IVX]) + (T(0.5) / rc) * (pl - pr);
  const bool left = ubar > T(0.);
  const T* s = left ? wl : wr;
  const T h = left ? hl : hr;
  const T m = ubar * s[IDN];
  flx[IDN] = m;
  flx[IVX] = m * s[IVX] + pbar;
  flx[IVY] = m * s[IVY];
  flx[IVZ] = m * s[IVZ];
  flx[IPR] = m * h;
}

// HLLC with PVRS wave speeds: pmid, the contact speed and the contact
// pressure from perturbations (the weights of the normal momentum flux and
// the contact pressure sum to one, so the background part is exactly p_f).
template <typename T>
__device__ __forceinline__ void hllc_pert(const T* wl, const T* wr, T rl, T rr, T pl, T pr,
                                          T pf, T gammad, T* flx) {
  (void)rl;
  (void)rr;
  const T TINY_NUMBER = T(1.0e-10);
  const T gm1 = gammad - T(1.);
  T el = wl[IPR] / gm1;
  T er = wr[IPR] / gm1;
  const T cl = sqrt(gammad * wl[IPR] / wl[IDN]);
  const T cr = sqrt(gammad * wr[IPR] / wr[IDN]);
  el += T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  er += T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));

  const T rhoa = T(0.5) * (wl[IDN] + wr[IDN]);
  const T ca = T(0.5) * (cl + cr);
  const T pmid = T(0.5) * ((pl + pr) + (wl[IVX] - wr[IVX]) * rhoa * ca);
  const T g2 = (gammad + T(1.)) / (T(2.) * gammad);
  const T ql = (pmid <= pl) ? T(1.) : sqrt(T(1.) + g2 * ((pmid - pl) / wl[IPR]));
  const T qr = (pmid <= pr) ? T(1.) : sqrt(T(1.) + g2 * ((pmid - pr) / wr[IPR]));

  const T al = wl[IVX] - cl * ql;
  const T ar = wr[IVX] + cr * qr;
  const T bp = ar > T(0.) ? ar : TINY_NUMBER;
  const T bm = al < T(0.) ? al : -TINY_NUMBER;

  T vxl = wl[IVX] - al;
  T vxr = wr[IVX] - ar;
  const T tl = pl + vxl * wl[IDN] * wl[IVX];
  const T tr = pr + vxr * wr[IDN] * wr[IVX];
  const T ml = wl[IDN] * vxl;
  const T mr = -(wr[IDN] * vxr);
  const T imm = T(1.) / (ml + mr);
  const T am = (tl - tr) * imm;
  T cp = (ml * tr + mr * tl) * imm;
  cp = cp > -pf ? cp : -pf;  // total contact pressure clamped at zero

  vxl = wl[IVX] - bm;
  vxr = wr[IVX] - bp;
  T fl[5], fr[5];
  fl[IDN] = wl[IDN] * vxl;
  fr[IDN] = wr[IDN] * vxr;
  fl[IVX] = wl[IDN] * wl[IVX] * vxl + pl;
  fr[IVX] = wr[IDN] * wr[IVX] * vxr + pr;
  fl[IVY] = wl[IDN] * wl[IVY] * vxl;
  fr[IVY] = wr[IDN] * wr[IVY] * vxr;
  fl[IVZ] = wl[IDN] * wl[IVZ] * vxl;
  fr[IVZ] = wr[IDN] * wr[IVZ] * vxr;
  fl[IPR] = el * vxl + wl[IPR] * wl[IVX];
  fr[IPR] = er * vxr + wr[IPR] * wr[IVX];

  T sl, sr, sm;
  if (am >= T(0.)) {
    const T i = T(1.) / (am - bm);
    sl = am * i;
    sr = T(0.);
    sm = -bm * i;
  } else {
    const T i = T(1.) / (bp - am);
    sl = T(0.);
    sr = -am * i;
    sm = bp * i;
  }
  flx[IDN] = sl * fl[IDN] + sr * fr[IDN];
  flx[IVX] = (sl * fl[IVX] + sr * fr[IVX]) + sm * cp;
  flx[IVY] = sl * fl[IVY] + sr * fr[IVY];
  flx[IVZ] = sl * fl[IVZ] + sr * fr[IVZ];
  flx[IPR] = (sl * fl[IPR] + sr * fr[IPR]) + sm * (pf + cp) * am;
}

// Roe with the LLF fallback: the jumps of density, pressure and total energy
// (E = p / (gamma - 1) + rho |v|^2 / 2) from the perturbations.
template <typename T>
__device__ __forceinline__ void roe_pert(const T* wl, const T* wr, T rl, T rr, T pl, T pr,
                                         T gammad, T* flx) {
  const T TINY_NUMBER = T(1.0e-10);
  const T gm1 = gammad - T(1.);
  const T cl = sqrt(gammad * wl[IPR] / wl[IDN]);
  const T cr = sqrt(gammad * wr[IPR] / wr[IDN]);
  const T kel = T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  const T ker = T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));
  const T etl = wl[IPR] / gm1 + kel;
  const T etr = wr[IPR] / gm1 + ker;

  const T sqrtdl = sqrt(wl[IDN]);
  const T sqrtdr = sqrt(wr[IDN]);
  const T isdlpdr = T(1.) / (sqrtdl + sqrtdr);
  const T rhobar = sqrtdl * sqrtdr;
  const T v1 = (sqrtdl * wl[IVX] + sqrtdr * wr[IVX]) * isdlpdr;
  const T v2 = (sqrtdl * wl[IVY] + sqrtdr * wr[IVY]) * isdlpdr;
  const T v3 = (sqrtdl * wl[IVZ] + sqrtdr * wr[IVZ]) * isdlpdr;
  const T hl = (etl + wl[IPR]) / wl[IDN];
  const T hr = (etr + wr[IPR]) / wr[IDN];
  const T h = (hl * sqrtdl + hr * sqrtdr) * isdlpdr;

  const T fl0 = wl[IDN] * wl[IVX];
  const T fr0 = wr[IDN] * wr[IVX];
  const T fl1 = wl[IDN] * wl[IVX] * wl[IVX] + pl;
  const T fr1 = wr[IDN] * wr[IVX] * wr[IVX] + pr;
  const T fl2 = wl[IDN] * wl[IVX] * wl[IVY];
  const T fr2 = wr[IDN] * wr[IVX] * wr[IVY];
  const T fl3 = wl[IDN] * wl[IVX] * wl[IVZ];
  const T fr3 = wr[IDN] * wr[IVX] * wr[IVZ];
  const T fl4 = (etl + wl[IPR]) * wl[IVX];
  const T fr4 = (etr + wr[IPR]) * wr[IVX];

  const T qbar0 = (wl[IDN] / sqrtdl + wr[IDN] / sqrtdr) * isdlpdr;
  const T drho = rr - rl;
  const T dp = pr - pl;
  const T du1 = fr0 - fl0;
  const T du2 = wr[IDN] * wr[IVY] - wl[IDN] * wl[IVY];
  const T du3 = wr[IDN] * wr[IVZ] - wl[IDN] * wl[IVZ];
  const T du4 = dp / gm1 + (ker - kel);

  T f[5];
  f[IDN] = T(0.5) * (fl0 + fr0);
  f[IVX] = T(0.5) * (fl1 + fr1);
  f[IVY] = T(0.5) * (fl2 + fr2);
  f[IVZ] = T(0.5) * (fl3 + fr3);
  f[IPR] = T(0.5) * (fl4 + fr4);

  const T vsq = (v1 * v1 + v2 * v2) + v3 * v3;
  const T q = h - T(0.5) * vsq;
  const T qm = q > TINY_NUMBER ? q : TINY_NUMBER;
  const T cs_sq = gm1 * qm > TINY_NUMBER ? gm1 * qm : TINY_NUMBER;
  const T cs = sqrt(cs_sq);
  const T ics = T(1.) / cs, ics2 = T(1.) / cs_sq;

  const T lam_m = v1 - cs;
  const T lam_p = v1 + cs;
  const T spd_m = fabs(lam_m);
  const T spd_0 = fabs(v1);
  const T spd_p = fabs(lam_p);

  const T du = wr[IVX] - wl[IVX];
  const T dv = wr[IVY] - wl[IVY];
  const T dw = wr[IVZ] - wl[IVZ];
  const T alpha_m = -T(0.5) * rhobar * ics * du + T(0.5) * dp * ics2;
  const T alpha_p = +T(0.5) * rhobar * ics * du + T(0.5) * dp * ics2;
  const T alpha_v = rhobar * dv;
  const T alpha_w = rhobar * dw;
  const T alpha0 = drho - dp * ics2 * qbar0;
  const T rho_first = wl[IDN] + alpha_m * qbar0;
  const T rho_second = rho_first + alpha0;
  const bool llf_flag = rho_first < T(0.) || rho_second < T(0.);

  f[IDN] -= T(0.5) * ((spd_m * alpha_m * qbar0 + spd_0 * alpha0) + spd_p * alpha_p * qbar0);
  f[IVX] -= T(0.5) * ((spd_m * alpha_m * (v1 - cs) + spd_0 * (v1 * alpha0)) +
                    spd_p * alpha_p * (v1 + cs));
  f[IVY] -= T(0.5) * ((spd_m * alpha_m * v2 + spd_0 * (v2 * alpha0 + alpha_v)) +
                    spd_p * alpha_p * v2);
  f[IVZ] -= T(0.5) * ((spd_m * alpha_m * v3 + spd_0 * (v3 * alpha0 + alpha_w)) +
                    spd_p * alpha_p * v3);
  f[IPR] -= T(0.5) * ((spd_m * alpha_m * (h - v1 * cs) +
                     spd_0 * ((rhobar * ((hr - hl) - v1 * du) + alpha0 * h) - dp)) +
                    spd_p * alpha_p * (h + v1 * cs));

  if (lam_m > T(0.)) {
    f[IDN] = fl0;
    f[IVX] = fl1;
    f[IVY] = fl2;
    f[IVZ] = fl3;
    f[IPR] = fl4;
  }
  if (lam_p < T(0.)) {
    f[IDN] = fr0;
    f[IVX] = fr1;
    f[IVY] = fr2;
    f[IVZ] = fr3;
    f[IPR] = fr4;
  }
  if (llf_flag) {
    const T x = fabs(wl[IVX]) + cl, y = fabs(wr[IVX]) + cr;
    const T a = T(0.5) * (x > y ? x : y);
    f[IDN] = T(0.5) * (fl0 + fr0) - a * drho;
    f[IVX] = T(0.5) * (fl1 + fr1) - a * du1;
    f[IVY] = T(0.5) * (fl2 + fr2) - a * du2;
    f[IVZ] = T(0.5) * (fl3 + fr3) - a * du3;
    f[IPR] = T(0.5) * (fl4 + fr4) - a * du4;
  }
#pragma unroll
  for (int n = 0; n < 5; ++n) flx[n] = f[n];
}

// dispatch on the solver (face frame, FP32)
template <typename T>
__device__ __forceinline__ void riemann_flux_pert(int solver, const T* wl, const T* wr, T rl, T rr,
                                                  T pl, T pr, T pf, T gammad, T* flx) {
  if (solver == SNAPX_RIEMANN_LMARS)
    lmars_pert(wl, wr, rl, rr, pl, pr, gammad, flx);
  else if (solver == SNAPX_RIEMANN_HLLC)
    hllc_pert(wl, wr, rl, rr, pl, pr, pf, gammad, flx);
  else
    roe_pert(wl, wr, rl, rr, pl, pr, gammad, flx);
}

}  // namespace snapx
