// Device Riemann solvers for the dry ideal gas (nvar = 5).
//
// wl / wr are primitive states (rho, v1, v2, v3, p) in registers; ivx is the
// normal-velocity slot (IPR - dim), ivy / ivz the cyclic tangential slots.
// The arithmetic order follows the reference kernels:
//   LMARS src/riemann/lmars_impl.h:16-78
//   HLLC  src/riemann/hllc_impl.h:16-130
//   Roe   src/riemann/roe_impl.h:23-211 (ideal_moist = false)
// with the ideal-gas helpers W->I = p/(gamma-1), W->A = gamma,
// WA->L = sqrt(gamma p / rho) (src/eos/ideal_gas.cpp:30-65).
#pragma once

#include "common.hpp"

namespace snapx {

template <typename T>
__device__ __forceinline__ void lmars_flux(const T* wl, const T* wr, T gammad, int ivx,
                                           int ivy, int ivz, T* flx, T* fp) {
  T hl = (wl[IPR] / (gammad - T(1))) / wl[IDN];
  T hr = (wr[IPR] / (gammad - T(1))) / wr[IDN];
  T wli[5] = {wl[IDN], wl[ivx], wl[ivy], wl[ivz], wl[IPR]};
  T wri[5] = {wr[IDN], wr[ivx], wr[ivy], wr[ivz], wr[IPR]};
  hl += T(0.5) * ((sq(wli[1]) + sq(wli[2])) + sq(wli[3])) + wli[4] / wli[0];
  hr += T(0.5) * ((sq(wri[1]) + sq(wri[2])) + sq(wri[3])) + wri[4] / wri[0];
  T rhobar = T(0.5) * (wli[0] + wri[0]);
  T gamma_bar = T(0.5) * (gammad + gammad);
  T cbar = sqrt(T(0.5) * gamma_bar * (wli[4] + wri[4]) / rhobar);
  T pbar = T(0.5) * (wli[4] + wri[4]) + T(0.5) * (rhobar * cbar) * (wli[1] - wri[1]);
  if (fp) *fp = pbar;
  T ubar = T(0.5) * (wli[1] + wri[1]) + T(0.5) / (rhobar * cbar) * (wli[4] - wri[4]);
  const T* s = ubar > T(0.0) ? wli : wri;
  T h = ubar > T(0.0) ? hl : hr;
  flx[IDN] = ubar * s[0] * T(1.0);
  flx[ivx] = ubar * s[0] * s[1] + pbar;
  flx[ivy] = ubar * s[0] * s[2];
  flx[ivz] = ubar * s[0] * s[3];
  flx[IPR] = ubar * s[0] * h;
}

template <typename T>
__device__ __forceinline__ void hllc_flux(const T* wl, const T* wr, T gammad, int ivx, int ivy,
                                          int ivz, T* flx, T* fp) {
  const T TINY_NUMBER = T(1.0e-10);
  T gammal = gammad, gammar = gammad;
  T el = wl[IPR] / (gammad - T(1));
  T er = wr[IPR] / (gammad - T(1));
  T cl = sqrt(gammal * wl[IPR] / wl[IDN]);
  T cr = sqrt(gammar * wr[IPR] / wr[IDN]);

  el += T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  er += T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));

  T rhoa = T(.5) * (wl[IDN] + wr[IDN]);
  T ca = T(.5) * (cl + cr);
  T pmid = T(.5) * ((wl[IPR] + wr[IPR]) + (wl[ivx] - wr[ivx]) * rhoa * ca);

  T ql = (pmid <= wl[IPR])
             ? T(1.0)
             : sqrt(T(1.0) + (gammal + T(1)) / (T(2) * gammal) * (pmid / wl[IPR] - T(1.0)));
  T qr = (pmid <= wr[IPR])
             ? T(1.0)
             : sqrt(T(1.0) + (gammar + T(1)) / (T(2) * gammar) * (pmid / wr[IPR] - T(1.0)));

  T al = wl[ivx] - cl * ql;
  T ar = wr[ivx] + cr * qr;
  T bp = ar > T(0.0) ? ar : TINY_NUMBER;
  T bm = al < T(0.0) ? al : -TINY_NUMBER;

  T vxl = wl[ivx] - al;
  T vxr = wr[ivx] - ar;
  T tl = wl[IPR] + vxl * wl[IDN] * wl[ivx];
  T tr = wr[IPR] + vxr * wr[IDN] * wr[ivx];
  T ml = wl[IDN] * vxl;
  T mr = -(wr[IDN] * vxr);
  T am = (tl - tr) / (ml + mr);
  T cp = (ml * tr + mr * tl) / (ml + mr);
  cp = cp > T(0.0) ? cp : T(0.0);
  if (fp) *fp = cp;

  vxl = wl[ivx] - bm;
  vxr = wr[ivx] - bp;

  T fl[5], fr[5];
  fl[IDN] = wl[IDN] * vxl * T(1.);
  fr[IDN] = wr[IDN] * vxr * T(1.);
  fl[ivx] = wl[IDN] * wl[ivx] * vxl + wl[IPR];
  fr[ivx] = wr[IDN] * wr[ivx] * vxr + wr[IPR];
  fl[ivy] = wl[IDN] * wl[ivy] * vxl;
  fr[ivy] = wr[IDN] * wr[ivy] * vxr;
  fl[ivz] = wl[IDN] * wl[ivz] * vxl;
  fr[ivz] = wr[IDN] * wr[ivz] * vxr;
  fl[IPR] = el * vxl + wl[IPR] * wl[ivx];
  fr[IPR] = er * vxr + wr[IPR] * wr[ivx];

  T sl, sr, sm;
  if (am >= T(0.0)) {
    sl = am / (am - bm);
    sr = T(0.0);
    sm = -bm / (am - bm);
  } else {
    sl = T(0.0);
    sr = -am / (bp - am);
    sm = bp / (bp - am);
  }
  flx[IDN] = sl * fl[IDN] + sr * fr[IDN];
  flx[ivx] = (sl * fl[ivx] + sr * fr[ivx]) + sm * cp;
  flx[ivy] = sl * fl[ivy] + sr * fr[ivy];
  flx[ivz] = sl * fl[ivz] + sr * fr[ivz];
  flx[IPR] = (sl * fl[IPR] + sr * fr[IPR]) + sm * cp * am;
}

template <typename T>
__device__ __forceinline__ void roe_flux(const T* wl, const T* wr, T gammad, int ivx, int ivy,
                                         int ivz, T* flx, T* fp) {
  const T TINY_NUMBER = T(1.0e-10);
  T gammal = gammad, gammar = gammad;
  T el = wl[IPR] / (gammad - T(1));
  T er = wr[IPR] / (gammad - T(1));
  T cl = sqrt(gammal * wl[IPR] / wl[IDN]);
  T cr = sqrt(gammar * wr[IPR] / wr[IDN]);

  T etl = el + T(0.5) * wl[IDN] * ((sq(wl[IVX]) + sq(wl[IVY])) + sq(wl[IVZ]));
  T etr = er + T(0.5) * wr[IDN] * ((sq(wr[IVX]) + sq(wr[IVY])) + sq(wr[IVZ]));

  T sqrtdl = sqrt(wl[IDN]);
  T sqrtdr = sqrt(wr[IDN]);
  T isdlpdr = T(1) / (sqrtdl + sqrtdr);

  T rhobar = sqrtdl * sqrtdr;
  T v1 = (sqrtdl * wl[ivx] + sqrtdr * wr[ivx]) * isdlpdr;
  T v2 = (sqrtdl * wl[ivy] + sqrtdr * wr[ivy]) * isdlpdr;
  T v3 = (sqrtdl * wl[ivz] + sqrtdr * wr[ivz]) * isdlpdr;
  T hl = (etl + wl[IPR]) / wl[IDN];
  T hr = (etr + wr[IPR]) / wr[IDN];
  T h = (hl * sqrtdl + hr * sqrtdr) * isdlpdr;

  T rhol0 = wl[IDN] * T(1);
  T rhor0 = wr[IDN] * T(1);
  T fl0 = rhol0 * wl[ivx];
  T fr0 = rhor0 * wr[ivx];
  T fl1 = wl[IDN] * wl[ivx] * wl[ivx] + wl[IPR];
  T fr1 = wr[IDN] * wr[ivx] * wr[ivx] + wr[IPR];
  T fl2 = wl[IDN] * wl[ivx] * wl[ivy];
  T fr2 = wr[IDN] * wr[ivx] * wr[ivy];
  T fl3 = wl[IDN] * wl[ivx] * wl[ivz];
  T fr3 = wr[IDN] * wr[ivx] * wr[ivz];
  T fl4 = (etl + wl[IPR]) * wl[ivx];
  T fr4 = (etr + wr[IPR]) * wr[ivx];

  T qbar0 = (rhol0 / sqrtdl + rhor0 / sqrtdr) * isdlpdr;
  T gamma_roe = T(0.5) * (gammal + gammar);

  T du1 = wr[IDN] * wr[ivx] - wl[IDN] * wl[ivx];
  T du2 = wr[IDN] * wr[ivy] - wl[IDN] * wl[ivy];
  T du3 = wr[IDN] * wr[ivz] - wl[IDN] * wl[ivz];
  T du4 = etr - etl;

  T f[5];
  f[IDN] = T(0.5) * (fl0 + fr0);
  f[ivx] = T(0.5) * (fl1 + fr1);
  f[ivy] = T(0.5) * (fl2 + fr2);
  f[ivz] = T(0.5) * (fl3 + fr3);
  f[IPR] = T(0.5) * (fl4 + fr4);

  T vsq = (v1 * v1 + v2 * v2) + v3 * v3;
  T gm1_roe = gamma_roe - T(1);
  T q = (h - T(0.5) * vsq) - T(0);
  T qm = q > TINY_NUMBER ? q : TINY_NUMBER;
  T cs_sq = gm1_roe * qm > TINY_NUMBER ? gm1_roe * qm : TINY_NUMBER;
  T cs = sqrt(cs_sq);
  if (fp) *fp = T(0.5) * ((wl[IPR] + wr[IPR]) + rhobar * cs * (wl[ivx] - wr[ivx]));

  T lam_m = v1 - cs;
  T lam_0 = v1;
  T lam_p = v1 + cs;
  T spd_m = fabs(lam_m);
  T spd_0 = fabs(lam_0);
  T spd_p = fabs(lam_p);

  bool llf_flag = false;
  T du = wr[ivx] - wl[ivx];
  T dv = wr[ivy] - wl[ivy];
  T dw = wr[ivz] - wl[ivz];
  T dp = wr[IPR] - wl[IPR];

  T alpha_m = -T(0.5) * rhobar / cs * du + T(0.5) * dp / cs_sq;
  T alpha_p = +T(0.5) * rhobar / cs * du + T(0.5) * dp / cs_sq;
  T alpha_v = rhobar * dv;
  T alpha_w = rhobar * dw;

  T alpha0 = T(0);
  T alpha_comp0 = (rhor0 - rhol0) - dp / cs_sq * qbar0;
  T rho_first = rhol0 + alpha_m * qbar0;
  if (rho_first < T(0)) llf_flag = true;
  T rho_second = rho_first + alpha_comp0;
  if (rho_second < T(0)) llf_flag = true;
  alpha0 += alpha_comp0;
  f[IDN] -= T(0.5) * ((spd_m * alpha_m * qbar0 + spd_0 * alpha_comp0) + spd_p * alpha_p * qbar0);

  f[ivx] -= T(0.5) * ((spd_m * alpha_m * (v1 - cs) + spd_0 * (v1 * alpha0)) +
                      spd_p * alpha_p * (v1 + cs));
  f[ivy] -= T(0.5) * ((spd_m * alpha_m * v2 + spd_0 * (v2 * alpha0 + alpha_v)) +
                      spd_p * alpha_p * v2);
  f[ivz] -= T(0.5) * ((spd_m * alpha_m * v3 + spd_0 * (v3 * alpha0 + alpha_w)) +
                      spd_p * alpha_p * v3);
  f[IPR] -= T(0.5) * ((spd_m * alpha_m * (h - v1 * cs) +
                       spd_0 * ((rhobar * ((hr - hl) - v1 * du) + alpha0 * h) - dp)) +
                      spd_p * alpha_p * (h + v1 * cs));

  if (lam_m > T(0)) {
    f[IDN] = fl0;
    f[ivx] = fl1;
    f[ivy] = fl2;
    f[ivz] = fl3;
    f[IPR] = fl4;
  }
  if (lam_p < T(0)) {
    f[IDN] = fr0;
    f[ivx] = fr1;
    f[ivy] = fr2;
    f[ivz] = fr3;
    f[IPR] = fr4;
  }
  if (llf_flag) {
    T x = fabs(wl[ivx]) + cl, y = fabs(wr[ivx]) + cr;
    T a = T(0.5) * (x > y ? x : y);
    f[IDN] = T(0.5) * (fl0 + fr0) - a * (rhor0 - rhol0);
    f[ivx] = T(0.5) * (fl1 + fr1) - a * du1;
    f[ivy] = T(0.5) * (fl2 + fr2) - a * du2;
    f[ivz] = T(0.5) * (fl3 + fr3) - a * du3;
    f[IPR] = T(0.5) * (fl4 + fr4) - a * du4;
  }
  for (int n = 0; n < 5; ++n) flx[n] = f[n];
}

template <typename T>
__device__ __forceinline__ void riemann_flux(int solver, const T* wl, const T* wr, T gammad,
                                             int dim, T* flx, T* fp) {
  int ivx = IPR - dim;
  int ivy = IVX + ((ivx - IVX) + 1) % 3;
  int ivz = IVX + ((ivx - IVX) + 2) % 3;
  if (solver == SNAPX_RIEMANN_LMARS)
    lmars_flux(wl, wr, gammad, ivx, ivy, ivz, flx, fp);
  else if (solver == SNAPX_RIEMANN_HLLC)
    hllc_flux(wl, wr, gammad, ivx, ivy, ivz, flx, fp);
  else
    roe_flux(wl, wr, gammad, ivx, ivy, ivz, flx, fp);
}

}  // namespace snapx
