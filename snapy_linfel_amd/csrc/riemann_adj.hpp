// Reverse-mode (transposed Jacobian) products of the Riemann fluxes.
//
// lmars_flux_adj / roe_flux_adj give lwl += (d flx / d wl)^T lf and
// lwr += (d flx / d wr)^T lf for lmars_flux / roe_flux in the x1 frame (normal
// velocity in slot IVX): one forward sweep that keeps the intermediates and
// one reverse sweep through them.  They differentiate the branch that the
// values select, as forward-mode duals on the solvers do: for LMARS the
// upwind side of the interface velocity (src/riemann/lmars_impl.h:16-78); for
// Roe the LLF fallback when an intermediate density of the Roe solution is
// negative, else the one-sided flux when both acoustic speeds share a sign,
// else the Roe flux (src/riemann/roe_impl.h:23-211).  |x| has derivative
// sign(x) with sign(0) = +1, the dual-number convention.
#pragma once

#include "common.hpp"

namespace snapx {

__device__ __forceinline__ double sgn1(double x) { return x < 0. ? -1. : 1.; }

// adjoint of the physical flux F(w) = (rho u, rho u u + p, rho u v, rho u w,
// (E + p) u) with E = p / (gamma - 1) + rho |v|^2 / 2, added into gw
__device__ __forceinline__ void euler_flux_adj(const double* w, double gm1, const double* g,
                                               double* gw) {
  double r = w[IDN], u = w[IVX], v = w[IVY], z = w[IVZ], p = w[IPR];
  double et = p / gm1 + 0.5 * r * ((u * u + v * v) + z * z);
  double g_et = u * g[IPR];
  gw[IDN] += u * g[IDN] + u * u * g[IVX] + u * v * g[IVY] + u * z * g[IVZ] +
             0.5 * ((u * u + v * v) + z * z) * g_et;
  gw[IVX] += r * g[IDN] + 2. * r * u * g[IVX] + r * v * g[IVY] + r * z * g[IVZ] +
             (et + p) * g[IPR] + r * u * g_et;
  gw[IVY] += r * u * g[IVY] + r * v * g_et;
  gw[IVZ] += r * u * g[IVZ] + r * z * g_et;
  gw[IPR] += g[IVX] + u * g[IPR] + g_et / gm1;
}

// adjoint of the conserved state U(w) = (rho, rho u, rho v, rho w, E), added into gw
__device__ __forceinline__ void cons_adj(const double* w, double gm1, const double* g,
                                         double* gw) {
  double r = w[IDN], u = w[IVX], v = w[IVY], z = w[IVZ];
  gw[IDN] += g[IDN] + u * g[IVX] + v * g[IVY] + z * g[IVZ] +
             0.5 * ((u * u + v * v) + z * z) * g[IPR];
  gw[IVX] += r * g[IVX] + r * u * g[IPR];
  gw[IVY] += r * g[IVY] + r * v * g[IPR];
  gw[IVZ] += r * g[IVZ] + r * z * g[IPR];
  gw[IPR] += g[IPR] / gm1;
}

__device__ __forceinline__ void lmars_flux_adj(const double* wl, const double* wr,
                                               double gammad, const double* lf, double* lwl,
                                               double* lwr) {
  const double gm1 = gammad - 1.;
  const double rl = wl[IDN], ul = wl[IVX], pl = wl[IPR];
  const double rr = wr[IDN], ur = wr[IVX], pr = wr[IPR];
  double rb = 0.5 * (rl + rr);
  double cbar = ::sqrt(0.5 * gammad * (pl + pr) / rb);
  double m = rb * cbar;
  double ubar = 0.5 * (ul + ur) + 0.5 / m * (pl - pr);
  bool left = ubar > 0.;
  const double* s = left ? wl : wr;
  double* gs = left ? lwl : lwr;
  double r = s[IDN], p = s[IPR];
  double h = (p / gm1) / r + (0.5 * ((s[IVX] * s[IVX] + s[IVY] * s[IVY]) + s[IVZ] * s[IVZ]) + p / r);
  // flx = ubar r (1, s1, s2, s3, h) + (0, pbar, 0, 0, 0)
  double g_ubar = r * (lf[IDN] + lf[IVX] * s[IVX] + lf[IVY] * s[IVY] + lf[IVZ] * s[IVZ] +
                       lf[IPR] * h);
  double ur_ = ubar * r;
  double g_h = ur_ * lf[IPR];
  gs[IDN] += ubar * (lf[IDN] + lf[IVX] * s[IVX] + lf[IVY] * s[IVY] + lf[IVZ] * s[IVZ] +
                     lf[IPR] * h) -
             g_h * (p / gm1 + p) / (r * r);
  gs[IVX] += ur_ * lf[IVX] + g_h * s[IVX];
  gs[IVY] += ur_ * lf[IVY] + g_h * s[IVY];
  gs[IVZ] += ur_ * lf[IVZ] + g_h * s[IVZ];
  gs[IPR] += g_h * (1. / gm1 + 1.) / r;
  // pbar = (pl + pr) / 2 + m (ul - ur) / 2, ubar = (ul + ur) / 2 + (pl - pr) / (2 m)
  double g_pbar = lf[IVX];
  double g_m = 0.5 * g_pbar * (ul - ur) - 0.5 * g_ubar * (pl - pr) / (m * m);
  lwl[IVX] += 0.5 * g_ubar + 0.5 * m * g_pbar;
  lwr[IVX] += 0.5 * g_ubar - 0.5 * m * g_pbar;
  lwl[IPR] += 0.5 * g_pbar + 0.5 * g_ubar / m;
  lwr[IPR] += 0.5 * g_pbar - 0.5 * g_ubar / m;
  // m = rb cbar, cbar = sqrt(gamma (pl + pr) / (2 rb))
  double g_cbar = rb * g_m;
  double g_rb = cbar * g_m - g_cbar * cbar / (2. * rb);
  double g_psum = g_cbar * gammad / (4. * rb * cbar);
  lwl[IPR] += g_psum;
  lwr[IPR] += g_psum;
  lwl[IDN] += 0.5 * g_rb;
  lwr[IDN] += 0.5 * g_rb;
}

__device__ __forceinline__ void roe_flux_adj(const double* wl, const double* wr, double gammad,
                                             const double* lf, double* lwl, double* lwr) {
  const double TINY_NUMBER = 1.0e-10;
  const double gm1 = gammad - 1.;
  const double rl = wl[IDN], ul = wl[IVX], vl = wl[IVY], zl = wl[IVZ], pl = wl[IPR];
  const double rr = wr[IDN], ur = wr[IVX], vr = wr[IVY], zr = wr[IVZ], pr = wr[IPR];

  // ---- forward sweep (same arithmetic as roe_flux) ----
  double etl = pl / gm1 + 0.5 * rl * ((ul * ul + vl * vl) + zl * zl);
  double etr = pr / gm1 + 0.5 * rr * ((ur * ur + vr * vr) + zr * zr);
  double cl = ::sqrt(gammad * pl / rl);
  double cr = ::sqrt(gammad * pr / rr);
  double sl = ::sqrt(rl), sr = ::sqrt(rr);
  double isd = 1. / (sl + sr);
  double rhobar = sl * sr;
  double v1 = (sl * ul + sr * ur) * isd;
  double v2 = (sl * vl + sr * vr) * isd;
  double v3 = (sl * zl + sr * zr) * isd;
  double hl = (etl + pl) / rl;
  double hr = (etr + pr) / rr;
  double h = (hl * sl + hr * sr) * isd;
  double qb = (rl / sl + rr / sr) * isd;
  double vsq = (v1 * v1 + v2 * v2) + v3 * v3;
  double q = h - 0.5 * vsq;
  bool q_on = q > TINY_NUMBER;
  double qm = q_on ? q : TINY_NUMBER;
  bool c_on = gm1 * qm > TINY_NUMBER;
  double cs_sq = c_on ? gm1 * qm : TINY_NUMBER;
  double cs = ::sqrt(cs_sq);
  double lam_m = v1 - cs, lam_p = v1 + cs;
  double sm = ::fabs(lam_m), s0 = ::fabs(v1), sp = ::fabs(lam_p);
  double du = ur - ul, dv = vr - vl, dw = zr - zl, dp = pr - pl;
  double am = -0.5 * rhobar / cs * du + 0.5 * dp / cs_sq;
  double ap = 0.5 * rhobar / cs * du + 0.5 * dp / cs_sq;
  double av = rhobar * dv, aw = rhobar * dw;
  double ac0 = (rr - rl) - dp / cs_sq * qb;
  double rho_first = rl + am * qb;
  bool llf = rho_first < 0. || rho_first + ac0 < 0.;

  double gfl[5], gfr[5];
  if (llf) {
    // f = (F(wl) + F(wr)) / 2 - a (U(wr) - U(wl)), a = max(|ul| + cl, |ur| + cr) / 2
    double x = ::fabs(ul) + cl, y = ::fabs(ur) + cr;
    double a = 0.5 * (x > y ? x : y);
    double dU[5] = {rr - rl, rr * ur - rl * ul, rr * vr - rl * vl, rr * zr - rl * zl, etr - etl};
    double g_a = 0.;
    double gU[5], mgU[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      gfl[k] = gfr[k] = 0.5 * lf[k];
      g_a -= lf[k] * dU[k];
      gU[k] = -a * lf[k];
      mgU[k] = a * lf[k];
    }
    cons_adj(wr, gm1, gU, lwr);
    cons_adj(wl, gm1, mgU, lwl);
    double g_c = 0.5 * g_a;
    if (x > y) {  // d cl / d pl = gamma / (2 cl rl), d cl / d rl = -cl / (2 rl)
      lwl[IVX] += sgn1(ul) * g_c;
      lwl[IPR] += g_c * gammad / (2. * cl * rl);
      lwl[IDN] -= g_c * cl / (2. * rl);
    } else {
      lwr[IVX] += sgn1(ur) * g_c;
      lwr[IPR] += g_c * gammad / (2. * cr * rr);
      lwr[IDN] -= g_c * cr / (2. * rr);
    }
    euler_flux_adj(wl, gm1, gfl, lwl);
    euler_flux_adj(wr, gm1, gfr, lwr);
    return;
  }
  if (lam_p < 0.) {
    euler_flux_adj(wr, gm1, lf, lwr);
    return;
  }
  if (lam_m > 0.) {
    euler_flux_adj(wl, gm1, lf, lwl);
    return;
  }

  // ---- reverse sweep of f = (F(wl) + F(wr)) / 2 - D / 2 ----
#pragma unroll
  for (int k = 0; k < 5; ++k) gfl[k] = gfr[k] = 0.5 * lf[k];
  const double g0 = -0.5 * lf[IDN], g1 = -0.5 * lf[IVX], g2 = -0.5 * lf[IVY],
               g3 = -0.5 * lf[IVZ], g4 = -0.5 * lf[IPR];
  double hm = h - v1 * cs, hp = h + v1 * cs;
  double e4 = (rhobar * ((hr - hl) - v1 * du) + ac0 * h) - dp;
  double g_sm = am * (g0 * qb + g1 * lam_m + g2 * v2 + g3 * v3 + g4 * hm);
  double g_sp = ap * (g0 * qb + g1 * lam_p + g2 * v2 + g3 * v3 + g4 * hp);
  double g_s0 = g0 * ac0 + g1 * v1 * ac0 + g2 * (v2 * ac0 + av) + g3 * (v3 * ac0 + aw) + g4 * e4;
  double g_am = sm * (g0 * qb + g1 * lam_m + g2 * v2 + g3 * v3 + g4 * hm);
  double g_ap = sp * (g0 * qb + g1 * lam_p + g2 * v2 + g3 * v3 + g4 * hp);
  double g_ac0 = s0 * (g0 + g1 * v1 + g2 * v2 + g3 * v3 + g4 * h);
  double g_av = s0 * g2, g_aw = s0 * g3;
  double wsum = (sm * am + s0 * ac0) + sp * ap;
  double g_qb = g0 * (sm * am + sp * ap);
  double g_v1 = g1 * wsum + g4 * ((sp * ap - sm * am) * cs - s0 * rhobar * du);
  double g_cs = (g1 + g4 * v1) * (sp * ap - sm * am);
  double g_v2 = g2 * wsum, g_v3 = g3 * wsum;
  double g_h = g4 * wsum;
  double g_rb = g4 * s0 * ((hr - hl) - v1 * du);
  double g_hr = g4 * s0 * rhobar, g_hl = -g_hr;
  double g_du = -g4 * s0 * rhobar * v1;
  double g_dp = -g4 * s0;
  double g_dv = 0., g_dw = 0.;
  // wave speeds
  g_v1 += sgn1(lam_m) * g_sm + sgn1(v1) * g_s0 + sgn1(lam_p) * g_sp;
  g_cs += sgn1(lam_p) * g_sp - sgn1(lam_m) * g_sm;
  // wave strengths
  double ics = 1. / cs, ics2 = 1. / cs_sq;
  double g_cs_sq = g_ac0 * dp * qb * ics2 * ics2;
  g_dp -= g_ac0 * qb * ics2;
  g_qb -= g_ac0 * dp * ics2;
  double g_rr = g_ac0, g_rl = -g_ac0;
  double gapm = g_ap - g_am;
  g_rb += 0.5 * du * ics * gapm;
  g_cs -= 0.5 * rhobar * du * ics * ics * gapm;
  g_du += 0.5 * rhobar * ics * gapm;
  g_dp += 0.5 * ics2 * (g_am + g_ap);
  g_cs_sq -= 0.5 * dp * ics2 * ics2 * (g_am + g_ap);
  g_rb += dv * g_av + dw * g_aw;
  g_dv += rhobar * g_av;
  g_dw += rhobar * g_aw;
  // sound speed of the Roe state
  g_cs_sq += 0.5 * g_cs * ics;
  double g_q = (c_on && q_on) ? gm1 * g_cs_sq : 0.;
  g_h += g_q;
  g_v1 -= v1 * g_q;
  g_v2 -= v2 * g_q;
  g_v3 -= v3 * g_q;
  // Roe averages
  double g_sl = hl * isd * g_h + ul * isd * g_v1 + vl * isd * g_v2 + zl * isd * g_v3 -
                g_qb * isd * rl / (sl * sl) + sr * g_rb;
  double g_sr = hr * isd * g_h + ur * isd * g_v1 + vr * isd * g_v2 + zr * isd * g_v3 -
                g_qb * isd * rr / (sr * sr) + sl * g_rb;
  double g_isd = (hl * sl + hr * sr) * g_h + (sl * ul + sr * ur) * g_v1 +
                 (sl * vl + sr * vr) * g_v2 + (sl * zl + sr * zr) * g_v3 +
                 (rl / sl + rr / sr) * g_qb;
  g_hl += sl * isd * g_h;
  g_hr += sr * isd * g_h;
  g_rl += g_qb * isd / sl;
  g_rr += g_qb * isd / sr;
  g_sl -= isd * isd * g_isd;
  g_sr -= isd * isd * g_isd;
  g_rl += 0.5 * g_sl / sl;
  g_rr += 0.5 * g_sr / sr;
  // enthalpies h = (E + p) / rho
  double g_etl = g_hl / rl, g_etr = g_hr / rr;
  g_rl -= hl / rl * g_hl;
  g_rr -= hr / rr * g_hr;
  lwl[IDN] += g_rl + 0.5 * ((ul * ul + vl * vl) + zl * zl) * g_etl;
  lwr[IDN] += g_rr + 0.5 * ((ur * ur + vr * vr) + zr * zr) * g_etr;
  lwl[IVX] += sl * isd * g_v1 - g_du + rl * ul * g_etl;
  lwr[IVX] += sr * isd * g_v1 + g_du + rr * ur * g_etr;
  lwl[IVY] += sl * isd * g_v2 - g_dv + rl * vl * g_etl;
  lwr[IVY] += sr * isd * g_v2 + g_dv + rr * vr * g_etr;
  lwl[IVZ] += sl * isd * g_v3 - g_dw + rl * zl * g_etl;
  lwr[IVZ] += sr * isd * g_v3 + g_dw + rr * zr * g_etr;
  lwl[IPR] += g_hl / rl - g_dp + g_etl / gm1;
  lwr[IPR] += g_hr / rr + g_dp + g_etr / gm1;
  euler_flux_adj(wl, gm1, gfl, lwl);
  euler_flux_adj(wr, gm1, gfr, lwr);
}

}  // namespace snapx
