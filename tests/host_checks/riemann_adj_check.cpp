// Host build of the device Riemann headers: the hand-written reverse sweeps
// (riemann_adj.hpp) against forward-mode duals (dual.hpp) on the same solver
// code (riemann_impl.hpp), over random states that reach every branch.
//
// Usage: riemann_adj_check <solver> <cases>   (solver: SNAPX_RIEMANN_* code)
// Prints the worst relative difference and the branch counts; exits 1 when
// the worst difference exceeds 1e-10.
#include <hip/hip_runtime.h>
// the headers' functions become plain host functions
#undef __device__
#define __device__
#undef __forceinline__
#define __forceinline__ inline

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

using std::fabs;
using std::sqrt;

#include "riemann_impl.hpp"
#include "dual.hpp"
#include "riemann_adj.hpp"

using namespace snapx;

int main(int argc, char** argv) {
  const int solver = argc > 1 ? atoi(argv[1]) : SNAPX_RIEMANN_ROE;
  const int cases = argc > 2 ? atoi(argv[2]) : 100000;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0., 1.);
  const double gam = 1.4;
  double worst = 0.;
  // Roe branches: LLF fallback, left flux, right flux, Roe flux;
  // LMARS: upwind side of the interface velocity
  long branch[4] = {0, 0, 0, 0};
  for (int t = 0; t < cases; ++t) {
    double wl[5], wr[5], lf[5];
    int kind = t % 4;  // subsonic, supersonic, strong density jump, strong pressure jump
    double vs = kind == 0 ? 30. : (kind == 1 ? 600. : (kind == 2 ? 2000. : 5.));
    wl[IDN] = 0.2 + 2. * U(rng);
    wr[IDN] = kind == 2 ? 1e-3 + 0.01 * U(rng) : 0.2 + 2. * U(rng);
    for (int v = IVX; v <= IVZ; ++v) {
      wl[v] = vs * (2. * U(rng) - 1.);
      wr[v] = vs * (2. * U(rng) - 1.);
    }
    wl[IPR] = 1e4 + 1e5 * U(rng);
    wr[IPR] = kind == 3 ? 1e2 + 1e3 * U(rng) : 1e4 + 1e5 * U(rng);
    for (int v = 0; v < 5; ++v) lf[v] = 2. * U(rng) - 1.;

    // (d flx / d w)^T lf column by column with one-direction duals
    double ref_l[5], ref_r[5];
    for (int p = 0; p < 10; ++p) {
      Dual<1> sl[5], sr[5], f[5];
      for (int v = 0; v < 5; ++v) {
        sl[v] = Dual<1>(wl[v]);
        sr[v] = Dual<1>(wr[v]);
        sl[v].d[0] = p == v ? 1. : 0.;
        sr[v].d[0] = p == 5 + v ? 1. : 0.;
      }
      riemann_flux<Dual<1>>(solver, sl, sr, Dual<1>(gam), 3, f, (Dual<1>*)nullptr);
      double acc = 0.;
      for (int m = 0; m < 5; ++m) acc += lf[m] * f[m].d[0];
      (p < 5 ? ref_l[p] : ref_r[p - 5]) = acc;
    }
    double gl[5] = {0., 0., 0., 0., 0.}, gr[5] = {0., 0., 0., 0., 0.};
    if (solver == SNAPX_RIEMANN_ROE)
      roe_flux_adj(wl, wr, gam, lf, gl, gr);
    else
      lmars_flux_adj(wl, wr, gam, lf, gl, gr);
    double scale = 0., err = 0.;
    for (int v = 0; v < 5; ++v) {
      scale = fmax(scale, fmax(fabs(ref_l[v]), fabs(ref_r[v])));
      err = fmax(err, fmax(fabs(gl[v] - ref_l[v]), fabs(gr[v] - ref_r[v])));
    }
    worst = fmax(worst, err / scale);

    // which branch the values took: the one-sided fluxes are the solver's
    // flux between equal states; the LLF test restates roe_impl.h's density check
    double f[5], fl[5], fr[5];
    riemann_flux<double>(solver, wl, wr, gam, 3, f, nullptr);
    riemann_flux<double>(solver, wl, wl, gam, 3, fl, nullptr);
    riemann_flux<double>(solver, wr, wr, gam, 3, fr, nullptr);
    bool isl = true, isr = true;
    for (int v = 0; v < 5; ++v) {
      isl = isl && f[v] == fl[v];
      isr = isr && f[v] == fr[v];
    }
    bool llf = false;
    if (solver == SNAPX_RIEMANN_ROE) {
      double sl = sqrt(wl[IDN]), sr = sqrt(wr[IDN]), isd = 1. / (sl + sr);
      double rb = sl * sr;
      double v1 = (sl * wl[IVX] + sr * wr[IVX]) * isd, v2 = (sl * wl[IVY] + sr * wr[IVY]) * isd,
             v3 = (sl * wl[IVZ] + sr * wr[IVZ]) * isd;
      double etl = wl[IPR] / (gam - 1.) +
                   0.5 * wl[IDN] * (wl[IVX] * wl[IVX] + wl[IVY] * wl[IVY] + wl[IVZ] * wl[IVZ]);
      double etr = wr[IPR] / (gam - 1.) +
                   0.5 * wr[IDN] * (wr[IVX] * wr[IVX] + wr[IVY] * wr[IVY] + wr[IVZ] * wr[IVZ]);
      double h = ((etl + wl[IPR]) / wl[IDN] * sl + (etr + wr[IPR]) / wr[IDN] * sr) * isd;
      double q = h - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3);
      double cs2 = fmax((gam - 1.) * fmax(q, 1e-10), 1e-10), cs = sqrt(cs2);
      double dp = wr[IPR] - wl[IPR], du = wr[IVX] - wl[IVX];
      double am = -0.5 * rb / cs * du + 0.5 * dp / cs2;
      double qb = (wl[IDN] / sl + wr[IDN] / sr) * isd;
      double ac0 = (wr[IDN] - wl[IDN]) - dp / cs2 * qb;
      llf = wl[IDN] + am * qb < 0. || wl[IDN] + am * qb + ac0 < 0.;
    }
    branch[llf ? 0 : (isl ? 1 : (isr ? 2 : 3))] += 1;
  }
  printf("worst %.3e llf %ld left %ld right %ld mixed %ld\n", worst, branch[0], branch[1],
         branch[2], branch[3]);
  return worst < 1e-10 ? 0 : 1;
}
