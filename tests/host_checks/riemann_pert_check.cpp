// Host build of the device Riemann headers: the perturbation-form solvers
// (riemann_pert.hpp) against the total-state solvers (riemann_impl.hpp).
//
// In FP64 the perturbation form plus the background flux (0, p_f, 0, 0, 0)
// must equal the total-state flux to rounding, for random states of every
// regime and an arbitrary face reference.  In FP32, on states a small
// perturbation away from a hydrostatic background at rest (p about 1e5), the
// perturbation form must reproduce the FP64 flux perturbation far better than
// the total-state FP32 solver does.
//
// Usage: riemann_pert_check <solver> <cases>   (solver: SNAPX_RIEMANN_* code)
// Prints "fp64 <worst rel diff> fp32_pert <err> fp32_total <err>"; exits 1
// when the FP64 difference exceeds 1e-11 or the FP32 perturbation form is not
// at least 100 times closer than the total form.
#include <hip/hip_runtime.h>
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
#include "riemann_pert.hpp"

using namespace snapx;

int main(int argc, char** argv) {
  const int solver = argc > 1 ? atoi(argv[1]) : SNAPX_RIEMANN_ROE;
  const int cases = argc > 2 ? atoi(argv[2]) : 20000;
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(0., 1.);
  const double gam = 1.4;
  double worst = 0.;
  for (int t = 0; t < cases; ++t) {
    // four regimes as in the branch tests: subsonic, supersonic either way,
    // strong rarefaction
    const int kind = t % 4;
    double wl[5], wr[5];
    wl[IDN] = 0.5 + 1.5 * U(rng);
    wl[IPR] = 0.5 + 1.5 * U(rng);
    for (int v = IVX; v <= IVZ; ++v) wl[v] = (U(rng) - 0.5);
    for (int v = 0; v < 5; ++v) wr[v] = wl[v] * (1.0 + 0.2 * (U(rng) - 0.5));
    if (kind == 1) wl[IVX] += 4.0, wr[IVX] += 4.0;
    if (kind == 2) wl[IVX] -= 4.0, wr[IVX] -= 4.0;
    if (kind == 3) wl[IVX] -= 4.0, wr[IVX] += 4.0;
    const double pf = 0.5 * (wl[IPR] + wr[IPR]) * (0.8 + 0.4 * U(rng));
    const double rf = 0.5 * (wl[IDN] + wr[IDN]);
    double ref[5], got[5];
    riemann_flux<double>(solver, wl, wr, gam, 3, ref, (double*)nullptr);
    riemann_flux_pert<double>(solver, wl, wr, wl[IDN] - rf, wr[IDN] - rf, wl[IPR] - pf,
                              wr[IPR] - pf, pf, gam, got);
    got[IVX] += pf;
    double scale = 0.;
    for (int v = 0; v < 5; ++v) scale = fmax(scale, fabs(ref[v]));
    for (int v = 0; v < 5; ++v) worst = fmax(worst, fabs(got[v] - ref[v]) / scale);
  }

  // FP32 near a hydrostatic background: p = pbar + p', rho = rhobar + rho'
  // with pbar = 1e5 Pa, rhobar = 1.2 kg/m^3, perturbations of 1 Pa and 1e-5
  // kg/m^3, velocities of 1e-2 m/s.  Error of each flux component (normal
  // momentum without p_f) against FP64, relative to its largest FP64 size.
  double e_pert[5] = {0., 0., 0., 0., 0.}, e_tot[5] = {0., 0., 0., 0., 0.};
  double mag[5] = {0., 0., 0., 0., 0.};
  for (int t = 0; t < cases; ++t) {
    const double pb = 1e5, rb = 1.2;
    double wl[5], wr[5], pl, pr, rl, rr;
    rl = 1e-5 * (U(rng) - 0.5), rr = 1e-5 * (U(rng) - 0.5);
    pl = (U(rng) - 0.5), pr = (U(rng) - 0.5);
    wl[IDN] = rb + rl, wr[IDN] = rb + rr;
    wl[IPR] = pb + pl, wr[IPR] = pb + pr;
    for (int v = IVX; v <= IVZ; ++v) {
      wl[v] = 1e-2 * (U(rng) - 0.5);
      wr[v] = 1e-2 * (U(rng) - 0.5);
    }
    double ref[5];
    riemann_flux_pert<double>(solver, wl, wr, rl, rr, pl, pr, pb, gam, ref);
    float wlf[5], wrf[5], fp[5], ft[5];
    for (int v = 0; v < 5; ++v) wlf[v] = (float)wl[v], wrf[v] = (float)wr[v];
    riemann_flux_pert<float>(solver, wlf, wrf, (float)rl, (float)rr, (float)pl, (float)pr,
                             (float)pb, (float)gam, fp);
    riemann_flux<float>(solver, wlf, wrf, (float)gam, 3, ft, (float*)nullptr);
    const double tot_mom = (double)ft[IVX] - pb;  // total form: subtract afterwards
    for (int v = 0; v < 5; ++v) {
      mag[v] = fmax(mag[v], fabs(ref[v]));
      e_pert[v] = fmax(e_pert[v], fabs((double)fp[v] - ref[v]));
      e_tot[v] = fmax(e_tot[v], fabs((v == IVX ? tot_mom : (double)ft[v]) - ref[v]));
    }
  }
  // worst component, each against the largest size of its flux perturbation
  double err_pert = 0., err_tot = 0.;
  for (int v = 0; v < 5; ++v) {
    err_pert = fmax(err_pert, e_pert[v] / mag[v]);
    err_tot = fmax(err_tot, e_tot[v] / mag[v]);
  }
  printf("fp64 %.3e fp32_pert %.3e fp32_total %.3e\n", worst, err_pert, err_tot);
  return (worst < 1e-11 && err_pert * 100. < err_tot) ? 0 : 1;
}
