// Library-wide entry points of libsnapx: error reporting, version, config
// validation.
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "interp_impl.hpp"

namespace snapx {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int validate_config(const snapx_config* c) {
  SNAPX_CHECK_ARG(c != nullptr, "config is null");
  SNAPX_CHECK_ARG(c->nx1 >= 1 && c->nx2 >= 1 && c->nx3 >= 1, "cell counts must be >= 1");
  SNAPX_CHECK_ARG(c->nghost >= 1 && c->nghost <= 8, "nghost must be in [1, 8]");
  // src/coord/coordinate.cpp:146-160
  SNAPX_CHECK_ARG(c->nx1 == 1 || c->nx1 >= c->nghost,
                  "Number of x1 grids must be greater than the ghost zone size");
  SNAPX_CHECK_ARG(c->nx2 == 1 || c->nx2 >= c->nghost,
                  "Number of x2 grids must be greater than the ghost zone size");
  SNAPX_CHECK_ARG(c->nx3 == 1 || c->nx3 >= c->nghost,
                  "Number of x3 grids must be greater than the ghost zone size");
  SNAPX_CHECK_ARG(c->gammad > 1.0, "gammad must be > 1");
  for (int r = 0; r < 2; ++r) {
    SNAPX_CHECK_ARG(c->recon_type[r] >= SNAPX_INTERP_DC && c->recon_type[r] <= SNAPX_INTERP_WENO5,
                    "Interp: unknown type");
    SNAPX_CHECK_ARG(interp_stencils(c->recon_type[r]) / 2 + 1 <= c->nghost,
                    "reconstruction stencil needs more ghost cells");
  }
  for (int d = 0; d < 3; ++d)
    SNAPX_CHECK_ARG(c->riemann[d] >= SNAPX_RIEMANN_ROE && c->riemann[d] <= SNAPX_RIEMANN_HLLC,
                    "RiemannSolver: unknown type");
  SNAPX_CHECK_ARG(c->implicit_scheme == 0 || c->implicit_scheme == 1 || c->implicit_scheme == 9,
                  "Unsupported implicit scheme");
  SNAPX_CHECK_ARG(c->non_hydrostatic >= 0. && c->non_hydrostatic <= 1.,
                  "non-hydrostatic must be in [0, 1]");
  for (int b = 0; b < 6; ++b)
    SNAPX_CHECK_ARG(c->bc[b] >= SNAPX_BC_NONE && c->bc[b] <= SNAPX_BC_CUSTOM,
                    "Boundary function is not defined");
  SNAPX_CHECK_ARG(c->coord == SNAPX_COORD_CARTESIAN || c->coord == SNAPX_COORD_SPHERICAL_POLAR ||
                      c->coord == SNAPX_COORD_GNOMONIC_EQUIANGLE,
                  "Unknown coordinate type");
  if (c->coord == SNAPX_COORD_GNOMONIC_EQUIANGLE) {
    // src/coord/gnomonic_equiangle.cpp:21-23
    SNAPX_CHECK_ARG(c->nx2 == c->nx3, "GnomonicEquiangleImpl::reset(): nx2 must equal nx3");
    SNAPX_CHECK_ARG(c->nx1 > 1 && c->nx2 > 1,
                    "gnomonic-equiangle panels need active x1, x2 and x3 axes");
    const double hw = 0.5 * M_PI;
    const double d2 = (c->x2max - c->x2min) / c->nx2, d3 = (c->x3max - c->x3min) / c->nx3;
    SNAPX_CHECK_ARG(c->x2min - c->nghost * d2 > -hw && c->x2max + c->nghost * d2 < hw &&
                        c->x3min - c->nghost * d3 > -hw && c->x3max + c->nghost * d3 < hw,
                    "gnomonic-equiangle panel coordinates (ghosts included) must lie in "
                    "(-pi/2, pi/2)");
    SNAPX_CHECK_ARG(!c->limiter, "gnomonic-equiangle panels: the EOS limiter is not supported");
  }
  if (c->coord == SNAPX_COORD_SPHERICAL_POLAR) {
    // src/coord/spherical_polar.cpp:46-48
    SNAPX_CHECK_ARG(c->nx2 <= 1 || (c->x2min >= 0.0 && c->x2max <= M_PI),
                    "Spherical-polar coordinates require x2 in [0, pi]");
    // the reference's spherical source reads the x1 fluxes, which exist only
    // on blocks with an active x1 axis
    SNAPX_CHECK_ARG(c->nx1 > 1, "spherical-polar blocks need an active x1 (radial) axis");
  }
  Grid g = make_grid(c);
  SNAPX_CHECK_ARG((int64_t)2 * 5 * g.n1 * g.n2 * g.n3 < ((int64_t)1 << 31),
                  "block too large for 32-bit indexing");
  return 0;
}

}  // namespace snapx

extern "C" const char* snapx_last_error(void) { return snapx::g_last_error.c_str(); }

extern "C" int snapx_version(void) { return SNAPX_VERSION; }

extern "C" int snapx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
