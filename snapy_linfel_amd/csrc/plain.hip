// Plain (module-by-module) path of the finite-volume RHS on MI355X.
//
// One kernel per reference module call, full-array intermediates in HBM, in
// the data layout of the reference ([nvar][nc3][nc2][nc1], x1 fastest).  Each
// kernel maps consecutive threads to consecutive x1 cells, so every load and
// store of a wavefront is coalesced for all three stencil directions.  This
// translation unit is built with -ffp-contract=off so that FP64 results follow
// the reference's operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "geom_impl.hpp"
#include "interp_impl.hpp"
#include "riemann_impl.hpp"

namespace snapx {

constexpr int kThreads = 256;
constexpr double kRgas = 8.31446261815324;  // CODATA; kintera::constants::Rgas

// ---------------------------------------------------------------- EOS ----

// EquationOfStateImpl::apply_conserved_limiter_ + ideal_gas_cons2prim +
// apply_primitive_limiter_ (src/eos/equation_of_state.cpp:110-179,
// src/eos/ideal_gas_impl.h:15-35).  The velocities are raised to
// contravariant components with the cell's cos(theta)
// (coord_vec_raise_impl, src/coord/coord_utils_impl.h:8-16) on gnomonic
// panels (gno); elsewhere cos(theta) = 0 makes the raise an identity.
template <typename T>
__device__ __forceinline__ void cons2prim_cell(T (&U)[5], T (&W)[5], T gammad, int limiter,
                                               T dfloor, T pfloor, T cv_tfloor,
                                               bool gno = false, T cth = T(0.)) {
  if (limiter) {
#pragma unroll
    for (int v = 0; v < 5; ++v) U[v] = isnan(U[v]) ? T(0.) : U[v];
    U[IDN] = U[IDN] < dfloor ? dfloor : U[IDN];
    T ke = T(0.5) * ((U[IVX] * U[IVX] + U[IVY] * U[IVY]) + U[IVZ] * U[IVZ]) / U[IDN];
    T lo = ke + U[IDN] * cv_tfloor;
    U[IPR] = U[IPR] < lo ? lo : U[IPR];
  }
  W[IDN] = U[IDN];
  W[IVX] = U[IVX] / W[IDN];
  W[IVY] = U[IVY] / W[IDN];
  W[IVZ] = U[IVZ] / W[IDN];
  if (gno) {
    const T v = W[IVY], x = W[IVZ], sth2 = T(1.) - cth * cth;
    W[IVY] = v / sth2 - x * cth / sth2;
    W[IVZ] = -v * cth / sth2 + x / sth2;
  }
  T ke = T(0.5) * ((W[IVX] * U[IVX] + W[IVY] * U[IVY]) + W[IVZ] * U[IVZ]);
  T ie = U[IPR] - ke;
  W[IPR] = (gammad - T(1.)) * ie;
  if (limiter) {
#pragma unroll
    for (int v = 0; v < 5; ++v) W[v] = isnan(W[v]) ? T(0.) : W[v];
    W[IDN] = W[IDN] < dfloor ? dfloor : W[IDN];
    W[IPR] = W[IPR] < pfloor ? pfloor : W[IPR];
  }
}

template <typename T>
__global__ void cons2prim_kernel(T* __restrict__ u, T* __restrict__ w, int ncell, T gammad,
                                 int limiter, T dfloor, T pfloor, T cv_tfloor, Geometry geo) {
  const bool gno = geo.coord == SNAPX_COORD_GNOMONIC_EQUIANGLE;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    T U[5], W[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) U[v] = u[v * ncell + q];
    const T cth = gno ? T(gno_cos_cell(geo, q / (geo.n1 * geo.n2), (q / geo.n1) % geo.n2))
                      : T(0.);
    cons2prim_cell(U, W, gammad, limiter, dfloor, pfloor, cv_tfloor, gno, cth);
    if (limiter) {
#pragma unroll
      for (int v = 0; v < 5; ++v) u[v * ncell + q] = U[v];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) w[v * ncell + q] = W[v];
  }
}

// IdealGasImpl::_prim2cons (src/eos/ideal_gas.cpp:67-88) with both limiters
template <typename T>
__global__ void prim2cons_kernel(T* __restrict__ w, T* __restrict__ u, int ncell, T gammad,
                                 int limiter, T dfloor, T pfloor, T cv_tfloor, Geometry geo) {
  const bool gno = geo.coord == SNAPX_COORD_GNOMONIC_EQUIANGLE;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    T W[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) W[v] = w[v * ncell + q];
    if (limiter) {
#pragma unroll
      for (int v = 0; v < 5; ++v) W[v] = isnan(W[v]) ? T(0.) : W[v];
      W[IDN] = W[IDN] < dfloor ? dfloor : W[IDN];
      W[IPR] = W[IPR] < pfloor ? pfloor : W[IPR];
#pragma unroll
      for (int v = 0; v < 5; ++v) w[v * ncell + q] = W[v];
    }
    T U[5];
    U[IDN] = W[IDN];
    U[IVX] = W[IVX] * W[IDN];
    U[IVY] = W[IVY] * W[IDN];
    U[IVZ] = W[IVZ] * W[IDN];
    if (gno) {  // coord_vec_lower_impl (src/coord/coord_utils_impl.h:18-24)
      const T cth = T(gno_cos_cell(geo, q / (geo.n1 * geo.n2), (q / geo.n1) % geo.n2));
      const T m2 = U[IVY], m3 = U[IVZ];
      U[IVY] = m2 + m3 * cth;
      U[IVZ] = m3 + m2 * cth;
    }
    T ke = T(0.5) * ((W[IVX] * U[IVX] + W[IVY] * U[IVY]) + W[IVZ] * U[IVZ]);
    U[IPR] = W[IPR] / (gammad - T(1));
    U[IPR] += ke;
    if (limiter) {
#pragma unroll
      for (int v = 0; v < 5; ++v) U[v] = isnan(U[v]) ? T(0.) : U[v];
      U[IDN] = U[IDN] < dfloor ? dfloor : U[IDN];
      T ke2 = T(0.5) * ((U[IVX] * U[IVX] + U[IVY] * U[IVY]) + U[IVZ] * U[IVZ]) / U[IDN];
      T lo = ke2 + U[IDN] * cv_tfloor;
      U[IPR] = U[IPR] < lo ? lo : U[IPR];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v * ncell + q] = U[v];
  }
}

// ------------------------------------------------------ reconstruction ----

// ReconstructImpl::forward (src/recon/reconstruct.cpp:75-154).  One thread
// per (variable, face): the right state at face f is the lower-face value of
// cell clamp(f, il, iu) and the left state the upper-face value of the cell
// below it, which reproduces both the computed faces and the dummy-face copies
// of _apply_inplace (:46-61).
template <typename T>
__global__ void recon_kernel(const T* __restrict__ w, T* __restrict__ wlr, int nvar, int n3,
                             int n2, int n1, int dim, int type, int scale, int shock,
                             int has_parent, int limiter, T dfloor, T pfloor, int revise, int ng) {
  int ncell = n1 * n2 * n3;
  int total = ncell * nvar;
  int N = interp_stencils(type);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int v = t / ncell;
    int q = t - v * ncell;
    int i = q % n1;
    int j = (q / n1) % n2;
    int k = q / (n1 * n2);
    int s, f, n;
    if (dim == 3) {
      s = 1, f = i, n = n1;
    } else if (dim == 2) {
      s = n1, f = j, n = n2;
    } else {
      s = n1 * n2, f = k, n = n3;
    }
    const T* p = w + v * ncell + q - f * s;
    int il = N / 2 + 1, iu = n - il;
    int tv = (has_parent && !shock && v >= IVX && v <= IPR) ? interp2_type(type) : type;
    int fc = f < il ? il : (f > iu ? iu : f);
    T rt = interp_face(tv, 0, scale, p, s, fc);
    T lt = interp_face(tv, 1, scale, p, s, fc - 1);
    if (!has_parent) {
      rt = rt < T(0.) ? T(0.) : rt;
      lt = lt < T(0.) ? T(0.) : lt;
    } else if (!shock && limiter) {
      T lo = v == IDN ? dfloor : (v == IPR ? pfloor : (v > IPR ? T(0.) : -INFINITY));
      rt = rt < lo ? lo : rt;
      lt = lt < lo ? lo : lt;
    }
    if (revise && (v == IDN || v == IPR)) {
      // HydroImpl::_revise_x1inner_lr / _revise_x1outer_lr (src/hydro/hydro.cpp:217-229)
      if (f == ng) lt = rt;
      if (f == n - ng) rt = lt;
    }
    wlr[t] = lt;
    wlr[total + t] = rt;
  }
}

// ------------------------------------------------------------ Riemann ----

// frame: on gnomonic panels LMARS and HLLC take the states in the face's
// orthonormal frame and return the flux to the panel's components
// (src/riemann/lmars.cpp:54-99, src/riemann/hllc.cpp:56-103); Roe does not
// (src/riemann/roe.cpp:14-59)
template <typename T>
__global__ void riemann_kernel(int solver, const T* __restrict__ wl, const T* __restrict__ wr,
                               T* __restrict__ flx, T* __restrict__ fp, int ncell, int dim,
                               T gammad, int frame, Geometry geo) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    T a[5], b[5], f[5], p;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      a[v] = wl[v * ncell + q];
      b[v] = wr[v * ncell + q];
    }
    T fc = T(0.), fs = T(1.);
    if (frame) {
      const GnoAngle ang = gno_face_angle(geo, dim, q / (geo.n1 * geo.n2), (q / geo.n1) % geo.n2);
      fc = T(ang.c), fs = T(ang.s);
      gno_prim2local(a, dim, fc, fs);
      gno_prim2local(b, dim, fc, fs);
    }
    riemann_flux(solver, a, b, gammad, dim, f, &p);
    if (frame) gno_flux2global(f, dim, fc, fs);
#pragma unroll
    for (int v = 0; v < 5; ++v) flx[v * ncell + q] = f[v];
    if (fp) fp[q] = p;
  }
}

// ----------------------------------------------- divergence + forcings ----

struct DuArgs {
  int n1, n2, n3, ng;
  double dx1, dx2, dx3;
  double dt, grav1, grav2, grav3, nh;
  const double* dtp;  // device time step; replaces dt when set
};

// CoordinateImpl::divergence (src/coord/coordinate.cpp:352-396), the
// interior assembly du = -dt D (src/hydro/hydro_forward.cpp:183-185), the
// constant-gravity forcing on every cell (src/forcing/const_gravity.cpp:26-45)
// and the hydrostatic correction (src/hydro/hydro_forward.cpp:191-195).
template <typename T>
__global__ void du_kernel(const T* __restrict__ f1, const T* __restrict__ f2,
                          const T* __restrict__ f3, const T* __restrict__ w,
                          const T* __restrict__ rho_grav, T* __restrict__ du, DuArgs a) {
  int n1 = a.n1, n2 = a.n2, n3 = a.n3, ng = a.ng;
  int ncell = n1 * n2 * n3;
  T dx1 = T(a.dx1), dx2 = T(a.dx2), dx3 = T(a.dx3);
  T dt = T(a.dtp ? *a.dtp : a.dt);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i = q % n1;
    int j = (q / n1) % n2;
    int k = q / (n1 * n2);
    bool in1 = n1 == 1 || (i >= ng && i < n1 - ng);
    bool in2 = n2 == 1 || (j >= ng && j < n2 - ng);
    bool in3 = n3 == 1 || (k >= ng && k < n3 - ng);
    bool interior = in1 && in2 && in3;
    T vol = dx3 * (dx2 * dx1);
    T d[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      int o = v * ncell + q;
      T dv = T(0.);
      if (interior) {
        if (f1) {
          T ar = dx3 * dx2;
          dv += ar * f1[o + 1] - ar * f1[o];
        }
        if (f2) {
          T ar = dx3 * dx1;
          dv += ar * f2[o + n1] - ar * f2[o];
        }
        if (f3) {
          T ar = dx2 * dx1;
          dv += ar * f3[o + n1 * n2] - ar * f3[o];
        }
        dv = -dt * (dv / vol);
      }
      d[v] = dv;
    }
    T rho = (a.grav1 != 0. || a.grav2 != 0. || a.grav3 != 0.) ? w[IDN * ncell + q] : T(0.);
    if (a.grav1 != 0.) {
      d[IVX] += dt * rho * T(a.grav1) * T(a.nh);
      d[IPR] += dt * rho * w[IVX * ncell + q] * T(a.grav1) * T(a.nh);
    }
    if (a.grav2 != 0.) {
      d[IVY] += dt * rho * T(a.grav2);
      d[IPR] += dt * rho * w[IVY * ncell + q] * T(a.grav2);
    }
    if (a.grav3 != 0.) {
      d[IVZ] += dt * rho * T(a.grav3);
      d[IPR] += dt * rho * w[IVZ * ncell + q] * T(a.grav3);
    }
    if (rho_grav) {
      T rg = rho_grav[q];
      d[IVX] += dt * rg * (T(1.) - T(a.nh));
      d[IPR] += dt * w[IVX * ncell + q] * rg * (T(1.) - T(a.nh));
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) du[v * ncell + q] = d[v];
  }
}

struct CoordArgs {
  Geometry g;
  double dx1, dx2, dx3;  // Cartesian widths
  double dt, grav1, grav2, grav3, nh;
  int du_mode;  // 0: div on every cell; 1: du = -dt div on the interior + forcings
  const double* dtp;  // device time step; replaces dt when set
};

// CoordinateImpl::forward (src/coord/coordinate.cpp:352-434) for both
// geometries and SphericalPolarImpl::forward (src/coord/spherical_polar.cpp:
// 190-279): each axis adds its flux difference on its own interior range
// (all cells of the other axes), then the momenta get the geometric sources.
// fp1 == nullptr selects the 2 p / r form of the radial pressure source.
// With du_mode the interior assembly, gravity and hydrostatic correction of
// du_kernel follow.
template <typename T>
__global__ void coord_forward_kernel(const T* __restrict__ w, const T* __restrict__ f1,
                                     const T* __restrict__ f2, const T* __restrict__ f3,
                                     const T* __restrict__ fp1, const T* __restrict__ rho_grav,
                                     T* __restrict__ out, CoordArgs a) {
  const Geometry& g = a.g;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3, ng = g.ng;
  const int ncell = n1 * n2 * n3;
  const bool sph = g.coord == SNAPX_COORD_SPHERICAL_POLAR;
  const bool gno = g.coord == SNAPX_COORD_GNOMONIC_EQUIANGLE;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i = q % n1;
    int j = (q / n1) % n2;
    int k = q / (n1 * n2);
    bool r1 = n1 == 1 || (i >= ng && i < n1 - ng);
    bool r2 = n2 == 1 || (j >= ng && j < n2 - ng);
    bool r3 = n3 == 1 || (k >= ng && k < n3 - ng);
    double vol, A1m, A1p, A2m, A2p, A3m, A3p;
    SphCell gc{};
    GnoCell gn{};
    if (sph) {
      gc = sph_cell(g, k, j, i);
      vol = gc.vol;
      A1m = gc.a1m, A1p = gc.a1p;
      A2m = gc.a2m, A2p = gc.a2p;
      A3m = A3p = gc.a3;
    } else if (gno) {
      gn = gno_cell(g, k, j, i);
      vol = gn.vol;
      A1m = gn.a1m, A1p = gn.a1p;
      A2m = gn.a2m, A2p = gn.a2p;
      A3m = gn.a3m, A3p = gn.a3p;
    } else {
      vol = a.dx3 * (a.dx2 * a.dx1);
      A1m = A1p = a.dx3 * a.dx2;
      A2m = A2p = a.dx3 * a.dx1;
      A3m = A3p = a.dx2 * a.dx1;
    }
    T d[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      int o = v * ncell + q;
      T acc = T(0.);
      if (f1 && r1) acc += T(A1p) * f1[o + 1] - T(A1m) * f1[o];
      if (f2 && r2) acc += T(A2p) * f2[o + n1] - T(A2m) * f2[o];
      if (f3 && r3) acc += T(A3p) * f3[o + n1 * n2] - T(A3m) * f3[o];
      d[v] = acc / T(vol);
    }
    if (gno) {
      // GnomonicEquiangleImpl::forward (src/coord/gnomonic_equiangle.cpp:362-418)
      T rho = w[IDN * ncell + q], v1 = w[IVX * ncell + q];
      T v2 = w[IVY * ncell + q], v3 = w[IVZ * ncell + q], pre = w[IPR * ncell + q];
      T cosine = T(gn.cos), sine2 = T(gn.sin) * T(gn.sin), radius = T(gn.r);
      T v_2 = v2 + v3 * cosine, v_3 = v3 + v2 * cosine;
      T src1 = rho * (v2 * v_2 + v3 * v_3) / radius;
      if (fp1) {
        if (r1)
          src1 += (T(A1p) * fp1[q + 1] - T(A1m) * fp1[q]) / T(vol) -
                  (fp1[q + 1] - fp1[q]) / T(gn.dx1);
      } else {
        src1 += T(2.0) * pre / radius;
      }
      d[IVX] -= src1;
      d[IVY] -= T(gn.xov) * (pre + rho * v3 * v3 * sine2) - rho * v1 * v_2 / radius;
      d[IVZ] -= T(gn.yov) * (pre + rho * v2 * v2 * sine2) - rho * v1 * v_3 / radius;
    }
    if (sph) {
      T rho = w[IDN * ncell + q], v2 = w[IVY * ncell + q], v3 = w[IVZ * ncell + q];
      T pre = w[IPR * ncell + q];
      T s1i = T(gc.src1_i), s1j = T(gc.src1_j);
      T src1 = s1i * rho * (v2 * v2 + v3 * v3);
      if (fp1) {
        if (r1)
          src1 += (T(A1p) * fp1[q + 1] - T(A1m) * fp1[q]) / T(vol) -
                  (fp1[q + 1] - fp1[q]) / T(gc.dx1);
      } else {
        src1 += T(2.) * s1i * pre;
      }
      d[IVX] -= src1;
      d[IVY] -= s1i * s1j * (rho * (v3 * v3) + pre);
      if (r1 && f1) {
        T s2i = T(gc.src2_i);
        d[IVY] += s2i * (T(A1m) * f1[IVY * ncell + q] + T(A1p) * f1[IVY * ncell + q + 1]);
        d[IVZ] += s2i * (T(A1m) * f1[IVZ * ncell + q] + T(A1p) * f1[IVZ * ncell + q + 1]);
      }
      if (n2 > 1) {
        if (r1 && r2 && f2)
          d[IVZ] += s1i * T(gc.src2_j) *
                    (T(A2m) * f2[IVZ * ncell + q] + T(A2p) * f2[IVZ * ncell + q + n1]);
      } else {
        d[IVZ] += s1i * s1j * (rho * v3 * v2);
      }
    }
    if (!a.du_mode) {
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v * ncell + q] = d[v];
      continue;
    }
    T dt = T(a.dtp ? *a.dtp : a.dt);
    bool interior = r1 && r2 && r3;
#pragma unroll
    for (int v = 0; v < 5; ++v) d[v] = interior ? -dt * d[v] : T(0.);
    T rho = (a.grav1 != 0. || a.grav2 != 0. || a.grav3 != 0.) ? w[IDN * ncell + q] : T(0.);
    if (a.grav1 != 0.) {
      d[IVX] += dt * rho * T(a.grav1) * T(a.nh);
      d[IPR] += dt * rho * w[IVX * ncell + q] * T(a.grav1) * T(a.nh);
    }
    if (a.grav2 != 0.) {
      d[IVY] += dt * rho * T(a.grav2);
      d[IPR] += dt * rho * w[IVY * ncell + q] * T(a.grav2);
    }
    if (a.grav3 != 0.) {
      d[IVZ] += dt * rho * T(a.grav3);
      d[IPR] += dt * rho * w[IVZ * ncell + q] * T(a.grav3);
    }
    if (rho_grav) {
      T rg = rho_grav[q];
      d[IVX] += dt * rg * (T(1.) - T(a.nh));
      d[IPR] += dt * w[IVX * ncell + q] * rg * (T(1.) - T(a.nh));
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v * ncell + q] = d[v];
  }
}

template <typename T>
__global__ void zero_kernel(T* __restrict__ x, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    x[q] = T(0.);
}

// hydrostatic pressure gradient of the reconstructed x1 states
// (src/hydro/hydro_forward.cpp:67-76)
template <typename T>
__global__ void rho_grav_kernel(const T* __restrict__ wlr, T* __restrict__ rg, int n1, int ncell,
                                int nvar, int ng, T dx1) {
  const T* pl = wlr + IPR * ncell;
  const T* pr = wlr + nvar * ncell + IPR * ncell;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i = q % n1;
    rg[q] = (i >= ng && i < n1 - ng) ? (pl[q + 1] - pr[q]) / dx1 : T(0.);
  }
}

// HydroImpl::_revise_x1inner_ghost / _revise_x1outer_ghost
// (src/hydro/hydro.cpp:231-269); one thread per column.
// dz of each ghost layer: dx1v[is - n - 1] inward and dx1v[ie + n] outward
// (the centre distances of the coordinate; the face width on Cartesian blocks)
struct GhostDz {
  double in[8], out[8];
};

template <typename T>
__global__ void revise_ghost_kernel(T* __restrict__ w, int n1, int ncol, int ncell, int ng,
                                    T grav1, T gammad, GhostDz dz) {
  for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < ncol;
       col += gridDim.x * blockDim.x) {
    T* rho = w + IDN * ncell + col * n1;
    T* pre = w + IPR * ncell + col * n1;
    T grav = -grav1;
    T gamma = gammad * T(1.0);
    T gm = gamma - T(1.);
    T a = gm / gamma;
    T inv_g = T(1.) / gamma, inv_a = T(1.) / a;
    int is = ng, ie = n1 - ng - 1;
    T K = pre[is] / pow(rho[is], gamma);
    for (int n = 0; n < ng; ++n) {
      T h = pow(pre[is - n], a) + a * grav * T(dz.in[n]) / pow(K, inv_g);
      pre[is - n - 1] = pow(h, inv_a);
      rho[is - n - 1] = pow(pre[is - n - 1] / K, inv_g);
    }
    K = pre[ie] / pow(rho[ie], gamma);
    for (int n = 0; n < ng; ++n) {
      T h = pow(pre[ie + n], a) - a * grav * T(dz.out[n]) / pow(K, inv_g);
      pre[ie + n + 1] = pow(h, inv_a);
      rho[ie + n + 1] = pow(pre[ie + n + 1] / K, inv_g);
    }
  }
}

// ---------------------------------------------------------- boundaries ----

// BC_FUNCTIONs of src/bc/bc_func.cpp:7-63 for both sides of one axis.
template <typename T>
__global__ void bc_kernel(T* __restrict__ var, int nvar, int n3, int n2, int n1, int ng, int dim,
                          int type_in, int type_out, int flip_normal) {
  int s, n, np;
  if (dim == 3) {
    s = 1, n = n1, np = n2 * n3;
  } else if (dim == 2) {
    s = n1, n = n2, np = n1 * n3;
  } else {
    s = n1 * n2, n = n3, np = n1 * n2;
  }
  int ncell = n1 * n2 * n3;
  int total = nvar * 2 * ng * np;
  int vnorm = 4 - dim;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int p = t % np;
    int g2 = (t / np) % (2 * ng);
    int v = t / (np * 2 * ng);
    int base;
    if (dim == 3)
      base = p * n1;
    else if (dim == 2)
      base = (p / n1) * n1 * n2 + p % n1;
    else
      base = p;
    bool outer = g2 >= ng;
    int g = outer ? g2 - ng : g2;
    int type = outer ? type_out : type_in;
    if (type < 0 || type == SNAPX_BC_CUSTOM) continue;  // custom: the caller's ghosts
    int dst, src;
    if (!outer) {
      dst = g;
      src = type == SNAPX_BC_REFLECTING ? 2 * ng - 1 - g
            : type == SNAPX_BC_OUTFLOW  ? ng
                                        : n - 2 * ng + g;
    } else {
      dst = n - ng + g;
      src = type == SNAPX_BC_REFLECTING ? n - ng - 1 - g
            : type == SNAPX_BC_OUTFLOW  ? n - ng - 1
                                        : ng + g;
    }
    T* x = var + v * ncell + base;
    T val = x[src * s];
    if (type == SNAPX_BC_REFLECTING && v == vnorm && flip_normal) val = val * T(-1);
    x[dst * s] = val;
  }
}

// -------------------------------------------------------------- RK / dt ----

template <typename T>
__global__ void rk_combine_kernel(T* __restrict__ u, const T* __restrict__ u0,
                                  const T* __restrict__ u1, const T* __restrict__ du, T w0, T w1,
                                  T w2, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    u[q] = (w0 * u0[q] + w1 * u1[q]) + w2 * du[q];
}

__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

// Order key of a CFL candidate for an unsigned atomic minimum.  Candidates
// are width / (|v| + cs) >= 0 (or +inf), whose bit patterns order like their
// values; shifted up by one they leave key 0 for NaN, so that one NaN
// candidate wins the minimum, as the reference's tensor min() propagates NaN.
__host__ __device__ inline unsigned long long cfl_key(double c) {
  if (c != c) return 0ull;
  unsigned long long b;
  memcpy(&b, &c, sizeof(b));
  return b + 1ull;
}

__device__ inline double cfl_value(unsigned long long key) {
  return key == 0ull ? __longlong_as_double(0x7ff8000000000000ll)
                     : __longlong_as_double((long long)(key - 1ull));
}

// HydroImpl::max_time_step (src/hydro/hydro.cpp:118-180) with the
// coordinate's center widths: Cartesian dx, spherical-polar dr, r dtheta and
// r sin(theta) dphi (src/coord/spherical_polar.cpp:146-155)
template <typename T>
__global__ void cfl_kernel(const T* __restrict__ w, unsigned long long* __restrict__ out, int n3,
                           int n2, int n1, int ng, T gammad, T dx1, T dx2, T dx3, int use1,
                           int use2, int use3, Geometry geo) {
  int ncell = n1 * n2 * n3;
  unsigned long long best = cfl_key(1.e9);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i = q % n1;
    int j = (q / n1) % n2;
    int k = q / (n1 * n2);
    bool in1 = n1 == 1 || (i >= ng && i < n1 - ng);
    bool in2 = n2 == 1 || (j >= ng && j < n2 - ng);
    bool in3 = n3 == 1 || (k >= ng && k < n3 - ng);
    if (!(in1 && in2 && in3)) continue;
    T cs = sqrt(gammad * T(1.0) * w[IPR * ncell + q] / w[IDN * ncell + q]);
    if (geo.coord == SNAPX_COORD_SPHERICAL_POLAR) {
      dx1 = T(axis_width(geo.ax[0], i));
      if (use2) dx2 = T(sph_width2(geo, j, i));
      if (use3) dx3 = T(sph_width3(geo, k, j, i));
    } else if (geo.coord == SNAPX_COORD_GNOMONIC_EQUIANGLE) {
      if (use2) dx2 = T(gno_width2(geo, k, j, i));
      if (use3) dx3 = T(gno_width3(geo, k, j, i));
    }
    if (use1) best = min(best, cfl_key((double)(dx1 / (fabs(w[IVX * ncell + q]) + cs))));
    if (use2) best = min(best, cfl_key((double)(dx2 / (fabs(w[IVY * ncell + q]) + cs))));
    if (use3) best = min(best, cfl_key((double)(dx3 / (fabs(w[IVZ * ncell + q]) + cs))));
  }
  // wavefront reduction, then one atomic per wavefront
  for (int off = warpSize / 2; off > 0; off >>= 1) best = min(best, __shfl_down(best, off));
  if ((threadIdx.x & (warpSize - 1)) == 0) atomicMin(out, best);
}

__global__ void cfl_finalize_kernel(double* dt, double cfl) {
  double v = cfl_value(*reinterpret_cast<const unsigned long long*>(dt));
  *dt = 1.0 * cfl * v;
}

// Positivity screen of MeshBlockImpl::check_redo (src/mesh/meshblock.cpp:974-1028)
// over the interior cells.  The reference takes tensor minima, which propagate
// NaN, and compares them with zero, so the accumulator records "some value <= 0"
// and "some value is NaN" separately for density, total energy and the
// cons2prim pressure; state_finalize_kernel turns them into the verdict.
enum : int { kRhoLe0 = 1, kRhoNan = 2, kELe0 = 4, kENan = 8, kPLe0 = 16, kPNan = 32,
             kNonFinite = 64 };

__global__ void fill_i32_kernel(int* p, int v) { *p = v; }

template <typename T>
__global__ void state_screen_kernel(const T* __restrict__ u, int* __restrict__ acc, int n3,
                                    int n2, int n1, int ng, T gammad, int limiter, T dfloor,
                                    T pfloor, T cv_tfloor, Geometry geo) {
  int ncell = n1 * n2 * n3;
  int bits = 0;
  const bool gno = geo.coord == SNAPX_COORD_GNOMONIC_EQUIANGLE;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    int i = q % n1;
    int j = (q / n1) % n2;
    int k = q / (n1 * n2);
    bool in1 = n1 == 1 || (i >= ng && i < n1 - ng);
    bool in2 = n2 == 1 || (j >= ng && j < n2 - ng);
    bool in3 = n3 == 1 || (k >= ng && k < n3 - ng);
    if (!(in1 && in2 && in3)) continue;
    T U[5], W[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) U[v] = u[v * ncell + q];
#pragma unroll
    for (int v = 0; v < 5; ++v) bits |= isfinite(U[v]) ? 0 : kNonFinite;
    bits |= isnan(U[IDN]) ? kRhoNan : (U[IDN] <= T(0.) ? kRhoLe0 : 0);
    bits |= isnan(U[IPR]) ? kENan : (U[IPR] <= T(0.) ? kELe0 : 0);
    const T cth = gno ? T(gno_cos_cell(geo, k, j)) : T(0.);
    cons2prim_cell(U, W, gammad, limiter, dfloor, pfloor, cv_tfloor, gno, cth);
    bits |= isnan(W[IPR]) ? kPNan : (W[IPR] <= T(0.) ? kPLe0 : 0);
  }
  for (int off = warpSize / 2; off > 0; off >>= 1) bits |= __shfl_down(bits, off);
  if ((threadIdx.x & (warpSize - 1)) == 0 && bits) atomicOr(acc, bits);
}

// min(rho) <= 0 is false when some rho is NaN; the pressure redo needs both the
// energy screen and the pressure test to trip (meshblock.cpp:994-998).
__global__ void state_finalize_kernel(int* flag) {
  int b = *flag;
  int rho = (b & kRhoLe0) && !(b & kRhoNan);
  int pres = (b & kELe0) && !(b & kENan) && (b & kPLe0) && !(b & kPNan);
  *flag = (rho ? SNAPX_REDO_DENSITY : 0) | (pres ? SNAPX_REDO_PRESSURE : 0) |
          ((b & kNonFinite) ? SNAPX_STATE_NONFINITE : 0);
}

// ------------------------------------------------------------ launches ----

template <typename T>
T* as(void* p) {
  return reinterpret_cast<T*>(p);
}

template <typename T>
int launch_cons2prim(const snapx_config* c, void* u, void* w, hipStream_t st) {
  Grid g = make_grid(c);
  double cvt = kRgas / c->weight / (c->gammad - 1.) * c->temperature_floor;
  hipLaunchKernelGGL(cons2prim_kernel<T>, blocks_for(g.ncell, kThreads), dim3(kThreads), 0, st,
                     as<T>(u), as<T>(w), g.ncell, T(c->gammad), c->limiter, T(c->density_floor),
                     T(c->pressure_floor), T(cvt), make_geometry(c));
  SNAPX_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_recon(const void* w, void* wlr, int nvar, int n3, int n2, int n1, int dim, int type,
                 int scale, int shock, int has_parent, int limiter, double dfloor, double pfloor,
                 int revise, int ng, hipStream_t st) {
  int64_t total = (int64_t)nvar * n1 * n2 * n3;
  hipLaunchKernelGGL(recon_kernel<T>, blocks_for(total, kThreads), dim3(kThreads), 0, st,
                     reinterpret_cast<const T*>(w), as<T>(wlr), nvar, n3, n2, n1, dim, type,
                     scale, shock, has_parent, limiter, T(dfloor), T(pfloor), revise, ng);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

// geo: the block's geometry when the face frames of a gnomonic panel apply
template <typename T>
int launch_riemann(int solver, const void* wl, const void* wr, void* flx, void* fp, int ncell,
                   int dim, double gammad, hipStream_t st, const Geometry* geo = nullptr) {
  const int frame = geo && geo->coord == SNAPX_COORD_GNOMONIC_EQUIANGLE &&
                    solver != SNAPX_RIEMANN_ROE;
  hipLaunchKernelGGL(riemann_kernel<T>, blocks_for(ncell, kThreads), dim3(kThreads), 0, st,
                     solver, reinterpret_cast<const T*>(wl), reinterpret_cast<const T*>(wr),
                     as<T>(flx), as<T>(fp), ncell, dim, T(gammad), frame,
                     geo ? *geo : Geometry{});
  SNAPX_LAUNCH_CHECK();
  return 0;
}

template <typename T>
int launch_bc(const snapx_config* c, void* var, int nvar, int kind, hipStream_t st) {
  Grid g = make_grid(c);
  int dims[3] = {3, 2, 1};
  int ns[3] = {g.n1, g.n2, g.n3};
  for (int a = 0; a < 3; ++a) {
    if (ns[a] == 1) continue;
    int tin = c->bc[2 * a], tout = c->bc[2 * a + 1];
    if (tin < 0 && tout < 0) continue;
    int np = g.ncell / ns[a];
    int64_t total = (int64_t)nvar * 2 * g.ng * np;
    hipLaunchKernelGGL(bc_kernel<T>, blocks_for(total, kThreads), dim3(kThreads), 0, st,
                       as<T>(var), nvar, g.n3, g.n2, g.n1, g.ng, dims[a], tin, tout,
                       kind == SNAPX_CONSERVED || kind == SNAPX_PRIMITIVE);
    SNAPX_LAUNCH_CHECK();
  }
  return 0;
}

int launch_bc_f32(const snapx_config* c, float* var, hipStream_t st) {
  return launch_bc<float>(c, var, 5, SNAPX_CONSERVED, st);
}

}  // namespace snapx

using namespace snapx;

// implemented in vic.hip
namespace snapx {
int implicit_correction(const snapx_config* cfg, double dt, void* du, const void* w,
                        void* workspace, size_t bytes, int dtype, hipStream_t st,
                        const double* dtp = nullptr);
size_t implicit_workspace_bytes(const snapx_config* cfg);
}  // namespace snapx

extern "C" int snapx_ideal_gas_cons2prim(const snapx_config* cfg, void* u, void* w, int dtype,
                                         void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(u && w, "snapx_ideal_gas_cons2prim: null pointer");
  return dtype == SNAPX_F64 ? launch_cons2prim<double>(cfg, u, w, as_stream(stream))
                            : launch_cons2prim<float>(cfg, u, w, as_stream(stream));
}

extern "C" int snapx_ideal_gas_prim2cons(const snapx_config* cfg, void* w, void* u, int dtype,
                                         void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(u && w, "snapx_ideal_gas_prim2cons: null pointer");
  Grid g = make_grid(cfg);
  double cvt = kRgas / cfg->weight / (cfg->gammad - 1.) * cfg->temperature_floor;
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(prim2cons_kernel<double>, blocks_for(g.ncell, kThreads), dim3(kThreads),
                       0, st, as<double>(w), as<double>(u), g.ncell, cfg->gammad, cfg->limiter,
                       cfg->density_floor, cfg->pressure_floor, cvt, make_geometry(cfg));
  else
    hipLaunchKernelGGL(prim2cons_kernel<float>, blocks_for(g.ncell, kThreads), dim3(kThreads),
                       0, st, as<float>(w), as<float>(u), g.ncell, (float)cfg->gammad,
                       cfg->limiter, (float)cfg->density_floor, (float)cfg->pressure_floor,
                       (float)cvt, make_geometry(cfg));
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_reconstruct(const void* w, void* wlr, int nvar, int n3, int n2, int n1,
                                 int dim, int interp_type, int scale, int shock, int has_parent,
                                 int limiter, double density_floor, double pressure_floor,
                                 int revise_x1_lr, int nghost, int dtype, void* stream) {
  SNAPX_CHECK_ARG(w && wlr, "snapx_reconstruct: null pointer");
  SNAPX_CHECK_ARG(nvar == 5, "snapx_reconstruct: nvar must be 5 (dry ideal gas)");
  SNAPX_CHECK_ARG(dim >= 1 && dim <= 3, "snapx_reconstruct: dim must be 1, 2 or 3");
  SNAPX_CHECK_ARG(interp_type >= SNAPX_INTERP_DC && interp_type <= SNAPX_INTERP_WENO5,
                  "snapx_reconstruct: unknown interpolation type");
  int n = dim == 3 ? n1 : (dim == 2 ? n2 : n3);
  int il = interp_stencils(interp_type) / 2 + 1;
  SNAPX_CHECK_ARG(il <= n - il, "snapx_reconstruct: il > iu (axis too short for the stencil)");
  SNAPX_CHECK_ARG((int64_t)2 * nvar * n1 * n2 * n3 < (int64_t)1 << 31,
                  "snapx_reconstruct: block too large for 32-bit indexing");
  SNAPX_CHECK_ARG(!revise_x1_lr || dim == 3, "snapx_reconstruct: x1 LR revision needs dim 3");
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    return launch_recon<double>(w, wlr, nvar, n3, n2, n1, dim, interp_type, scale, shock,
                                has_parent, limiter, density_floor, pressure_floor, revise_x1_lr,
                                nghost, st);
  return launch_recon<float>(w, wlr, nvar, n3, n2, n1, dim, interp_type, scale, shock, has_parent,
                             limiter, density_floor, pressure_floor, revise_x1_lr, nghost, st);
}

extern "C" int snapx_riemann(int solver, const void* wl, const void* wr, void* flx,
                             void* face_pressure, int nvar, int n3, int n2, int n1, int dim,
                             double gammad, int dtype, void* stream) {
  SNAPX_CHECK_ARG(wl && wr && flx, "snapx_riemann: null pointer");
  SNAPX_CHECK_ARG(nvar == 5, "snapx_riemann: nvar must be 5 (dry ideal gas)");
  SNAPX_CHECK_ARG(dim >= 1 && dim <= 3, "snapx_riemann: dim must be 1, 2 or 3");
  SNAPX_CHECK_ARG(solver >= SNAPX_RIEMANN_ROE && solver <= SNAPX_RIEMANN_HLLC,
                  "snapx_riemann: unknown solver");
  SNAPX_CHECK_ARG((int64_t)nvar * n1 * n2 * n3 < (int64_t)1 << 31,
                  "snapx_riemann: block too large for 32-bit indexing");
  int ncell = n1 * n2 * n3;
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    return launch_riemann<double>(solver, wl, wr, flx, face_pressure, ncell, dim, gammad, st);
  return launch_riemann<float>(solver, wl, wr, flx, face_pressure, ncell, dim, gammad, st);
}

static DuArgs du_args(const snapx_config* c, double dt, const double* dtp) {
  Grid g = make_grid(c);
  DuArgs a;
  a.n1 = g.n1, a.n2 = g.n2, a.n3 = g.n3, a.ng = g.ng;
  a.dx1 = g.dx1, a.dx2 = g.dx2, a.dx3 = g.dx3;
  a.dt = dt;
  a.dtp = dtp;
  a.grav1 = c->grav1, a.grav2 = c->grav2, a.grav3 = c->grav3, a.nh = c->non_hydrostatic;
  return a;
}

template <typename T>
static int launch_du(const snapx_config* c, double dt, const void* f1, const void* f2,
                     const void* f3, const void* w, const void* rg, void* du, hipStream_t st,
                     const double* dtp = nullptr) {
  Grid g = make_grid(c);
  if (g.n1 == 1) f1 = nullptr;  // no faces on an inactive axis
  if (g.n2 == 1) f2 = nullptr;
  if (g.n3 == 1) f3 = nullptr;
  hipLaunchKernelGGL(du_kernel<T>, blocks_for(g.ncell, kThreads), dim3(kThreads), 0, st,
                     reinterpret_cast<const T*>(f1), reinterpret_cast<const T*>(f2),
                     reinterpret_cast<const T*>(f3), reinterpret_cast<const T*>(w),
                     reinterpret_cast<const T*>(rg), as<T>(du), du_args(c, dt, dtp));
  SNAPX_LAUNCH_CHECK();
  return 0;
}

static GhostDz ghost_dz(const snapx_config* c) {
  Grid g = make_grid(c);
  Geometry geo = make_geometry(c);
  GhostDz dz;
  int is = g.ng, ie = g.n1 - g.ng - 1;
  for (int n = 0; n < g.ng; ++n) {
    bool sph = c->coord == SNAPX_COORD_SPHERICAL_POLAR;
    dz.in[n] = sph ? sph_dx1v(geo, is - n - 1) : g.dx1;
    dz.out[n] = sph ? sph_dx1v(geo, ie + n) : g.dx1;
  }
  return dz;
}

static CoordArgs coord_args(const snapx_config* c, double dt, int du_mode,
                            const double* dtp) {
  Grid g = make_grid(c);
  CoordArgs a;
  a.g = make_geometry(c);
  a.dx1 = g.dx1, a.dx2 = g.dx2, a.dx3 = g.dx3;
  a.dt = dt;
  a.dtp = dtp;
  a.grav1 = c->grav1, a.grav2 = c->grav2, a.grav3 = c->grav3, a.nh = c->non_hydrostatic;
  a.du_mode = du_mode;
  return a;
}

template <typename T>
static int launch_coord(const snapx_config* c, double dt, int du_mode, const void* w,
                        const void* f1, const void* f2, const void* f3, const void* fp1,
                        const void* rg, void* out, hipStream_t st, const double* dtp = nullptr) {
  Grid g = make_grid(c);
  // an inactive axis has no faces: its flux (and the x1 face pressure) is
  // never read, so a one-cell axis cannot index past the field
  if (g.n1 == 1) f1 = fp1 = nullptr;
  if (g.n2 == 1) f2 = nullptr;
  if (g.n3 == 1) f3 = nullptr;
  hipLaunchKernelGGL(coord_forward_kernel<T>, blocks_for(g.ncell, kThreads), dim3(kThreads), 0,
                     st, reinterpret_cast<const T*>(w), reinterpret_cast<const T*>(f1),
                     reinterpret_cast<const T*>(f2), reinterpret_cast<const T*>(f3),
                     reinterpret_cast<const T*>(fp1), reinterpret_cast<const T*>(rg), as<T>(out),
                     coord_args(c, dt, du_mode, dtp));
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_coord_forward(const snapx_config* cfg, const void* w, const void* f1,
                                   const void* f2, const void* f3, const void* fp1, void* div,
                                   int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(div, "snapx_coord_forward: null output");
  SNAPX_CHECK_ARG(w || cfg->coord == SNAPX_COORD_CARTESIAN,
                  "snapx_coord_forward: spherical-polar sources need the primitives w");
  hipStream_t st = as_stream(stream);
  return dtype == SNAPX_F64
             ? launch_coord<double>(cfg, 0., 0, w, f1, f2, f3, fp1, nullptr, div, st)
             : launch_coord<float>(cfg, 0., 0, w, f1, f2, f3, fp1, nullptr, div, st);
}

extern "C" int snapx_divergence(const snapx_config* cfg, const void* f1, const void* f2,
                                const void* f3, void* div, int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(div, "snapx_divergence: null output");
  SNAPX_CHECK_ARG(cfg->coord == SNAPX_COORD_CARTESIAN,
                  "snapx_divergence: Cartesian only (snapx_coord_forward covers spherical-polar)");
  // D = -(du / dt) with dt = -1 and no forcing: du_kernel with dt = -1 gives
  // +D on the interior; the reference leaves non-interior cells at 0 / vol = 0
  // except on the ghost rows of the other axes, which no caller reads.
  snapx_config c = *cfg;
  c.grav1 = c.grav2 = c.grav3 = 0.;
  hipStream_t st = as_stream(stream);
  // w is not read when there is no gravity
  return dtype == SNAPX_F64 ? launch_du<double>(&c, -1.0, f1, f2, f3, nullptr, nullptr, div, st)
                            : launch_du<float>(&c, -1.0, f1, f2, f3, nullptr, nullptr, div, st);
}

extern "C" int snapx_revise_x1_ghost(const snapx_config* cfg, void* w, int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  Grid g = make_grid(cfg);
  SNAPX_CHECK_ARG(g.n1 > 1, "snapx_revise_x1_ghost: x1 axis is inactive");
  int ncol = g.n2 * g.n3;
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(revise_ghost_kernel<double>, blocks_for(ncol, kThreads), dim3(kThreads), 0,
                       st, as<double>(w), g.n1, ncol, g.ncell, g.ng, cfg->grav1, cfg->gammad,
                       ghost_dz(cfg));
  else
    hipLaunchKernelGGL(revise_ghost_kernel<float>, blocks_for(ncol, kThreads), dim3(kThreads), 0,
                       st, as<float>(w), g.n1, ncol, g.ncell, g.ng, (float)cfg->grav1,
                       (float)cfg->gammad, ghost_dz(cfg));
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_apply_bc(const snapx_config* cfg, void* var, int nvar, int var_kind,
                              int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(var && nvar > 0, "snapx_apply_bc: bad variable array");
  return dtype == SNAPX_F64 ? launch_bc<double>(cfg, var, nvar, var_kind, as_stream(stream))
                            : launch_bc<float>(cfg, var, nvar, var_kind, as_stream(stream));
}

extern "C" int snapx_rk_combine(void* u, const void* u0, const void* u1, const void* du,
                                double w0, double w1, double w2, int64_t n, int dtype,
                                void* stream) {
  SNAPX_CHECK_ARG(u && u0 && u1 && du && n >= 0, "snapx_rk_combine: bad arguments");
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(rk_combine_kernel<double>, blocks_for(n, kThreads), dim3(kThreads), 0, st,
                       as<double>(u), reinterpret_cast<const double*>(u0),
                       reinterpret_cast<const double*>(u1), reinterpret_cast<const double*>(du),
                       w0, w1, w2, n);
  else
    hipLaunchKernelGGL(rk_combine_kernel<float>, blocks_for(n, kThreads), dim3(kThreads), 0, st,
                       as<float>(u), reinterpret_cast<const float*>(u0),
                       reinterpret_cast<const float*>(u1), reinterpret_cast<const float*>(du),
                       (float)w0, (float)w1, (float)w2, n);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_max_time_step(const snapx_config* cfg, const void* w, double cfl,
                                   double* dt_out, int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(w && dt_out, "snapx_max_time_step: null pointer");
  Grid g = make_grid(cfg);
  int sch = cfg->implicit_scheme;
  int use1 = g.n1 > 1 && (!(sch & 1) || (g.n3 == 1 && g.n2 == 1));
  int use2 = g.n2 > 1 && !((sch >> 1) & 1);
  int use3 = g.n3 > 1 && !((sch >> 2) & 1);
  hipStream_t st = as_stream(stream);
  auto* acc = reinterpret_cast<unsigned long long*>(dt_out);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(1), 0, st, acc, cfl_key(1.e9));
  int nb = std::min<int64_t>((g.ncell + kThreads - 1) / kThreads, 4096);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(cfl_kernel<double>, dim3(nb), dim3(kThreads), 0, st,
                       reinterpret_cast<const double*>(w), acc, g.n3, g.n2, g.n1, g.ng,
                       cfg->gammad, g.dx1, g.dx2, g.dx3, use1, use2, use3, make_geometry(cfg));
  else
    hipLaunchKernelGGL(cfl_kernel<float>, dim3(nb), dim3(kThreads), 0, st,
                       reinterpret_cast<const float*>(w), acc, g.n3, g.n2, g.n1, g.ng,
                       (float)cfg->gammad, (float)g.dx1, (float)g.dx2, (float)g.dx3, use1, use2,
                       use3, make_geometry(cfg));
  hipLaunchKernelGGL(cfl_finalize_kernel, dim3(1), dim3(1), 0, st, dt_out, cfl);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_check_state(const snapx_config* cfg, const void* u, int* flag, int dtype,
                                 void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(u && flag, "snapx_check_state: null pointer");
  Grid g = make_grid(cfg);
  hipStream_t st = as_stream(stream);
  double cvt = kRgas / cfg->weight / (cfg->gammad - 1.) * cfg->temperature_floor;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(1), 0, st, flag, 0);
  int nb = std::min<int64_t>((g.ncell + kThreads - 1) / kThreads, 4096);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(state_screen_kernel<double>, dim3(nb), dim3(kThreads), 0, st,
                       reinterpret_cast<const double*>(u), flag, g.n3, g.n2, g.n1, g.ng,
                       cfg->gammad, cfg->limiter, cfg->density_floor, cfg->pressure_floor, cvt,
                       make_geometry(cfg));
  else
    hipLaunchKernelGGL(state_screen_kernel<float>, dim3(nb), dim3(kThreads), 0, st,
                       reinterpret_cast<const float*>(u), flag, g.n3, g.n2, g.n1, g.ng,
                       (float)cfg->gammad, cfg->limiter, (float)cfg->density_floor,
                       (float)cfg->pressure_floor, (float)cvt, make_geometry(cfg));
  hipLaunchKernelGGL(state_finalize_kernel, dim3(1), dim3(1), 0, st, flag);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

// --------------------------------------------------- composed plain path ----

namespace {

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct HydroSpace {
  void *wlr, *f[3], *rg, *fp1, *vic;
  size_t bytes;
};

HydroSpace hydro_layout(const snapx_config* c, int dtype, char* base) {
  Grid g = make_grid(c);
  size_t es = dtype == SNAPX_F64 ? 8 : 4;
  size_t field = align256(es * 5 * (size_t)g.ncell);
  HydroSpace h;
  size_t off = 0;
  h.wlr = base + off;
  off += 2 * field;
  for (int d = 0; d < 3; ++d) {
    h.f[d] = base + off;
    off += field;
  }
  h.rg = base + off;
  off += align256(es * (size_t)g.ncell);
  h.fp1 = base + off;  // x1 face pressure (spherical-polar sources)
  off += align256(es * (size_t)g.ncell);
  h.vic = base + off;
  off += align256(implicit_workspace_bytes(c));
  h.bytes = off;
  return h;
}

}  // namespace

extern "C" size_t snapx_hydro_workspace_bytes(const snapx_config* cfg, int dtype) {
  if (validate_config(cfg)) return 0;
  return hydro_layout(cfg, dtype, nullptr).bytes;
}

template <typename T>
static int hydro_forward_impl(const snapx_config* c, double dt, void* u, void* w, void* du,
                              void* f1, void* f2, void* f3, void* ws, size_t bytes, int dtype,
                              hipStream_t st, const double* dtp = nullptr) {
  Grid g = make_grid(c);
  HydroSpace h = hydro_layout(c, dtype, reinterpret_cast<char*>(ws));
  SNAPX_CHECK_ARG(bytes >= h.bytes, "snapx_hydro_forward: workspace too small");
  void* F[3] = {f1 ? f1 : h.f[0], f2 ? f2 : h.f[1], f3 ? f3 : h.f[2]};
  int ns[3] = {g.n1, g.n2, g.n3};
  int dims[3] = {3, 2, 1};
  const void* Fuse[3] = {nullptr, nullptr, nullptr};
  const void* rg = nullptr;
  // curvilinear blocks (spherical-polar, gnomonic panel): geometric sources
  // with the x1 face pressure
  const bool sph = c->coord != SNAPX_COORD_CARTESIAN;
  const Geometry geo = make_geometry(c);

  // (1) primitives
  if (int e = launch_cons2prim<T>(c, u, w, st)) return e;

  for (int a = 0; a < 3; ++a) {
    if (ns[a] == 1) continue;
    int grav1 = a == 0 && c->grav1 != 0.;
    if (grav1) {
      hipLaunchKernelGGL(revise_ghost_kernel<T>, blocks_for(g.n2 * g.n3, kThreads),
                         dim3(kThreads), 0, st, as<T>(w), g.n1, g.n2 * g.n3, g.ncell, g.ng,
                         T(c->grav1), T(c->gammad), ghost_dz(c));
      SNAPX_LAUNCH_CHECK();
    }
    int r = a == 0 ? 0 : 1;
    if (int e = launch_recon<T>(w, h.wlr, 5, g.n3, g.n2, g.n1, dims[a], c->recon_type[r],
                                c->recon_scale[r], c->recon_shock[r], 1, c->limiter,
                                c->density_floor, c->pressure_floor, grav1, g.ng, st))
      return e;
    if (grav1 && c->non_hydrostatic < 1.) {
      hipLaunchKernelGGL(rho_grav_kernel<T>, blocks_for(g.ncell, kThreads), dim3(kThreads), 0,
                         st, reinterpret_cast<const T*>(h.wlr), as<T>(h.rg), g.n1, g.ncell, 5,
                         g.ng, T(g.dx1));
      SNAPX_LAUNCH_CHECK();
      rg = h.rg;
    }
    // the x1 face pressure feeds the curvilinear radial source; the
    // reference keeps it zero when the x1 flux is disabled
    void* fp = (a == 0 && sph) ? h.fp1 : nullptr;
    if (!c->disable_flux[a]) {
      T* wl = as<T>(h.wlr);
      if (int e = launch_riemann<T>(c->riemann[a], wl, wl + 5 * g.ncell, F[a], fp, g.ncell,
                                    dims[a], c->gammad, st, &geo))
        return e;
      Fuse[a] = F[a];
    } else {
      if (a == 0 ? f1 : (a == 1 ? f2 : f3)) {
        hipLaunchKernelGGL(zero_kernel<T>, blocks_for(5 * g.ncell, kThreads), dim3(kThreads), 0,
                           st, as<T>(F[a]), (int64_t)5 * g.ncell);
        SNAPX_LAUNCH_CHECK();
      }
      if (fp) {
        hipLaunchKernelGGL(zero_kernel<T>, blocks_for(g.ncell, kThreads), dim3(kThreads), 0, st,
                           as<T>(fp), (int64_t)g.ncell);
        SNAPX_LAUNCH_CHECK();
      }
    }
  }
  if (sph) {
    if (int e = launch_coord<T>(c, dt, 1, w, Fuse[0], Fuse[1], Fuse[2], h.fp1, rg, du, st, dtp))
      return e;
  } else if (int e = launch_du<T>(c, dt, Fuse[0], Fuse[1], Fuse[2], w, rg, du, st, dtp)) {
    return e;
  }
  if (c->implicit_scheme)
    return implicit_correction(c, dt, du, w, h.vic, implicit_workspace_bytes(c), dtype, st, dtp);
  return 0;
}

extern "C" int snapx_hydro_forward(const snapx_config* cfg, double dt, void* u, void* w, void* du,
                                   void* f1, void* f2, void* f3, void* workspace,
                                   size_t workspace_bytes, int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(u && w && du && workspace, "snapx_hydro_forward: null pointer");
  hipStream_t st = as_stream(stream);
  return dtype == SNAPX_F64 ? hydro_forward_impl<double>(cfg, dt, u, w, du, f1, f2, f3, workspace,
                                                         workspace_bytes, dtype, st)
                            : hydro_forward_impl<float>(cfg, dt, u, w, du, f1, f2, f3, workspace,
                                                        workspace_bytes, dtype, st);
}

extern "C" size_t snapx_rk3_workspace_bytes(const snapx_config* cfg, int dtype) {
  if (validate_config(cfg)) return 0;
  Grid g = make_grid(cfg);
  size_t es = dtype == SNAPX_F64 ? 8 : 4;
  size_t field = align256(es * 5 * (size_t)g.ncell);
  return 3 * field + snapx_hydro_workspace_bytes(cfg, dtype);
}

template <typename T>
static int rk3_step_impl(const snapx_config* c, double dt, void* u, void* w, void* ws,
                         size_t bytes, int dtype, hipStream_t st, const double* dtp = nullptr) {
  Grid g = make_grid(c);
  size_t field = align256(sizeof(T) * 5 * (size_t)g.ncell);
  char* base = reinterpret_cast<char*>(ws);
  T* u0 = reinterpret_cast<T*>(base);
  T* u1 = reinterpret_cast<T*>(base + field);
  T* du = reinterpret_cast<T*>(base + 2 * field);
  void* hws = base + 3 * field;
  size_t hbytes = bytes - 3 * field;
  size_t n = 5 * (size_t)g.ncell;
  // Shu-Osher SSP-RK3 stage weights of u <- w0 u0 + w1 u1 + w2 du
  static const double W[3][3] = {
      {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};
  static const char* const kStageRange[3] = {"snapx plain stage 0", "snapx plain stage 1",
                                             "snapx plain stage 2"};
  TraceRange step_range("snapx plain step");
  SNAPX_HIP(hipMemcpyAsync(u0, u, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  SNAPX_HIP(hipMemcpyAsync(u1, u, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  for (int s = 0; s < 3; ++s) {
    TraceRange stage_range(kStageRange[s]);
    if (int e = hydro_forward_impl<T>(c, dt, u, w, du, nullptr, nullptr, nullptr, hws, hbytes,
                                      dtype, st, dtp))
      return e;
    hipLaunchKernelGGL(rk_combine_kernel<T>, blocks_for(n, kThreads), dim3(kThreads), 0, st,
                       as<T>(u), u0, u1, du, T(W[s][0]), T(W[s][1]), T(W[s][2]), (int64_t)n);
    SNAPX_LAUNCH_CHECK();
    if (c->limiter) {
      // apply_conserved_limiter_ on the combined state (meshblock.cpp:676):
      // the cons2prim kernel applies it to u in place (w is rewritten by the
      // next stage anyway)
      if (int e = launch_cons2prim<T>(c, u, w, st)) return e;
    }
    if (int e = launch_bc<T>(c, u, 5, SNAPX_CONSERVED, st)) return e;
    if (s < 2) SNAPX_HIP(hipMemcpyAsync(u1, u, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
  }
  return 0;
}

extern "C" int snapx_rk3_step(const snapx_config* cfg, double dt, void* u, void* w,
                              void* workspace, size_t workspace_bytes, int dtype, void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(u && w && workspace, "snapx_rk3_step: null pointer");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_rk3_workspace_bytes(cfg, dtype),
                  "snapx_rk3_step: workspace too small");
  hipStream_t st = as_stream(stream);
  return dtype == SNAPX_F64
             ? rk3_step_impl<double>(cfg, dt, u, w, workspace, workspace_bytes, dtype, st)
             : rk3_step_impl<float>(cfg, dt, u, w, workspace, workspace_bytes, dtype, st);
}

extern "C" int snapx_rk3_step_dt(const snapx_config* cfg, const double* dt, void* u, void* w,
                                 void* workspace, size_t workspace_bytes, int dtype,
                                 void* stream) {
  if (int e = validate_config(cfg)) return e;
  SNAPX_CHECK_ARG(dt && u && w && workspace, "snapx_rk3_step_dt: null pointer");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_rk3_workspace_bytes(cfg, dtype),
                  "snapx_rk3_step_dt: workspace too small");
  hipStream_t st = as_stream(stream);
  return dtype == SNAPX_F64
             ? rk3_step_impl<double>(cfg, 0., u, w, workspace, workspace_bytes, dtype, st, dt)
             : rk3_step_impl<float>(cfg, 0., u, w, workspace, workspace_bytes, dtype, st, dt);
}
