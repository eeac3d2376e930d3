// Coordinate geometry of a block for the plain and fused paths, on host and device.
//
// Faces are torch::linspace(lo, hi, nc + 1) over the block including its
// ghost cells (CoordinateImpl constructor, src/coord/coordinate.cpp:201-221).
// torch's CPU linspace in double counts up from lo for the first half of the
// points and down from hi for the rest, each point with one rounding
// (a fused multiply-add); axis_face reproduces that exactly.
//
// Cartesian blocks use widths only.  Spherical-polar (r, theta, phi) blocks
// add the volume-weighted centres, face areas, cell volumes and source
// coefficients of SphericalPolarImpl (src/coord/spherical_polar.cpp:15-144,
// 162-230).  Everything is evaluated per cell from the face coordinates, so
// no geometry arrays live in device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "common.hpp"
#include "snapx.h"

namespace snapx {

struct Axis {
  double lo, hi, step;  // torch::linspace(lo, hi, npts), step = (hi - lo) / (npts - 1)
  int npts;             // cells + 1
};

struct Geometry {
  int coord;  // SNAPX_COORD_*
  int n1, n2, n3, ng;
  Axis ax[3];    // x1, x2, x3
  double dx[3];  // (xmax - xmin) / nx: the uniform widths dx?f of the gnomonic panel
};

__host__ __device__ inline double axis_face(const Axis& a, int i) {
  return i < a.npts / 2 ? fma(a.step, (double)i, a.lo) : fma(-a.step, (double)(a.npts - 1 - i), a.hi);
}

__host__ __device__ inline double axis_width(const Axis& a, int i) {
  return axis_face(a, i + 1) - axis_face(a, i);
}

// src/coord/coordinate.cpp:201-221: ghost cells extend the axis by nghost
// widths on both sides when it has more than one cell
inline Axis make_axis(double xmin, double xmax, int nx, int ng) {
  double dx = (xmax - xmin) / nx;
  Axis a;
  a.lo = nx > 1 ? xmin - ng * dx : xmin;
  a.hi = nx > 1 ? xmax + ng * dx : xmax;
  a.npts = (nx > 1 ? nx + 2 * ng : 1) + 1;
  a.step = (a.hi - a.lo) / (a.npts - 1);
  return a;
}

inline Geometry make_geometry(const snapx_config* c) {
  Geometry g;
  g.coord = c->coord;
  g.ng = c->nghost;
  g.ax[0] = make_axis(c->x1min, c->x1max, c->nx1, c->nghost);
  g.ax[1] = make_axis(c->x2min, c->x2max, c->nx2, c->nghost);
  g.ax[2] = make_axis(c->x3min, c->x3max, c->nx3, c->nghost);
  g.n1 = g.ax[0].npts - 1, g.n2 = g.ax[1].npts - 1, g.n3 = g.ax[2].npts - 1;
  g.dx[0] = (c->x1max - c->x1min) / c->nx1;
  g.dx[1] = (c->x2max - c->x2min) / c->nx2;
  g.dx[2] = (c->x3max - c->x3min) / c->nx3;
  return g;
}

// ---- spherical-polar (src/coord/spherical_polar.cpp) ----

// radial_centers (:17-21): volume-weighted radius of the shell [rm, rp]
__host__ __device__ inline double sph_x1v(const Geometry& g, int i) {
  double rm = axis_face(g.ax[0], i), rp = axis_face(g.ax[0], i + 1);
  return 0.75 * (pow(rp, 4) - pow(rm, 4)) / (pow(rp, 3) - pow(rm, 3));
}

// polar_centers (:23-28), or the midpoint when x2 has one cell (:64-67)
__host__ __device__ inline double sph_x2v(const Geometry& g, int j) {
  double tm = axis_face(g.ax[1], j), tp = axis_face(g.ax[1], j + 1);
  if (g.n2 == 1) return 0.5 * (tm + tp);
  return ((sin(tp) - tp * cos(tp)) - (sin(tm) - tm * cos(tm))) / (cos(tm) - cos(tp));
}

// center_distances (:30-38) of x1: dx1v[0] = dx1f[0], dx1v[i] = x1v[i] - x1v[i-1]
__host__ __device__ inline double sph_dx1v(const Geometry& g, int i) {
  return i == 0 ? axis_width(g.ax[0], 0) : sph_x1v(g, i) - sph_x1v(g, i - 1);
}

// |cos(theta_m) - cos(theta_p)| of cell j (the polar factor of areas and volumes)
__host__ __device__ inline double sph_polar(const Geometry& g, int j) {
  return fabs(cos(axis_face(g.ax[1], j)) - cos(axis_face(g.ax[1], j + 1)));
}

// face_area1 (:176-181) at x1 face i of row (k, j)
__host__ __device__ inline double sph_area1(const Geometry& g, int k, int j, int i) {
  double r = axis_face(g.ax[0], i);
  return r * r * sph_polar(g, j) * axis_width(g.ax[2], k);
}

// 0.5 (rp^2 - rm^2) of cell i
__host__ __device__ inline double sph_radial_area(const Geometry& g, int i) {
  double rm = axis_face(g.ax[0], i), rp = axis_face(g.ax[0], i + 1);
  return 0.5 * (rp * rp - rm * rm);
}

// (rp^3 - rm^3) / 3 of cell i
__host__ __device__ inline double sph_radial_volume(const Geometry& g, int i) {
  double rm = axis_face(g.ax[0], i), rp = axis_face(g.ax[0], i + 1);
  return (pow(rp, 3) - pow(rm, 3)) / 3.0;
}

// face_area2 (:183-189) at x2 face j of cell (k, ., i)
__host__ __device__ inline double sph_area2(const Geometry& g, int k, int j, int i) {
  return sph_radial_area(g, i) * fabs(sin(axis_face(g.ax[1], j))) * axis_width(g.ax[2], k);
}

// face_area3 (:191-196): independent of k
__host__ __device__ inline double sph_area3(const Geometry& g, int j, int i) {
  return sph_radial_area(g, i) * axis_width(g.ax[1], j);
}

// cell_volume (:198-210)
__host__ __device__ inline double sph_volume(const Geometry& g, int k, int j, int i) {
  return sph_radial_volume(g, i) * sph_polar(g, j) * axis_width(g.ax[2], k);
}

// coord_src1_i, coord_src2_i (:88-93)
__host__ __device__ inline double sph_src1_i(const Geometry& g, int i) {
  return sph_radial_area(g, i) / sph_radial_volume(g, i);
}
__host__ __device__ inline double sph_src2_i(const Geometry& g, int i) {
  double rm = axis_face(g.ax[0], i), rp = axis_face(g.ax[0], i + 1);
  return axis_width(g.ax[0], i) / ((rm + rp) * sph_radial_volume(g, i));
}

// coord_src1_j (= coord_src3_j) and coord_src2_j (:94-104)
__host__ __device__ inline double sph_src1_j(const Geometry& g, int j) {
  double sm = fabs(sin(axis_face(g.ax[1], j))), sp = fabs(sin(axis_face(g.ax[1], j + 1)));
  return (sp - sm) / sph_polar(g, j);
}
__host__ __device__ inline double sph_src2_j(const Geometry& g, int j) {
  double sm = fabs(sin(axis_face(g.ax[1], j))), sp = fabs(sin(axis_face(g.ax[1], j + 1)));
  return (sp - sm) / ((sm + sp) * sph_polar(g, j));
}

// Everything one cell needs from the coordinate, from its six face
// coordinates with one sine / cosine pair per polar face (the functions above
// evaluate each factor on its own): the volume, the areas of its faces, and
// the source coefficients of its row and layer.
struct SphCell {
  double vol, a1m, a1p, a2m, a2p, a3, src1_i, src2_i, src1_j, src2_j, dx1;
};

__host__ __device__ inline SphCell sph_cell(const Geometry& g, int k, int j, int i) {
  const double rm = axis_face(g.ax[0], i), rp = axis_face(g.ax[0], i + 1);
  const double tm = axis_face(g.ax[1], j), tp = axis_face(g.ax[1], j + 1);
  const double dphi = axis_width(g.ax[2], k);
  double sm, cm, sp, cp;
  sincos(tm, &sm, &cm);
  sincos(tp, &sp, &cp);
  sm = fabs(sm), sp = fabs(sp);
  const double polar = fabs(cm - cp);
  const double vr = (rp * rp * rp - rm * rm * rm) / 3.0, ra = 0.5 * (rp * rp - rm * rm);
  SphCell c;
  c.vol = vr * polar * dphi;
  c.a1m = rm * rm * polar * dphi;
  c.a1p = rp * rp * polar * dphi;
  c.a2m = ra * sm * dphi;
  c.a2p = ra * sp * dphi;
  c.a3 = ra * (tp - tm);
  c.src1_i = ra / vr;
  c.src2_i = (rp - rm) / ((rm + rp) * vr);
  c.src1_j = (sp - sm) / polar;
  c.src2_j = (sp - sm) / ((sm + sp) * polar);
  c.dx1 = rp - rm;
  return c;
}

// center_width2 / center_width3 (:146-155), the CFL lengths of x2 and x3
__host__ __device__ inline double sph_width2(const Geometry& g, int j, int i) {
  return sph_x1v(g, i) * axis_width(g.ax[1], j);
}
__host__ __device__ inline double sph_width3(const Geometry& g, int k, int j, int i) {
  return sph_x1v(g, i) * fabs(sin(sph_x2v(g, j))) * axis_width(g.ax[2], k);
}

// ---- gnomonic-equiangle panel (src/coord/gnomonic_equiangle.cpp) ----
//
// x1 = r with uniform widths dx1f, centres at the face midpoints; x2, x3 are
// equiangular coordinates (radians) of one cube face, a point on the unit
// sphere being (1, x, y) / |.| with x = tan x2, y = tan x3.  The x2 and x3
// coordinate lines meet at an angle theta with cos = -x y / (C D) and
// sin = sqrt(1 + x^2 + y^2) / (C D), C = sqrt(1 + x^2), D = sqrt(1 + y^2)
// (:60-77); the angular widths of a cell are great-circle arcs (:79-110).
// Every expression keeps the reference's evaluation order.

__host__ __device__ inline double gno_center(const Axis& a, int j) {
  return 0.5 * (axis_face(a, j) + axis_face(a, j + 1));
}

struct GnoAngle {
  double c, s;  // cos, sin of theta
};

__host__ __device__ inline GnoAngle gno_angle(double x, double y) {
  const double C = sqrt(1.0 + x * x), D = sqrt(1.0 + y * y);
  return {-x * y / (C * D), sqrt(1.0 + x * x + y * y) / (C * D)};
}

// dx2f_ang: arc between the x2 lines x1 = tan x2f[j], x2 = tan x2f[j + 1] at y
__host__ __device__ inline double gno_ang2(double x1, double x2, double y) {
  const double d1 = sqrt(1. + x1 * x1 + y * y), d2 = sqrt(1. + x2 * x2 + y * y);
  return acos((1. + x1 * x2 + y * y) / (d1 * d2));
}

// dx3f_ang: arc between the x3 lines y1, y2 at x
__host__ __device__ inline double gno_ang3(double x, double y1, double y2) {
  const double d1 = sqrt(1. + x * x + y1 * y1), d2 = sqrt(1. + x * x + y2 * y2);
  return acos((1. + x * x + y1 * y2) / (d1 * d2));
}

// cosine_cell_kj of row (k, j)
__host__ __device__ inline double gno_cos_cell(const Geometry& g, int k, int j) {
  return gno_angle(tan(gno_center(g.ax[1], j)), tan(gno_center(g.ax[2], k))).c;
}

// The angle at the face of dim (3: x1 faces, the cell's angle; 2: the lower
// x2 face of column j; 1: the lower x3 face of row k), for the face frames
__host__ __device__ inline GnoAngle gno_face_angle(const Geometry& g, int dim, int k, int j) {
  const double x = dim == 2 ? tan(axis_face(g.ax[1], j)) : tan(gno_center(g.ax[1], j));
  const double y = dim == 1 ? tan(axis_face(g.ax[2], k)) : tan(gno_center(g.ax[2], k));
  return gno_angle(x, y);
}

// Everything one cell (k, j, i) needs: face areas (:161-176), volume
// (:178-184), the metric source coefficients x_ov_rD, y_ov_rC (:112-118),
// the cell's angle and radius
struct GnoCell {
  double vol, a1m, a1p, a2m, a2p, a3m, a3p, xov, yov, cos, sin, r, dx1;
};

__host__ __device__ inline GnoCell gno_cell(const Geometry& g, int k, int j, int i) {
  const double rm = axis_face(g.ax[0], i), rp = axis_face(g.ax[0], i + 1);
  const double dr = g.dx[0], rv = 0.5 * (rm + rp);
  const double xm = tan(axis_face(g.ax[1], j)), xp = tan(axis_face(g.ax[1], j + 1));
  const double ym = tan(axis_face(g.ax[2], k)), yp = tan(axis_face(g.ax[2], k + 1));
  const double x = tan(gno_center(g.ax[1], j)), y = tan(gno_center(g.ax[2], k));
  const GnoAngle cc = gno_angle(x, y);
  const double ang = gno_ang2(xm, xp, y) * gno_ang3(x, ym, yp) * cc.s;
  GnoCell c;
  c.a1m = rm * rm * ang;
  c.a1p = rp * rp * ang;
  c.vol = 0.5 * (c.a1m + c.a1p) * dr;
  const double ring = rv * dr;
  c.a2m = ring * gno_ang3(xm, ym, yp);
  c.a2p = ring * gno_ang3(xp, ym, yp);
  c.a3m = ring * gno_ang2(xm, xp, ym);
  c.a3p = ring * gno_ang2(xm, xp, yp);
  c.xov = (c.a2p * gno_angle(xp, y).s - c.a2m * gno_angle(xm, y).s) / c.vol;
  c.yov = (c.a3p * gno_angle(x, yp).s - c.a3m * gno_angle(x, ym).s) / c.vol;
  c.cos = cc.c, c.sin = cc.s;
  c.r = rv;
  c.dx1 = dr;
  return c;
}

// center_width2 / center_width3 (:157-163)
__host__ __device__ inline double gno_width2(const Geometry& g, int k, int j, int i) {
  const double rv = gno_center(g.ax[0], i);
  return rv * gno_ang2(tan(axis_face(g.ax[1], j)), tan(axis_face(g.ax[1], j + 1)),
                       tan(gno_center(g.ax[2], k)));
}
__host__ __device__ inline double gno_width3(const Geometry& g, int k, int j, int i) {
  const double rv = gno_center(g.ax[0], i);
  return rv * gno_ang3(tan(gno_center(g.ax[1], j)), tan(axis_face(g.ax[2], k)),
                       tan(axis_face(g.ax[2], k + 1)));
}

// GnomonicEquiangleImpl::prim2local{1,2,3}_ (:236-279): the tangential
// velocities (v2, v3) of an interface state in the orthonormal frame of a
// face of dim (3: x1, 2: x2, 1: x3) with angle (c, s)
template <typename T>
__host__ __device__ inline void gno_prim2local(T* w, int dim, T c, T s) {
  const T a = w[IVY], b = w[IVZ];
  if (dim == 3) {
    w[IVY] = a + b * c;
    w[IVZ] = b * s;
  } else if (dim == 2) {
    w[IVZ] = c * a + b;
    w[IVY] = (T(1.0) / sqrt(T(1.) / (s * s))) * a;
  } else {
    w[IVY] = a + c * b;
    w[IVZ] = (T(1.0) / sqrt(T(1.) / (s * s))) * b;
  }
}

// GnomonicEquiangleImpl::flux2global{1,2,3}_ (:281-360): face-frame momentum
// fluxes back to the panel's covariant components
template <typename T>
__host__ __device__ inline void gno_flux2global(T* f, int dim, T c, T s) {
  T ty, tz;
  if (dim == 3) {
    tz = f[IVZ] / s;
    ty = f[IVY] - tz * c;
  } else if (dim == 2) {
    const T t22 = sqrt(T(1.) / (s * s)), t32 = -t22 * c / T(1.0);
    tz = t32 * f[IVY] + f[IVZ];
    ty = t22 * f[IVY];
  } else {
    const T t33 = sqrt(T(1.) / (s * s)), t23 = -c / T(1.0) * t33;
    ty = f[IVY] + t23 * f[IVZ];
    tz = t33 * f[IVZ];
  }
  f[IVY] = ty + tz * c;
  f[IVZ] = tz + ty * c;
}

}  // namespace snapx
