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

#include "snapx.h"

namespace snapx {

struct Axis {
  double lo, hi, step;  // torch::linspace(lo, hi, npts), step = (hi - lo) / (npts - 1)
  int npts;             // cells + 1
};

struct Geometry {
  int coord;  // SNAPX_COORD_*
  int n1, n2, n3, ng;
  Axis ax[3];  // x1, x2, x3
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

}  // namespace snapx
