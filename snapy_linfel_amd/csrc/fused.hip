// Fused FP32 SSP-RK3 stage on MI355X.
//
// Per stage two kernels replace the ~15 module launches of the plain path:
//
//  K1 fused_rhs_kernel (FP32): one workgroup per (x1, x2, x3) tile.  The tile
//     plus a 3-cell halo of the conserved state is read from HBM once
//     (coalesced along x1), converted to primitives into LDS, the x1 ghost
//     cells get the isentropic hydrostatic refill, and then for each
//     direction the interface states (WENO5/CP5), Riemann fluxes and the flux
//     differences are computed from LDS; interface states and fluxes never
//     leave the chip.  Output: the explicit tendency du (dt included) of the
//     tile's cells, gravity added.
//  K2 fused_column_pcr_kernel (FP64 arithmetic): one lane per cell, whole
//     columns per workgroup.  Rebuilds the primitives in FP64, assembles the
//     block rows (interface |A| shared through LDS), solves the
//     block-tridiagonal system by parallel cyclic reduction, applies the
//     MS-VIC redistribution (column sums by wave reductions) and the SSP-RK3
//     stage combine; writes the new conserved state.  Columns taller than
//     the workgroup use fused_column_kernel (one lane per column, rolling
//     interface Jacobians, serial block Thomas sweep).
//
// Ghost cells are never read during the step: halo loads map outside indices
// through the external boundary functions (reflecting: mirror and flip the
// normal momentum; outflow: copy the edge cell; periodic: wrap), which gives
// the values the reference's boundary pass would have written.  One boundary
// kernel at the end of the step fills the stored ghosts.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "geom_impl.hpp"
#include "interp_impl.hpp"
#include "riemann_impl.hpp"
#include "vic_impl.hpp"
#include "pcr.hpp"
#include "bg_impl.hpp"
#include "riemann_pert.hpp"

namespace snapx {

// implemented in plain.hip
int launch_bc_f32(const snapx_config* c, float* var, hipStream_t st);

constexpr int kRhsThreads = 256;
constexpr int kMaxCellsPerThread = 4;  // tile cells <= 1024
constexpr int kColThreads = 64;

// Isentropic hydrostatic refill of the x1 ghosts (src/hydro/hydro.cpp:231-269)
// in FP32.  The reference steps p^a, a = (gamma - 1) / gamma, outward one
// layer at a time, p_n^a = p_{n-1}^a +- a g dz / K^(1/gamma) with
// K = p_b / rho_b^gamma, and sets rho_n = (p_n / K)^(1/gamma).  In exact
// arithmetic that is h_n = p_n^a = (p_b +- n a g dz rho_b) / P with
// P = p_b^(1/gamma), p_n = h_n^(1/a) = h_n H_n and rho_n = rho_b H_n / P with
// H_n = h_n^(1/(gamma - 1)): one powf per column end and one per layer
// instead of two per end and three per layer, and the layers no longer wait
// on each other.  (The FP64 plain path keeps the reference's order.)
struct IsentropicRefill {
  float rho_b, p_b, inv_P, step, inv_gm;
  // gdz = -grav1 * dx1 (positive for gravity pointing to -x1); dir -1: the
  // ghosts below the first interior cell, +1: above the last one
  __device__ __forceinline__ IsentropicRefill(float rho, float p, float gamma, float gdz,
                                              int dir)
      : rho_b(rho), p_b(p) {
    float gm = gamma - 1.f;
    inv_P = 1.f / powf(p, 1.f / gamma);
    step = (dir < 0 ? 1.f : -1.f) * (gm / gamma) * gdz * rho;
    inv_gm = 1.f / gm;
  }
  __device__ __forceinline__ void layer(int n, float& rho, float& p) const {
    float h = (p_b + (float)n * step) * inv_P;
    float H = powf(h, inv_gm);
    p = h * H;
    rho = rho_b * H * inv_P;
  }
  // layers of uneven thickness (spherical-polar x1 centres): built with
  // gdz = -grav1 (per unit length), z is the summed distance of the layer
  // from the boundary cell, sum_m dx1v of the layers in between
  __device__ __forceinline__ void layer_at(float z, float& rho, float& p) const {
    float h = (p_b + z * step) * inv_P;
    float H = powf(h, inv_gm);
    p = h * H;
    rho = rho_b * H * inv_P;
  }
};

struct FusedArgs {
  int n1, n2, n3, ng;
  int nx1, nx2, nx3;
  int tx1, tx2, tx3;  // tile extents (interior cells)
  int h2, h3;         // halo widths along x2 / x3 (3 or 0)
  int nt1, nt2, nt3;  // tiles per axis
  int recon_type[2], recon_scale[2], recon_shock[2];
  int riemann[3];
  int bc[6];
  int merged;  // faces of all directions in one pass (small tiles)
  float gammad, dx1, dx2, dx3, dt, grav1, grav2, grav3;
  // spherical-polar blocks: summed ghost-layer distances below / above
  // (dx1v[is - 1], + dx1v[is - 2], ...; dx1v[ie], + dx1v[ie + 1], ...)
  float zin[3], zout[3];
  Geometry geo;
  // hydrostatic background of the perturbation path (null: total state);
  // tables of bg_impl.hpp
  const float* bgt;
  const double* bgd;
  // device time step (double); replaces dt when set
  const double* dtp;
};

// Maps an interior-relative index x (may lie in the ghost range) of an axis
// with nx interior cells through that axis' boundary functions.  Returns the
// source index and sets flip when the normal momentum changes sign.
__device__ __forceinline__ int bc_map(int x, int nx, int bin, int bout, bool& flip) {
  if (x < 0) {
    if (bin == SNAPX_BC_REFLECTING) {
      flip = !flip;
      return -1 - x;
    }
    if (bin == SNAPX_BC_PERIODIC) return x + nx;
    return 0;  // outflow (and no boundary function)
  }
  if (x >= nx) {
    if (bout == SNAPX_BC_REFLECTING) {
      flip = !flip;
      return 2 * nx - 1 - x;
    }
    if (bout == SNAPX_BC_PERIODIC) return x - nx;
    return nx - 1;
  }
  return x;
}

// offset within an x1-x2 plane of the full array of interior-relative
// (i, j), mapping through the x1 / x2 boundaries (flips[0], flips[1])
__device__ __forceinline__ int plane_cell(const FusedArgs& a, int i, int j, bool* flips) {
  flips[0] = flips[1] = false;
  int lim = a.nx1 + a.ng - 1;
  i = i > lim ? lim : i;
  i = bc_map(i, a.nx1, a.bc[0], a.bc[1], flips[0]);
  int qj = 0;
  if (a.n2 > 1) {
    int l2 = a.nx2 + a.ng - 1;
    j = j > l2 ? l2 : j;
    j = bc_map(j, a.nx2, a.bc[2], a.bc[3], flips[1]);
    qj = j + a.ng;
  }
  return qj * a.n1 + i + a.ng;
}

// full-array index of the first cell of x3 slice k (interior-relative),
// mapped through the x3 boundaries (flip: the x3 momentum changes sign)
__device__ __forceinline__ int slice_start(const FusedArgs& a, int k, bool& flip) {
  flip = false;
  if (a.n3 <= 1) return 0;
  int l3 = a.nx3 + a.ng - 1;
  k = k > l3 ? l3 : k;
  k = bc_map(k, a.nx3, a.bc[4], a.bc[5], flip);
  return (k + a.ng) * a.n2 * a.n1;
}

// cell index in the full array of interior-relative (i, j, k), mapping
// through the boundaries; flips[a] reports a sign change of the axis-a
// momentum
__device__ __forceinline__ int mapped_cell(const FusedArgs& a, int i, int j, int k, bool* flips) {
  const int p = plane_cell(a, i, j, flips);
  return slice_start(a, k, flips[2]) + p;
}

// reciprocals of the fused kernels: the FP32 hardware reciprocal (this unit
// is built without correctly rounded FP32 division) and rcp_fast in FP64
__device__ __forceinline__ float recip(float x) { return 1.f / x; }
__device__ __forceinline__ double recip(double x) { return rcp_fast(x); }

template <typename T>
__device__ __forceinline__ void cons_to_prim(const T* U, T gammad, T* W) {
  W[IDN] = U[IDN];
  T inv = recip(U[IDN]);
  W[IVX] = U[IVX] * inv;
  W[IVY] = U[IVY] * inv;
  W[IVZ] = U[IVZ] * inv;
  T ke = T(0.5) * ((W[IVX] * U[IVX] + W[IVY] * U[IVY]) + W[IVZ] * U[IVZ]);
  W[IPR] = (gammad - T(1.)) * (U[IPR] - ke);
}

// cons_to_prim on a gnomonic panel: the covariant momenta over the density
// raised to contravariant velocities with the cell's cos(theta) before the
// kinetic energy (src/eos/ideal_gas_impl.h:15-35, coord_utils_impl.h:8-16)
template <typename T>
__device__ __forceinline__ void cons_to_prim_gno(const T* U, T gammad, T cth, T* W) {
  W[IDN] = U[IDN];
  T inv = recip(U[IDN]);
  W[IVX] = U[IVX] * inv;
  const T v = U[IVY] * inv, x = U[IVZ] * inv, sth2 = T(1.) - cth * cth;
  W[IVY] = v / sth2 - x * cth / sth2;
  W[IVZ] = -v * cth / sth2 + x / sth2;
  T ke = T(0.5) * ((W[IVX] * U[IVX] + W[IVY] * U[IVY]) + W[IVZ] * U[IVZ]);
  W[IPR] = (gammad - T(1.)) * (U[IPR] - ke);
}

// The face frame of a gnomonic panel on an interface state in the RHS
// kernel's slot order (the normal velocity in slot IVX; d = 0, 1, 2 for x1,
// x2, x3 faces): gno_prim2local on the panel velocities (v2, v3)
__device__ __forceinline__ void gno_slots_to_local(float* w, int d, float c, float s) {
  float g[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  const int i2 = d == 0 ? IVY : (d == 1 ? IVX : IVZ), i3 = d == 0 ? IVZ : (d == 1 ? IVY : IVX);
  g[IVY] = w[i2], g[IVZ] = w[i3];
  gno_prim2local(g, d == 0 ? 3 : (d == 1 ? 2 : 1), c, s);
  w[i2] = g[IVY], w[i3] = g[IVZ];
}
__device__ __forceinline__ void gno_slots_to_global(float* f, int d, float c, float s) {
  float g[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  const int i2 = d == 0 ? IVY : (d == 1 ? IVX : IVZ), i3 = d == 0 ? IVZ : (d == 1 ? IVY : IVX);
  g[IVY] = f[i2], g[IVZ] = f[i3];
  gno_flux2global(g, d == 0 ? 3 : (d == 1 ? 2 : 1), c, s);
  f[i2] = g[IVY], f[i3] = g[IVZ];
}


// the step's dt: read from device memory when the caller passed a pointer
// (the kernel argument itself stays unmodified, so that it is not copied to
// scratch memory)
__device__ __forceinline__ float step_dt(const FusedArgs& a) {
  return a.dtp ? (float)*a.dtp : a.dt;
}

// ---- perturbation path (hydrostatic background, bg_impl.hpp) ----

// FP32 background tables as seen by one kernel, in the layout of BgLayout:
// five rows over nl levels (rho_c, p_c, gdr, gdp, res; index = level -
// lbase), then ten rows over nf faces (rho_f, p_f, three WENO offsets below,
// three above, the two CP5 offsets; index = face - fbase).  Either the
// global tables (lbase = fbase = 0) or the slice a tile copied to LDS.
struct BgView {
  const float* t;
  int nl, nf, lbase, fbase;
  __device__ __forceinline__ float lv(int row, int level) const {
    return t[row * nl + level - lbase];
  }
  __device__ __forceinline__ float fc(int row, int face) const {
    return t[kBgLevelRows * nl + row * nf + face - fbase];
  }
};
enum { BG_RHO_C = 0, BG_P_C, BG_GDR, BG_GDP, BG_RES, BG_CN, BG_QK, BG_QK2, BG_DPG, BG_DRG };
enum { BG_RHO_F = 0, BG_P_F, BG_DRL, BG_DRR = 5, BG_EPL = 8, BG_EPR };

__device__ __forceinline__ BgView bg_global_view(const float* t, const BgLayout& B) {
  return BgView{t, B.L, B.F, 0, 0};
}

// Copies the tables of levels lbase .. lbase + nl - 1 and faces fbase ..
// fbase + nf - 1 into LDS at s (kBgLevelRows nl + kBgFaceRows nf floats;
// the caller synchronises before any read) and returns the view of the copy.
__device__ __forceinline__ BgView bg_stage_lds(const float* t, const BgLayout& B, int lbase,
                                               int nl, int fbase, int nf, float* s) {
  for (int k = threadIdx.x; k < kBgLevelRows * nl; k += blockDim.x) {
    const int r = k / nl, l = k - r * nl;
    s[k] = t[r * B.L + lbase + l];
  }
  for (int k = threadIdx.x; k < kBgFaceRows * nf; k += blockDim.x) {
    const int r = k / nf, f = k - r * nf;
    s[kBgLevelRows * nl + k] = t[kBgLevelRows * B.L + r * B.F + fbase + f];
  }
  return BgView{s, nl, nf, lbase, fbase};
}

// Primitive perturbation (rho', v1, v2, v3, p') of a loaded cell U' (flips
// applied) at the LDS position of x1 level lg, and the total density.  U' is
// relative to its source level; a ghost position (mirror or edge source)
// adds the background difference of the source and ghost levels.
__device__ __forceinline__ void bg_load_prim(const FusedArgs& a, const BgView& V,
                                             const float* U, int lg, float* W, float& rtot) {
  W[IDN] = U[IDN] + V.lv(BG_GDR, lg);
  rtot = V.lv(BG_RHO_C, lg) + W[IDN];
  const float inv = 1.f / rtot;
  W[IVX] = U[IVX] * inv;
  W[IVY] = U[IVY] * inv;
  W[IVZ] = U[IVZ] * inv;
  const float ke = 0.5f * ((W[IVX] * U[IVX] + W[IVY] * U[IVY]) + W[IVZ] * U[IVZ]);
  W[IPR] = (a.gammad - 1.f) * (U[IPR] - ke) + V.lv(BG_GDP, lg);
}

// Density and pressure perturbations of the boundary cell U' (conserved
// perturbation of an interior cell; the velocities need the total density)
__device__ __forceinline__ void bg_cell_rho_p(const BgView& V, const float* U, int lb,
                                              float gammad, float& rp, float& pp) {
  rp = U[IDN];
  const float rt = V.lv(BG_RHO_C, lb) + rp;
  const float ke = 0.5f * ((U[IVX] * U[IVX] + U[IVY] * U[IVY]) + U[IVZ] * U[IVZ]) / rt;
  pp = (gammad - 1.f) * (U[IPR] - ke);
}

// isentropic refill of the physical x1 ghost at x1 level lg from the
// boundary cell U' of level lb (bg_refill_pert): density and pressure
// perturbations against the ghost level
__device__ __forceinline__ void bg_refill_ghost(const BgView& V, const float* U, int lb, int lg,
                                                float gammad, float& rg, float& pg) {
  float rp, pp;
  bg_cell_rho_p(V, U, lb, gammad, rp, pp);
  bg_refill_pert(V.lv(BG_RHO_C, lb), V.lv(BG_P_C, lb), rp, pp, V.lv(BG_CN, lg), V.lv(BG_QK, lg),
                 V.lv(BG_QK2, lg), V.lv(BG_DPG, lg), V.lv(BG_DRG, lg), gammad, rg, pg);
}

// the refill of the tile's own x1 ghost at interior-relative index gi from
// the boundary cell qb, into W and the total density
__device__ __forceinline__ void bg_refill_prim(const FusedArgs& a, const BgView& V,
                                               const float* __restrict__ u, int ncell, int qb,
                                               int gi, float* W, float& rtot) {
  const bool upper = gi >= a.nx1;
  const int lb = upper ? a.ng + a.nx1 - 1 : a.ng, lg = gi + a.ng;
  float U[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) U[v] = u[v * ncell + qb];
  bg_refill_ghost(V, U, lb, lg, a.gammad, W[IDN], W[IPR]);
  rtot = V.lv(BG_RHO_C, lg) + W[IDN];
}

// Flux of an x2 / x3 face (perturbation form) from the perturbation states
// wl, wr in the face frame: the reference is the background of level lv.
__device__ __forceinline__ void bg_horizontal_flux(const BgView& V, int lv, int solver, float* wl,
                                                   float* wr, float gammad, float* flx) {
  const float rb = V.lv(BG_RHO_C, lv), pb = V.lv(BG_P_C, lv);
  const float rl = wl[IDN], rr = wr[IDN], pl = wl[IPR], pr = wr[IPR];
  wl[IDN] += rb;
  wr[IDN] += rb;
  wl[IPR] += pb;
  wr[IPR] += pb;
  riemann_flux_pert<float>(solver, wl, wr, rl, rr, pl, pr, pb, gammad, flx);
}

// Flux of the x1 face below LDS cell c of a pencil (start ws + base, stride
// 1) in perturbation form: WENO5 density from the total-density plane for
// the weights plus the background offsets, CP5 pressure perturbation plus
// its offset, CP5 velocities; face reference (rho_f, p_f) of face level fl.
// rev_lo / rev_hi: physical bottom / top face with gravity (the reference's
// LR revision copies density and pressure across the face).
__device__ __forceinline__ void bg_x1_face_flux(const float* ws, int HS, int base, int c, int fl,
                                                const BgView& V, bool rev_lo, bool rev_hi,
                                                int solver, float gammad, float* flx) {
  const float* rp = ws + IDN * HS + base;
  const float* rt = ws + 5 * HS + base;
  const float* pp = ws + IPR * HS + base;
  float tw[6], fw[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    tw[k] = rt[c - 3 + k];
    fw[k] = rp[c - 3 + k];
  }
  const float dl[3] = {V.fc(BG_DRL, fl), V.fc(BG_DRL + 1, fl), V.fc(BG_DRL + 2, fl)};
  const float dr[3] = {V.fc(BG_DRR, fl), V.fc(BG_DRR + 1, fl), V.fc(BG_DRR + 2, fl)};
  float rl = weno5_off<true>(tw, fw, dl);
  float rr = weno5_off<false>(tw + 1, fw + 1, dr);
  float pl = cp5_hi(pp[c - 3], pp[c - 2], pp[c - 1], pp[c], pp[c + 1]) + V.fc(BG_EPL, fl);
  float pr = cp5_lo(pp[c - 2], pp[c - 1], pp[c], pp[c + 1], pp[c + 2]) + V.fc(BG_EPR, fl);
  if (rev_lo) rl = rr, pl = pr;
  if (rev_hi) rr = rl, pr = pl;
  const float rf = V.fc(BG_RHO_F, fl), pf = V.fc(BG_P_F, fl);
  float wl[5], wr[5];
  wl[IDN] = rf + rl;
  wr[IDN] = rf + rr;
  wl[IPR] = pf + pl;
  wr[IPR] = pf + pr;
#pragma unroll
  for (int v = IVX; v <= IVZ; ++v) {
    const float* q = ws + v * HS + base;
    wl[v] = cp5_hi(q[c - 3], q[c - 2], q[c - 1], q[c], q[c + 1]);
    wr[v] = cp5_lo(q[c - 2], q[c - 1], q[c], q[c + 1], q[c + 2]);
  }
  riemann_flux_pert<float>(solver, wl, wr, rl, rr, pl, pr, pf, gammad, flx);
}

// Flux differences, geometric sources and du of the tile's cells on a
// spherical-polar block (merged face pass): the per-face areas and cell
// volumes of the coordinate, then SphericalPolarImpl::forward
// (src/coord/spherical_polar.cpp:212-279) with the x1 face pressure, as the
// plain path's coord_forward_kernel does per cell.  fs holds the fluxes of all
// directions ([5][faces]) and, after them, the x1 face pressures.
// du of the cells of a gnomonic panel tile: flux differences with the
// panel's face areas and volume, the metric sources of
// GnomonicEquiangleImpl::forward (src/coord/gnomonic_equiangle.cpp:362-418)
// with the x1 face pressure, and gravity
__device__ __forceinline__ void gno_cell_du(const FusedArgs& a, const float* ws,
                                            const float* fs, int ftot, int fo1, int fo2, int o1,
                                            int o2, int o3, int e1, int e2, int e3, float dt,
                                            float* __restrict__ du) {
  const Geometry& G = a.geo;
  const int H1 = a.tx1 + 6, H2 = a.tx2 + 2 * a.h2, H = H1 * H2 * (a.tx3 + 2 * a.h3);
  const int ncell = a.n1 * a.n2 * a.n3;
  const bool act2 = a.n2 > 1, act3 = a.n3 > 1;
  const float* fp = fs + 5 * ftot;
  const DivMod de1(e1), de2(e2);
  for (int cid = threadIdx.x; cid < e1 * e2 * e3; cid += blockDim.x) {
    int cx, cy;
    int cz = de2.div(de1.div(cid, cx), cy);
    const int i = o1 + cx + a.ng;
    const int j = act2 ? o2 + cy + a.ng : 0;
    const int k = act3 ? o3 + cz + a.ng : 0;
    const GnoCell gc = gno_cell(G, k, j, i);
    const float vol = (float)gc.vol;
    const float A1m = (float)gc.a1m, A1p = (float)gc.a1p;
    const float A2m = (float)gc.a2m, A2p = (float)gc.a2p;
    const float A3m = (float)gc.a3m, A3p = (float)gc.a3p;
    const int lo1 = (cz * e2 + cy) * (e1 + 1) + cx;
    const int lo2 = fo1 + (cz * (e2 + 1) + cy) * e1 + cx;
    const int lo3 = fo2 + (cz * e2 + cy) * e1 + cx;
    float div[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      float acc = A1p * fs[v * ftot + lo1 + 1] - A1m * fs[v * ftot + lo1];
      if (act2) acc += A2p * fs[v * ftot + lo2 + e1] - A2m * fs[v * ftot + lo2];
      if (act3) acc += A3p * fs[v * ftot + lo3 + e1 * e2] - A3m * fs[v * ftot + lo3];
      div[v] = acc / vol;
    }
    const int lin = ((cz + a.h3) * H2 + cy + a.h2) * H1 + cx + 3;
    const float rho = ws[IDN * H + lin], v1 = ws[IVX * H + lin], v2 = ws[IVY * H + lin],
                v3 = ws[IVZ * H + lin], pre = ws[IPR * H + lin];
    const float cosine = (float)gc.cos, sine2 = (float)gc.sin * (float)gc.sin;
    const float radius = (float)gc.r;
    const float v_2 = v2 + v3 * cosine, v_3 = v3 + v2 * cosine;
    const float fpm = fp[lo1], fpp = fp[lo1 + 1];
    float src1 = rho * (v2 * v_2 + v3 * v_3) / radius;
    src1 += (A1p * fpp - A1m * fpm) / vol - (fpp - fpm) / (float)gc.dx1;
    div[IVX] -= src1;
    div[IVY] -= (float)gc.xov * (pre + rho * v3 * v3 * sine2) - rho * v1 * v_2 / radius;
    div[IVZ] -= (float)gc.yov * (pre + rho * v2 * v2 * sine2) - rho * v1 * v_3 / radius;
    float d[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) d[v] = -dt * div[v];
    if (a.grav1 != 0.f) {
      d[IVX] += dt * rho * a.grav1;
      d[IPR] += dt * rho * v1 * a.grav1;
    }
    if (a.grav2 != 0.f) {
      d[IVY] += dt * rho * a.grav2;
      d[IPR] += dt * rho * v2 * a.grav2;
    }
    if (a.grav3 != 0.f) {
      d[IVZ] += dt * rho * a.grav3;
      d[IPR] += dt * rho * v3 * a.grav3;
    }
    const int q = (k * a.n2 + j) * a.n1 + i;
#pragma unroll
    for (int v = 0; v < 5; ++v) du[v * ncell + q] = d[v];
  }
}

__device__ __forceinline__ void sph_cell_du(const FusedArgs& a, const float* ws,
                                            const float* fs, int ftot, int fo1, int fo2, int o1,
                                            int o2, int o3, int e1, int e2, int e3, float dt,
                                            float* __restrict__ du) {
  const Geometry& G = a.geo;
  const int H1 = a.tx1 + 6, H2 = a.tx2 + 2 * a.h2, H = H1 * H2 * (a.tx3 + 2 * a.h3);
  const int ncell = a.n1 * a.n2 * a.n3;
  const bool act2 = a.n2 > 1, act3 = a.n3 > 1;
  const float* fp = fs + 5 * ftot;
  const DivMod de1(e1), de2(e2);
  for (int cid = threadIdx.x; cid < e1 * e2 * e3; cid += blockDim.x) {
    int cx, cy;
    int cz = de2.div(de1.div(cid, cx), cy);
    const int i = o1 + cx + a.ng;
    const int j = act2 ? o2 + cy + a.ng : 0;
    const int k = act3 ? o3 + cz + a.ng : 0;
    const SphCell gc = sph_cell(G, k, j, i);
    const float vol = (float)gc.vol;
    const float A1m = (float)gc.a1m, A1p = (float)gc.a1p;
    const int lo1 = (cz * e2 + cy) * (e1 + 1) + cx;
    const int lo2 = fo1 + (cz * (e2 + 1) + cy) * e1 + cx;
    const int lo3 = fo2 + (cz * e2 + cy) * e1 + cx;
    const float A2m = (float)gc.a2m, A2p = (float)gc.a2p, A3 = (float)gc.a3;
    float div[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      float acc = A1p * fs[v * ftot + lo1 + 1] - A1m * fs[v * ftot + lo1];
      if (act2) acc += A2p * fs[v * ftot + lo2 + e1] - A2m * fs[v * ftot + lo2];
      if (act3) acc += A3 * fs[v * ftot + lo3 + e1 * e2] - A3 * fs[v * ftot + lo3];
      div[v] = acc / vol;
    }
    const int lin = ((cz + a.h3) * H2 + cy + a.h2) * H1 + cx + 3;
    const float rho = ws[IDN * H + lin], v1 = ws[IVX * H + lin], v2 = ws[IVY * H + lin],
                v3 = ws[IVZ * H + lin], pre = ws[IPR * H + lin];
    const float s1i = (float)gc.src1_i, s1j = (float)gc.src1_j;
    const float s2i = (float)gc.src2_i;
    const float fpm = fp[lo1], fpp = fp[lo1 + 1];
    float src1 = s1i * rho * (v2 * v2 + v3 * v3);
    src1 += (A1p * fpp - A1m * fpm) / vol - (fpp - fpm) / (float)gc.dx1;
    div[IVX] -= src1;
    div[IVY] -= s1i * s1j * (rho * (v3 * v3) + pre);
    div[IVY] += s2i * (A1m * fs[IVY * ftot + lo1] + A1p * fs[IVY * ftot + lo1 + 1]);
    div[IVZ] += s2i * (A1m * fs[IVZ * ftot + lo1] + A1p * fs[IVZ * ftot + lo1 + 1]);
    if (act2)
      div[IVZ] += s1i * (float)gc.src2_j *
                  (A2m * fs[IVZ * ftot + lo2] + A2p * fs[IVZ * ftot + lo2 + e1]);
    else
      div[IVZ] += s1i * s1j * (rho * v3 * v2);
    float d[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) d[v] = -dt * div[v];
    if (a.grav1 != 0.f) {
      d[IVX] += dt * rho * a.grav1;
      d[IPR] += dt * rho * v1 * a.grav1;
    }
    if (a.grav2 != 0.f) {
      d[IVY] += dt * rho * a.grav2;
      d[IPR] += dt * rho * v2 * a.grav2;
    }
    if (a.grav3 != 0.f) {
      d[IVZ] += dt * rho * a.grav3;
      d[IPR] += dt * rho * v3 * a.grav3;
    }
    const int q = (k * a.n2 + j) * a.n1 + i;
#pragma unroll
    for (int v = 0; v < 5; ++v) du[v * ncell + q] = d[v];
  }
}

// Face options fixed at compile time (-1: read from the arguments at run
// time): the density interpolant and the interpolant of velocities and
// pressure for x1 (kTy0, kTy20) and for x2 / x3 (kTy1, kTy21), and the
// Riemann solver per direction.  A fixed instance has the scale option off.
// The instances of the configurations in use compile one interpolant pair and
// one solver per direction, without the run-time dispatch (the marching K1
// and rhs_tile).
template <int kTy0, int kTy20, int kTy1, int kTy21, int kR1, int kR2, int kR3>
struct MarchFaces {
  static constexpr int ty0 = kTy0, ty20 = kTy20, ty1 = kTy1, ty21 = kTy21;
  static constexpr int r1 = kR1, r2 = kR2, r3 = kR3;
};
using MarchGeneric = MarchFaces<-1, -1, -1, -1, -1, -1, -1>;
// the compiled instances: WENO5 density with CP5 velocities and pressure on
// both axis groups, with LMARS in x1 and Roe in x2 / x3, or one solver in
// every direction
using FacesLmRoRo = MarchFaces<SNAPX_INTERP_WENO5, SNAPX_INTERP_CP5, SNAPX_INTERP_WENO5,
                               SNAPX_INTERP_CP5, SNAPX_RIEMANN_LMARS, SNAPX_RIEMANN_ROE,
                               SNAPX_RIEMANN_ROE>;
using FacesLm = MarchFaces<SNAPX_INTERP_WENO5, SNAPX_INTERP_CP5, SNAPX_INTERP_WENO5,
                           SNAPX_INTERP_CP5, SNAPX_RIEMANN_LMARS, SNAPX_RIEMANN_LMARS,
                           SNAPX_RIEMANN_LMARS>;
using FacesRo = MarchFaces<SNAPX_INTERP_WENO5, SNAPX_INTERP_CP5, SNAPX_INTERP_WENO5,
                           SNAPX_INTERP_CP5, SNAPX_RIEMANN_ROE, SNAPX_RIEMANN_ROE,
                           SNAPX_RIEMANN_ROE>;
using FacesHl = MarchFaces<SNAPX_INTERP_WENO5, SNAPX_INTERP_CP5, SNAPX_INTERP_WENO5,
                           SNAPX_INTERP_CP5, SNAPX_RIEMANN_HLLC, SNAPX_RIEMANN_HLLC,
                           SNAPX_RIEMANN_HLLC>;
enum { kFacesGeneric = 0, kFacesLmRoRo, kFacesLm, kFacesRo, kFacesHl };

// The compiled face-option instance that matches a block (kFacesGeneric:
// none; the solvers of inactive directions do not matter).
static int face_policy(const snapx_config* c) {
  constexpr int W5 = SNAPX_INTERP_WENO5;
  constexpr int RO = SNAPX_RIEMANN_ROE, LM = SNAPX_RIEMANN_LMARS, HL = SNAPX_RIEMANN_HLLC;
  const bool weno5_cp5 = c->recon_type[0] == W5 && c->recon_type[1] == W5 && !c->recon_shock[0] &&
                         !c->recon_shock[1] && !c->recon_scale[0] && !c->recon_scale[1];
  if (!weno5_cp5) return kFacesGeneric;
  const int r1 = c->riemann[0];
  const int r2 = c->nx2 > 1 ? c->riemann[1] : r1;
  const int r3 = c->nx3 > 1 ? c->riemann[2] : r2;
  if (r1 == LM && r2 == RO && r3 == RO) return kFacesLmRoRo;
  if (r1 == LM && r2 == LM && r3 == LM) return kFacesLm;
  if (r1 == RO && r2 == RO && r3 == RO) return kFacesRo;
  if (r1 == HL && r2 == HL && r3 == HL) return kFacesHl;
  return kFacesGeneric;
}


// The face options of direction d (row r = 0 for x1, 1 for x2 / x3): the
// constants of P where it fixes them, else the argument fields passed in
// (by value: a reference to the by-value argument struct would place it in
// scratch memory).
template <class P>
__device__ __forceinline__ int face_type(int r, int t0, int t1) {
  if (P::ty0 >= 0 && P::ty1 >= 0) return r ? P::ty1 : P::ty0;
  return r ? t1 : t0;
}
template <class P>
__device__ __forceinline__ int face_scale(int r, int s0, int s1) {
  if (P::ty0 >= 0 && P::ty1 >= 0) return 0;
  return r ? s1 : s0;
}
template <class P>
__device__ __forceinline__ int face_type2(int r, int type, int shock0, int shock1) {
  if (P::ty20 >= 0 && P::ty21 >= 0) return r ? P::ty21 : P::ty20;
  const int shock = r ? shock1 : shock0;
  return shock ? type : interp2_type(type);
}
template <class P>
__device__ __forceinline__ int face_solver(int d, int r0, int r1, int r2) {
  if (P::r1 >= 0 && P::r2 >= 0 && P::r3 >= 0) return d == 0 ? P::r1 : (d == 1 ? P::r2 : P::r3);
  return d == 0 ? r0 : (d == 1 ? r1 : r2);
}

// K1 body for one tile; every thread of the workgroup calls it with the same
// tile (it synchronises the workgroup internally).  The arguments come by
// value: a reference to the kernel argument makes the compiler keep a copy
// of the whole struct in scratch memory.
// kMerged: 1 / 0 when the launch knows that every tile takes the merged
// face pass / the per-direction passes (the other path is not compiled in),
// -1 to follow a.merged at run time (batches of mixed members).
// kBg (merged face pass, Cartesian): perturbation path about the
// hydrostatic background of a.bgt / a.bgd, as in fused_rhs_march_kernel.
// kSph: 0 Cartesian, 1 spherical-polar, 2 gnomonic panel (curvilinear
// tiles take the merged face pass).
// The arguments reach rhs_tile by value on Cartesian tiles and by reference
// on curvilinear ones: with these forms the compiler keeps every instance's
// copy of FusedArgs out of scratch memory (the per-direction field selects
// of the Cartesian two-pass tile and the geometry reads of the curvilinear
// ones defeat it otherwise).
template <int kSph>
using RhsArgs = std::conditional_t<kSph != 0, const FusedArgs&, const FusedArgs>;

template <int kMerged, int kSph = 0, bool kBg = false, class P = MarchGeneric>
__device__ __forceinline__ void rhs_tile(const float* __restrict__ u, float* __restrict__ du,
                                         RhsArgs<kSph> a, int b, float* smem, float dt) {
  static_assert(!kBg || (kMerged == 1 && !kSph), "background path: merged Cartesian tiles");
  const bool merged = kMerged < 0 ? (a.merged != 0) : kMerged == 1;
  const int H1 = a.tx1 + 6, H2 = a.tx2 + 2 * a.h2, H3 = a.tx3 + 2 * a.h3;
  const int H = H1 * H2 * H3;
  constexpr int NP = kBg ? 6 : 5;  // LDS planes: primitives (+ total density)
  float* ws = smem;           // primitives [5][H3][H2][H1] (+ total density)
  float* fs = smem + NP * H;  // face fluxes of one direction [5][faces]
  BgView BV;
  if (kBg) BV = bg_global_view(a.bgt, BgLayout(a.n1));
  const int ncell = a.n1 * a.n2 * a.n3;

  int t1 = b % a.nt1;
  int t2 = (b / a.nt1) % a.nt2;
  int t3 = b / (a.nt1 * a.nt2);
  int o1 = t1 * a.tx1, o2 = t2 * a.tx2, o3 = t3 * a.tx3;
  int e1 = min(a.tx1, a.nx1 - o1);
  int e2 = a.n2 > 1 ? min(a.tx2, a.nx2 - o2) : 1;
  int e3 = a.n3 > 1 ? min(a.tx3, a.nx3 - o3) : 1;
  const float gammad = a.gammad;

  // ---- 1. halo'd tile: conserved -> primitive into LDS, and the isentropic
  //         refill of the physical x1 ghosts (hydro.cpp:231-269): for small
  //         tiles (merged face pass) inside the load pass, by the ghost cell's
  //         own thread from the boundary cell of its column; for large tiles
  //         in a pass of its own, one thread per column end (fewer registers
  //         in the load loop, which each thread runs many times) ----
  const DivMod dH1(H1), dH2(H2);
  const bool refill = a.grav1 != 0.f;
  const float gdz = -a.grav1 * a.dx1;
  for (int l = threadIdx.x; l < H; l += blockDim.x) {
    int x, y;
    int z = dH2.div(dH1.div(l, x), y);
    const int gi = o1 + x - 3, gj = o2 + y - a.h2, gk = o3 + z - a.h3;
    bool fl[3];
    int q = mapped_cell(a, gi, gj, gk, fl);
    float U[5], W[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) U[v] = u[v * ncell + q];
    if (fl[0]) U[IVX] = -U[IVX];
    if (fl[1]) U[IVY] = -U[IVY];
    if (fl[2]) U[IVZ] = -U[IVZ];
    if (kBg) {
      float rtot;
      bg_load_prim(a, BV, U, gi + a.ng, W, rtot);
      if (refill && (gi < 0 || gi >= a.nx1) && (unsigned)(y - a.h2) < (unsigned)e2 &&
          (unsigned)(z - a.h3) < (unsigned)e3) {
        bool flb[3];
        const int qb = mapped_cell(a, gi >= a.nx1 ? a.nx1 - 1 : 0, gj, gk, flb);
        bg_refill_prim(a, BV, u, ncell, qb, gi, W, rtot);
      }
      ws[5 * H + l] = rtot;
#pragma unroll
      for (int v = 0; v < 5; ++v) ws[v * H + l] = W[v];
      continue;
    }
    // cos(theta) of a gnomonic panel cell (the cell's own position, ghosts
    // included, as the reference's cosine_cell_kj)
    constexpr bool gno = kSph == 2;
    const float cth = gno ? (float)gno_cos_cell(a.geo, a.n3 > 1 ? gk + a.ng : 0,
                                                a.n2 > 1 ? gj + a.ng : 0)
                          : 0.f;
    if (gno)
      cons_to_prim_gno(U, gammad, cth, W);
    else
      cons_to_prim(U, gammad, W);
    // (only the tile's own columns: the x1 ghosts of halo columns are never read)
    if (merged && refill && (gi < 0 || gi >= a.nx1) && (unsigned)(y - a.h2) < (unsigned)e2 &&
        (unsigned)(z - a.h3) < (unsigned)e3) {
      const bool upper = gi >= a.nx1;
      bool flb[3];
      int qb = mapped_cell(a, upper ? a.nx1 - 1 : 0, gj, gk, flb);
      float Ub[5], Wb[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) Ub[v] = u[v * ncell + qb];
      // rho and p do not depend on the flips
      if (gno)
        cons_to_prim_gno(Ub, gammad, cth, Wb);
      else
        cons_to_prim(Ub, gammad, Wb);
      if (kSph) {
        const IsentropicRefill R(Wb[IDN], Wb[IPR], gammad, -a.grav1, upper ? 1 : -1);
        R.layer_at(upper ? a.zout[gi - a.nx1] : a.zin[-gi - 1], W[IDN], W[IPR]);
      } else {
        const IsentropicRefill R(Wb[IDN], Wb[IPR], gammad, gdz, upper ? 1 : -1);
        R.layer(upper ? gi - a.nx1 + 1 : -gi, W[IDN], W[IPR]);
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) ws[v * H + l] = W[v];
  }
  __syncthreads();
  if (!merged && refill) {
    // the tile's LDS holds physical x1 ghosts below when it starts at the
    // bottom, and above whenever its 3-cell halo passes the last cell: the
    // top tile, and also the one below a last tile of fewer than 3 cells
    const bool bot = o1 == 0, top = o1 + e1 + 3 > a.nx1;
    const int ncolt = e2 * e3;
    for (int c = threadIdx.x; c < 2 * ncolt; c += blockDim.x) {
      const bool upper = c >= ncolt;
      if (upper ? !top : !bot) continue;
      const int cc = upper ? c - ncolt : c;
      const int y = a.h2 + cc % e2, z = a.h3 + cc / e2;
      float* rho = ws + IDN * H + (z * H2 + y) * H1;
      float* pre = ws + IPR * H + (z * H2 + y) * H1;
      // LDS index of the boundary cell (the last interior cell above)
      const int is = upper ? 3 + (a.nx1 - 1 - o1) : 3;
      const IsentropicRefill R(rho[is], pre[is], gammad, gdz, upper ? 1 : -1);
#pragma unroll
      for (int n = 1; n <= 3; ++n) {
        const int x = upper ? is + n : is - n;
        if (x < H1) R.layer(n, rho[x], pre[x]);  // layers past this tile's halo: another tile's
      }
    }
    __syncthreads();
  }

  // ---- 2. interface states + Riemann fluxes of every direction at once ----
  // The faces of all active directions form one index space (direction d
  // occupies [fo[d], fo[d] + nf[d])), so that a single pass and a single
  // barrier cover them; the fluxes go to fs [5][total faces].
  float acc[kMaxCellsPerThread][5];
#pragma unroll
  for (int r = 0; r < kMaxCellsPerThread; ++r)
#pragma unroll
    for (int v = 0; v < 5; ++v) acc[r][v] = 0.f;
  const int ncells_t = e1 * e2 * e3;
  const DivMod de1(e1), de2(e2);
  const bool act2 = a.n2 > 1, act3 = a.n3 > 1;
  const int nf0 = (e1 + 1) * e2 * e3;
  const int nf1 = act2 ? e1 * (e2 + 1) * e3 : 0;
  const int nf2 = act3 ? e1 * e2 * (e3 + 1) : 0;
  const int fo1 = nf0, fo2 = nf0 + nf1, ftot = fo2 + nf2;
  const DivMod dq1(e1 + 1), dq2(e2 + 1);

  if (merged) {
    // few faces per thread: every direction in one pass, one barrier
    for (int g = threadIdx.x; g < ftot; g += blockDim.x) {
      const int d = g < fo1 ? 0 : (g < fo2 ? 1 : 2);
      const int f = g - (d == 0 ? 0 : (d == 1 ? fo1 : fo2));
      // face space of d: extent along d grows by one
      const int f1 = e1 + (d == 0), f2 = e2 + (d == 1);
      const int sd = d == 0 ? 1 : (d == 1 ? H1 : H1 * H2);  // LDS stride along d
      const int r = d == 0 ? 0 : 1;
      // selects, not a runtime index into the argument struct (which would
      // place it in scratch memory); the options P fixes are constants
      const int type = face_type<P>(r, a.recon_type[0], a.recon_type[1]);
      const int scale = face_scale<P>(r, a.recon_scale[0], a.recon_scale[1]);
      const int type2 = face_type2<P>(r, type, a.recon_shock[0], a.recon_shock[1]);
      const int solver = face_solver<P>(d, a.riemann[0], a.riemann[1], a.riemann[2]);
      int fx, fy, fz;
      if (d == 0) {
        fz = de2.div(dq1.div(f, fx), fy);
      } else if (d == 1) {
        fz = dq2.div(de1.div(f, fx), fy);
      } else {
        fz = de2.div(de1.div(f, fx), fy);
      }
      // LDS coordinates of the cell above the face along d
      int x = fx + 3, y = fy + a.h2, z = fz + a.h3;
      int lin = (z * H2 + y) * H1 + x;
      int c = d == 0 ? x : (d == 1 ? y : z);
      int base = lin - c * sd;  // pencil start along d
      // face frame: slot v holds variable vmap[v] (normal velocity in slot
      // IVX); register arrays are indexed by constants only, the frame is
      // entered and left through LDS addresses
      const int vmap[5] = {IDN, IVX + d, IVX + (d + 1) % 3, IVX + (d + 2) % 3, IPR};
      if (kBg) {
        float flx[5];
        const int lv = o1 + fx + a.ng;  // x1 level of the face / of the x2, x3 face's cells
        if (d == 0) {
          const int gi = o1 + fx;
          const bool rev = a.grav1 != 0.f;
          bg_x1_face_flux(ws, H, base, c, lv, BV, rev && gi == 0, rev && gi == a.nx1, solver,
                          gammad, flx);
        } else {
          float wl[5], wr[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            int tv = (v == IDN) ? type : type2;
            const float* p = ws + vmap[v] * H + base;
            wr[v] = interp_face<float>(tv, 0, scale, p, sd, c);
            wl[v] = interp_face<float>(tv, 1, scale, p, sd, c - 1);
          }
          bg_horizontal_flux(BV, lv, solver, wl, wr, gammad, flx);
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) fs[vmap[v] * ftot + g] = flx[v];
        continue;
      }
      float wl[5], wr[5];
  #pragma unroll
      for (int v = 0; v < 5; ++v) {
        int tv = (v == IDN) ? type : type2;
        const float* p = ws + vmap[v] * H + base;
        wr[v] = interp_face<float>(tv, 0, scale, p, sd, c);
        wl[v] = interp_face<float>(tv, 1, scale, p, sd, c - 1);
      }
      if (d == 0 && a.grav1 != 0.f) {
        // _revise_x1inner_lr / _revise_x1outer_lr (hydro.cpp:217-229)
        int gi = o1 + fx;
        if (gi == 0) {
          wl[IDN] = wr[IDN];
          wl[IPR] = wr[IPR];
        }
        if (gi == a.nx1) {
          wr[IDN] = wl[IDN];
          wr[IPR] = wl[IPR];
        }
      }
      float flx[5];
      if (kSph) {
        // gnomonic panels: LMARS and HLLC in the face's orthonormal frame
        // (src/riemann/lmars.cpp:54-99, src/riemann/hllc.cpp:56-103)
        const bool frame = kSph == 2 && solver != SNAPX_RIEMANN_ROE;
        float fc = 0.f, fsn = 1.f;
        if (frame) {
          const int kk = act3 ? o3 + fz + a.ng : 0, jj = act2 ? o2 + fy + a.ng : 0;
          const GnoAngle ang = gno_face_angle(a.geo, d == 0 ? 3 : (d == 1 ? 2 : 1), kk, jj);
          fc = (float)ang.c, fsn = (float)ang.s;
          gno_slots_to_local(wl, d, fc, fsn);
          gno_slots_to_local(wr, d, fc, fsn);
        }
        // the x1 face pressure feeds the radial momentum source
        float pf;
        riemann_flux<float>(solver, wl, wr, gammad, 3, flx, &pf);
        if (frame) gno_slots_to_global(flx, d, fc, fsn);
        if (d == 0) fs[5 * ftot + g] = pf;
      } else {
        riemann_flux<float>(solver, wl, wr, gammad, 3, flx, (float*)nullptr);
      }
  #pragma unroll
      for (int v = 0; v < 5; ++v) fs[vmap[v] * ftot + g] = flx[v];
      (void)f1;
      (void)f2;
    }
    __syncthreads();
    if (kSph) {
      if (kSph == 2)
        gno_cell_du(a, ws, fs, ftot, fo1, fo2, o1, o2, o3, e1, e2, e3, dt, du);
      else
        sph_cell_du(a, ws, fs, ftot, fo1, fo2, o1, o2, o3, e1, e2, e3, dt, du);
      return;
    }
    // flux differences, directions in order (x1, x2, x3)
  #pragma unroll
    for (int rr = 0; rr < kMaxCellsPerThread; ++rr) {
      int cid = threadIdx.x + rr * kRhsThreads;
      if (cid < ncells_t) {
        int cx, cy;
        int cz = de2.div(de1.div(cid, cx), cy);
        {
          int lo = (cz * e2 + cy) * (e1 + 1) + cx;
          float area = a.dx3 * a.dx2;
  #pragma unroll
          for (int v = 0; v < 5; ++v)
            acc[rr][v] += area * (fs[v * ftot + lo + 1] - fs[v * ftot + lo]);
        }
        if (act2) {
          int lo = fo1 + (cz * (e2 + 1) + cy) * e1 + cx;
          float area = a.dx3 * a.dx1;
  #pragma unroll
          for (int v = 0; v < 5; ++v)
            acc[rr][v] += area * (fs[v * ftot + lo + e1] - fs[v * ftot + lo]);
        }
        if (act3) {
          int lo = fo2 + (cz * e2 + cy) * e1 + cx;
          float area = a.dx2 * a.dx1;
  #pragma unroll
          for (int v = 0; v < 5; ++v)
            acc[rr][v] += area * (fs[v * ftot + lo + e1 * e2] - fs[v * ftot + lo]);
        }
      }
    }
  } else {
    // many faces per thread: one direction at a time, fs holds one direction
    for (int d = 0; d < 3; ++d) {
      int nd = d == 0 ? a.nx1 : (d == 1 ? a.nx2 : a.nx3);
      int ncd = d == 0 ? a.n1 : (d == 1 ? a.n2 : a.n3);
      if (ncd == 1) continue;
      // face space: extent along d grows by one
      int f1 = e1 + (d == 0), f2 = e2 + (d == 1), f3 = e3 + (d == 2);
      int nf = f1 * f2 * f3;
      int sd = d == 0 ? 1 : (d == 1 ? H1 : H1 * H2);  // LDS stride along d
      int r = d == 0 ? 0 : 1;
      // selects, not a runtime index into the argument struct (which would
      // place it in scratch memory); the options P fixes are constants
      const int type = face_type<P>(r, a.recon_type[0], a.recon_type[1]);
      const int scale = face_scale<P>(r, a.recon_scale[0], a.recon_scale[1]);
      const int type2 = face_type2<P>(r, type, a.recon_shock[0], a.recon_shock[1]);
      const int solver = face_solver<P>(d, a.riemann[0], a.riemann[1], a.riemann[2]);
      const DivMod df1(f1), df2(f2);
      for (int f = threadIdx.x; f < nf; f += blockDim.x) {
        int fx, fy;
        int fz = df2.div(df1.div(f, fx), fy);
        // LDS coordinates of the cell above the face along d
        int x = fx + 3, y = fy + a.h2, z = fz + a.h3;
        int lin = (z * H2 + y) * H1 + x;
        int c = d == 0 ? x : (d == 1 ? y : z);
        int base = lin - c * sd;  // pencil start along d
        // face frame: slot v holds variable vmap[v] (normal velocity in slot
        // IVX); register arrays are indexed by constants only, the frame is
        // entered and left through LDS addresses
        const int vmap[5] = {IDN, IVX + d, IVX + (d + 1) % 3, IVX + (d + 2) % 3, IPR};
        float wl[5], wr[5];
  #pragma unroll
        for (int v = 0; v < 5; ++v) {
          int tv = (v == IDN) ? type : type2;
          const float* p = ws + vmap[v] * H + base;
          wr[v] = interp_face<float>(tv, 0, scale, p, sd, c);
          wl[v] = interp_face<float>(tv, 1, scale, p, sd, c - 1);
        }
        if (d == 0 && a.grav1 != 0.f) {
          // _revise_x1inner_lr / _revise_x1outer_lr (hydro.cpp:217-229)
          int gi = o1 + fx;
          if (gi == 0) {
            wl[IDN] = wr[IDN];
            wl[IPR] = wr[IPR];
          }
          if (gi == nd) {
            wr[IDN] = wl[IDN];
            wr[IPR] = wl[IPR];
          }
        }
        float flx[5];
        riemann_flux<float>(solver, wl, wr, gammad, 3, flx, (float*)nullptr);
  #pragma unroll
        for (int v = 0; v < 5; ++v) fs[vmap[v] * nf + f] = flx[v];
      }
      __syncthreads();
      float area = d == 0 ? a.dx3 * a.dx2 : (d == 1 ? a.dx3 * a.dx1 : a.dx2 * a.dx1);
  #pragma unroll
      for (int rr = 0; rr < kMaxCellsPerThread; ++rr) {
        int cid = threadIdx.x + rr * kRhsThreads;
        if (cid < ncells_t) {
          int cx, cy;
          int cz = de2.div(de1.div(cid, cx), cy);
          int lo = (cz * f2 + cy) * f1 + cx;
          int hi = lo + (d == 0 ? 1 : (d == 1 ? f1 : f1 * f2));
  #pragma unroll
          for (int v = 0; v < 5; ++v) acc[rr][v] += area * (fs[v * nf + hi] - fs[v * nf + lo]);
        }
      }
      __syncthreads();
    }
  }

  // ---- 3. du = -dt D + gravity ----
  float vol = a.dx3 * (a.dx2 * a.dx1);
#pragma unroll
  for (int rr = 0; rr < kMaxCellsPerThread; ++rr) {
    int cid = threadIdx.x + rr * kRhsThreads;
    if (cid >= ncells_t) continue;
    int cx, cy;
    int cz = de2.div(de1.div(cid, cx), cy);
    int lin = ((cz + a.h3) * H2 + cy + a.h2) * H1 + cx + 3;
    float d[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) d[v] = -dt * (acc[rr][v] / vol);
    float rho = ws[IDN * H + lin];
    if (kBg) {
      // the background's flux divergence and gravity are the residual
      const int lv = o1 + cx + a.ng;
      d[IVX] += dt * (rho * a.grav1 + BV.lv(BG_RES, lv));
      rho = ws[5 * H + lin];
      d[IPR] += dt * rho * ws[IVX * H + lin] * a.grav1;
    } else if (a.grav1 != 0.f) {
      d[IVX] += dt * rho * a.grav1;
      d[IPR] += dt * rho * ws[IVX * H + lin] * a.grav1;
    }
    if (a.grav2 != 0.f) {
      d[IVY] += dt * rho * a.grav2;
      d[IPR] += dt * rho * ws[IVY * H + lin] * a.grav2;
    }
    if (a.grav3 != 0.f) {
      d[IVZ] += dt * rho * a.grav3;
      d[IPR] += dt * rho * ws[IVZ * H + lin] * a.grav3;
    }
    int qi = o1 + cx + a.ng;
    int qj = a.n2 > 1 ? o2 + cy + a.ng : 0;
    int qk = a.n3 > 1 ? o3 + cz + a.ng : 0;
    int q = (qk * a.n2 + qj) * a.n1 + qi;
#pragma unroll
    for (int v = 0; v < 5; ++v) du[v * ncell + q] = d[v];
  }
}

template <int kMerged, class P = MarchGeneric>
__global__ void __launch_bounds__(kRhsThreads)
    fused_rhs_kernel(const float* __restrict__ u, float* __restrict__ du, FusedArgs a) {
  extern __shared__ float smem[];
  rhs_tile<kMerged, 0, false, P>(u, du, a, blockIdx.x, smem, step_dt(a));
}

// K1 of the perturbation path (merged face pass, Cartesian)
template <class P = MarchGeneric>
__global__ void __launch_bounds__(kRhsThreads)
    fused_rhs_bg_kernel(const float* __restrict__ u, float* __restrict__ du, FusedArgs a) {
  extern __shared__ float smem[];
  rhs_tile<1, 0, true, P>(u, du, a, blockIdx.x, smem, step_dt(a));
}

// K1 of spherical-polar blocks (merged face pass, geometric sources)
__global__ void __launch_bounds__(kRhsThreads)
    fused_rhs_sph_kernel(const float* __restrict__ u, float* __restrict__ du, FusedArgs a) {
  extern __shared__ float smem[];
  rhs_tile<1, 1>(u, du, a, blockIdx.x, smem, step_dt(a));
}

// K1 of gnomonic panels (merged face pass, face frames, metric sources)
__global__ void __launch_bounds__(kRhsThreads)
    fused_rhs_gno_kernel(const float* __restrict__ u, float* __restrict__ du, FusedArgs a) {
  extern __shared__ float smem[];
  rhs_tile<1, 2>(u, du, a, blockIdx.x, smem, step_dt(a));
}

// ------------------------------------------------ K1 for 3-D blocks (2.5-D) ----
//
// For 3-D blocks the x1-x2 tile (tx1 x tx2 <= 256 cells, one per thread)
// marches along x3 through a slab of tx3 slices.  Per slice the halo'd x1-x2
// plane of primitives goes to LDS once (one buffer), the x1 and x2 faces of
// the slice are computed from LDS, and every thread keeps its own column's
// last six slices in registers for the x3 faces; the three halo slices at
// each end of the slab load only the thread's own cell.  Halo loads per
// interior cell fall from 8.6 (16 x 4 x 4 tiles of rhs_tile) to
// ((tx1 + 6)(tx2 + 6) tx3 + 6 tx1 tx2) / (tx1 tx2 tx3), 2.45 for 32 x 8 x 16.
// The operations (interface states, fluxes, the order of the flux
// differences, the ghost refill) are those of rhs_tile; the two agree to FP32
// rounding (the compiler contracts a few multiply-adds differently).
// (four wavefronts per SIMD: 128 registers, a few spilled)
//

// Interface states of one face in the face frame of direction kD (normal
// velocity in slot IVX) from the pencil of each variable, p_v = ws + vmap[v]
// * HS + base with stride sd; c is the pencil index of the cell above.
template <int kD, int kTy, int kTy2>
__device__ __forceinline__ void face_states(const float* ws, int HS, int base, int sd, int c,
                                            int type, int type2, int scale, float* wl,
                                            float* wr) {
  constexpr int vmap[5] = {IDN, IVX + kD, IVX + (kD + 1) % 3, IVX + (kD + 2) % 3, IPR};
  const int t = kTy >= 0 ? kTy : type, t2 = kTy2 >= 0 ? kTy2 : type2;
  const int sc = kTy >= 0 ? 0 : scale;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const float* p = ws + vmap[v] * HS + base;
    wr[v] = interp_face<float>(v == IDN ? t : t2, 0, sc, p, sd, c);
    wl[v] = interp_face<float>(v == IDN ? t : t2, 1, sc, p, sd, c - 1);
  }
}

// kBg: the state is the perturbation about the hydrostatic background of
// a.bgt / a.bgd (bg_impl.hpp); the LDS planes hold (rho', v1, v2, v3, p') and
// the total density, face states are references plus perturbations and the
// solvers run in perturbation form (riemann_pert.hpp).
template <class P, bool kBg = false>
__global__ void __launch_bounds__(kRhsThreads, 4)
    fused_rhs_march_kernel(const float* __restrict__ u, float* __restrict__ du, FusedArgs a) {
  extern __shared__ float smem[];
  constexpr int NP = kBg ? 6 : 5;  // LDS planes per slice
  const int H1 = a.tx1 + 6, H2 = a.tx2 + 6, HS = H1 * H2;
  const int ncell = a.n1 * a.n2 * a.n3;
  // workgroups go to the eight XCDs round robin (workgroup i to XCD i mod
  // 8); each XCD takes a contiguous run of tiles, so that the halos shared
  // by neighbouring tiles meet in the same L2
  const int nb = gridDim.x, xq = nb >> 3, xr = nb & 7, xcd = blockIdx.x & 7;
  const int b = xcd * xq + min(xcd, xr) + (blockIdx.x >> 3);
  const int t1 = b % a.nt1, t2 = (b / a.nt1) % a.nt2, t3 = b / (a.nt1 * a.nt2);
  const int o1 = t1 * a.tx1, o2 = t2 * a.tx2, o3 = t3 * a.tx3;
  const int e1 = min(a.tx1, a.nx1 - o1), e2 = min(a.tx2, a.nx2 - o2), ez = min(a.tx3, a.nx3 - o3);
  const float gammad = a.gammad;
  const bool refill = a.grav1 != 0.f;
  const float gdz = -a.grav1 * a.dx1;
  // the thread's own cell of the x1-x2 tile
  const DivMod de1(e1), dq1(e1 + 1), dH1(H1);
  int cx;
  const int cy = de1.div(threadIdx.x, cx);
  const bool own = (int)threadIdx.x < e1 * e2;
  const int lin_own = own ? (cy + 3) * H1 + cx + 3 : 0;
  // reconstruction options: x1 row 0, x2 / x3 row 1 (selects, not runtime
  // indices into the argument struct)
  const int type0 = a.recon_type[0], scale0 = a.recon_scale[0];
  const int type1 = a.recon_type[1], scale1 = a.recon_scale[1];
  const int ty20 = a.recon_shock[0] ? type0 : interp2_type(type0);
  const int ty21 = a.recon_shock[1] ? type1 : interp2_type(type1);
  const int sol1 = P::r1 >= 0 ? P::r1 : a.riemann[0];
  const int sol2 = P::r2 >= 0 ? P::r2 : a.riemann[1];
  const int sol3 = P::r3 >= 0 ? P::r3 : a.riemann[2];
  const int nf0 = (e1 + 1) * e2, ftot = nf0 + e1 * (e2 + 1);
  float* fs = smem + NP * HS;  // x1 then x2 face fluxes of one slice [5][ftot]
  // perturbation path: the tables of the tile's H1 levels and tx1 + 1 x1
  // faces, copied to LDS after the face fluxes
  BgView BV;
  if (kBg) {
    BV = bg_stage_lds(a.bgt, BgLayout(a.n1), o1, H1, o1 + a.ng, a.tx1 + 1,
                      fs + 5 * ((a.tx1 + 1) * a.tx2 + a.tx1 * (a.tx2 + 1)));
    __syncthreads();
  }
  const float area1 = a.dx3 * a.dx2, area2 = a.dx3 * a.dx1, area3 = a.dx2 * a.dx1;
  const float vol = a.dx3 * (a.dx2 * a.dx1);
  const float dt = step_dt(a);

  float win[5][6];   // own column, primitives of slices kk - 5 .. kk
  float part[3][5];  // x1 + x2 flux differences of slices kk - 1 .. kk - 3
  float f3lo[5];     // x3 flux at the lower face of the cell finished next
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    f3lo[v] = 0.f;
#pragma unroll
    for (int q = 0; q < 6; ++q) win[v][q] = 1.f;
#pragma unroll
    for (int q = 0; q < 3; ++q) part[q][v] = 0.f;
  }

  // the own cell's offset in its plane (interior in x1 and x2)
  const int own_pq = (o2 + cy + a.ng) * a.n1 + o1 + cx + a.ng;

  for (int kk = o3 - 3; kk <= o3 + ez + 2; ++kk) {
    // one plane buffer: a slice's plane is read only between the barrier
    // after its load and the barrier after its faces, so no thread writes
    // the next plane while another still reads this one
    float* ws = smem;
    bool f3;
    const int s0 = slice_start(a, kk, f3);
    if (kk >= o3 && kk < o3 + ez) {
      // ---- halo'd x1-x2 plane of slice kk: conserved -> primitive ----
      for (int l = threadIdx.x; l < HS; l += blockDim.x) {
        int x;
        const int y = dH1.div(l, x);
        const int gi = o1 + x - 3, gj = o2 + y - 3;
        bool fl[2];
        const int q = s0 + plane_cell(a, gi, gj, fl);
        // a physical x1 ghost of the tile's own rows: its density and
        // pressure come from the refill pass below (velocities from the
        // mirror source)
        const bool rf = refill && (gi < 0 || gi >= a.nx1) && (unsigned)(y - 3) < (unsigned)e2;
        float U[5], W[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) U[v] = u[v * ncell + q];
        if (fl[0]) U[IVX] = -U[IVX];
        if (fl[1]) U[IVY] = -U[IVY];
        if (f3) U[IVZ] = -U[IVZ];
        if (kBg) {
          float rtot;
          bg_load_prim(a, BV, U, gi + a.ng, W, rtot);
          if (!rf) ws[5 * HS + l] = rtot;
        } else {
          cons_to_prim(U, gammad, W);
        }
#pragma unroll
        for (int v = 0; v < 5; ++v)
          if (!rf || (v != IDN && v != IPR)) ws[v * HS + l] = W[v];
      }
      // ---- x1 ghost refill (src/hydro/hydro.cpp:231-269) of the tile's own
      //      rows, from the last thread down: these cells are few, and one
      //      pass keeps the refill's branch out of the other wavefronts ----
      if (refill) {
        const int nlo = max(0, 3 - o1), nhi = max(0, H1 - (a.nx1 - o1 + 3));
        const int per = nlo + nhi;
        for (int k = blockDim.x - 1 - threadIdx.x; k < per * e2; k += blockDim.x) {
          const int row = k / per, c = k - row * per;
          const bool upper = c >= nlo;
          const int x = upper ? a.nx1 - o1 + 3 + (c - nlo) : c, y = row + 3;
          const int l = y * H1 + x, gi = o1 + x - 3, gj = o2 + row;
          bool flb[2];
          const int qb = s0 + plane_cell(a, upper ? a.nx1 - 1 : 0, gj, flb);
          if (kBg) {
            float W[5], rtot;
            bg_refill_prim(a, BV, u, ncell, qb, gi, W, rtot);
            ws[IDN * HS + l] = W[IDN];
            ws[IPR * HS + l] = W[IPR];
            ws[5 * HS + l] = rtot;
          } else {
            float Ub[5], Wb[5], rho, pre;
#pragma unroll
            for (int v = 0; v < 5; ++v) Ub[v] = u[v * ncell + qb];
            cons_to_prim(Ub, gammad, Wb);  // rho and p do not depend on the flips
            const IsentropicRefill R(Wb[IDN], Wb[IPR], gammad, gdz, upper ? 1 : -1);
            R.layer(upper ? gi - a.nx1 + 1 : -gi, rho, pre);
            ws[IDN * HS + l] = rho;
            ws[IPR * HS + l] = pre;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int v = 0; v < 5; ++v) {
#pragma unroll
        for (int q = 0; q < 5; ++q) win[v][q] = win[v][q + 1];
        win[v][5] = own ? ws[v * HS + lin_own] : 1.f;
      }
    } else {
      // ---- x3 halo slice: only the own column's cell is used (x3 faces) ----
      float W[5] = {1.f, 1.f, 1.f, 1.f, 1.f};
      if (own) {
        float U[5];
        const int q = s0 + own_pq;
#pragma unroll
        for (int v = 0; v < 5; ++v) U[v] = u[v * ncell + q];
        if (f3) U[IVZ] = -U[IVZ];
        if (kBg) {
          float rtot;
          bg_load_prim(a, BV, U, o1 + cx + a.ng, W, rtot);
        } else {
          cons_to_prim(U, gammad, W);
        }
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
#pragma unroll
        for (int q = 0; q < 5; ++q) win[v][q] = win[v][q + 1];
        win[v][5] = W[v];
      }
    }
    // ---- x3 face f = kk - 2 (below slice f) from the own column ----
    const int f = kk - 2;
    float f3hi[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (own && f >= o3 && f <= o3 + ez) {
      // face frame of x3: slots (IDN, IVZ, IVX, IVY, IPR)
      const int vmap[5] = {IDN, IVZ, IVX, IVY, IPR};
      const int t = P::ty1 >= 0 ? P::ty1 : type1, t2 = P::ty21 >= 0 ? P::ty21 : ty21;
      const int sc = P::ty1 >= 0 ? 0 : scale1;
      float wl[5], wr[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const int tv = (v == IDN) ? t : t2;
        wr[v] = interp_face<float>(tv, 0, sc, &win[vmap[v]][0], 1, 3);
        wl[v] = interp_face<float>(tv, 1, sc, &win[vmap[v]][0], 1, 2);
      }
      float flx[5];
      if (kBg) {
        // reference: the background of the own column's level
        const int lv = o1 + cx + a.ng;
        bg_horizontal_flux(BV, lv, sol3, wl, wr, gammad, flx);
      } else {
        riemann_flux<float>(sol3, wl, wr, gammad, 3, flx, (float*)nullptr);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) f3hi[vmap[v]] = flx[v];
    }

    // ---- finish cell kk - 3 before the faces of slice kk (its x1 + x2
    //      part is the oldest in the ring): x3 flux difference, du = -dt D
    //      + gravity ----
    const int cz = kk - 3;
    if (own && cz >= o3 && cz < o3 + ez) {
      float d5[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        float acc = part[2][v];
        acc += area3 * (f3hi[v] - f3lo[v]);
        d5[v] = -dt * (acc / vol);
      }
      float rho = win[IDN][2];
      if (kBg) {
        // the background's flux divergence and gravity are the residual
        const int lv = o1 + cx + a.ng;
        d5[IVX] += dt * (rho * a.grav1 + BV.lv(BG_RES, lv));
        rho += BV.lv(BG_RHO_C, lv);
        d5[IPR] += dt * rho * win[IVX][2] * a.grav1;
      } else if (a.grav1 != 0.f) {
        d5[IVX] += dt * rho * a.grav1;
        d5[IPR] += dt * rho * win[IVX][2] * a.grav1;
      }
      if (a.grav2 != 0.f) {
        d5[IVY] += dt * rho * a.grav2;
        d5[IPR] += dt * rho * win[IVY][2] * a.grav2;
      }
      if (a.grav3 != 0.f) {
        d5[IVZ] += dt * rho * a.grav3;
        d5[IPR] += dt * rho * win[IVZ][2] * a.grav3;
      }
      const int q = ((cz + a.ng) * a.n2 + o2 + cy + a.ng) * a.n1 + o1 + cx + a.ng;
#pragma unroll
      for (int v = 0; v < 5; ++v) du[v * ncell + q] = d5[v];
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) f3lo[v] = f3hi[v];

    // ---- x1 and x2 faces of slice kk, then their flux differences ----
    float pn[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (kk >= o3 && kk < o3 + ez) {
      for (int g = threadIdx.x; g < ftot; g += blockDim.x) {
        float wl[5], wr[5], flx[5];
        if (g < nf0) {
          // x1 face below cell (fx, fy) of the slice
          int fx;
          const int fy = dq1.div(g, fx);
          if (kBg) {
            const int gi = o1 + fx;
            bg_x1_face_flux(ws, HS, (fy + 3) * H1, fx + 3, gi + a.ng, BV, refill && gi == 0,
                            refill && gi == a.nx1, sol1, gammad, flx);
          } else {
            face_states<0, P::ty0, P::ty20>(ws, HS, (fy + 3) * H1, 1, fx + 3, type0, ty20,
                                           scale0, wl, wr);
            if (refill) {
              // _revise_x1inner_lr / _revise_x1outer_lr (hydro.cpp:217-229)
              const int gi = o1 + fx;
              if (gi == 0) {
                wl[IDN] = wr[IDN];
                wl[IPR] = wr[IPR];
              }
              if (gi == a.nx1) {
                wr[IDN] = wl[IDN];
                wr[IPR] = wl[IPR];
              }
            }
            riemann_flux<float>(sol1, wl, wr, gammad, 3, flx, (float*)nullptr);
          }
          fs[IDN * ftot + g] = flx[0];
          fs[IVX * ftot + g] = flx[1];
          fs[IVY * ftot + g] = flx[2];
          fs[IVZ * ftot + g] = flx[3];
          fs[IPR * ftot + g] = flx[4];
        } else {
          // x2 face below cell (fx, fy) of the slice
          int fx;
          const int fy = de1.div(g - nf0, fx);
          face_states<1, P::ty1, P::ty21>(ws, HS, fx + 3, H1, fy + 3, type1, ty21, scale1, wl,
                                         wr);
          if (kBg)
            bg_horizontal_flux(BV, o1 + fx + a.ng, sol2, wl, wr, gammad, flx);
          else
            riemann_flux<float>(sol2, wl, wr, gammad, 3, flx, (float*)nullptr);
          fs[IDN * ftot + g] = flx[0];
          fs[IVY * ftot + g] = flx[1];
          fs[IVZ * ftot + g] = flx[2];
          fs[IVX * ftot + g] = flx[3];
          fs[IPR * ftot + g] = flx[4];
        }
      }
      __syncthreads();
      if (own) {
        const int lo1 = cy * (e1 + 1) + cx, lo2 = nf0 + cy * e1 + cx;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          pn[v] += area1 * (fs[v * ftot + lo1 + 1] - fs[v * ftot + lo1]);
          pn[v] += area2 * (fs[v * ftot + lo2 + e1] - fs[v * ftot + lo2]);
        }
      }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      part[2][v] = part[1][v];
      part[1][v] = part[0][v];
      part[0][v] = pn[v];
    }

  }
}

struct ColArgs {
  int n1, n2, n3, ng;
  int nl, ncol, ncol2;
  int bc_in, bc_out;  // x1 boundary functions
  int scheme;         // 0 none, 1 partial, 9 full
  int stage;
  int lanes, lshift;  // cell-parallel kernel: lanes per column (power of two), log2
  int rows2;          // two rows per lane (N = 3 columns of more than 64 cells)
  double gammad, dx1, area, vol, dt, grav1;
  double w0, w1, w2;  // stage weights
  // spherical-polar blocks: the first ghost layer's distance below / above
  // (dx1v[is - 1], dx1v[ie]) and the coordinate
  double zin1, zout1;
  Geometry geo;
  // hydrostatic background of the perturbation path (FP64 rho_b, p_b per x1
  // level and the FP32 tables of bg_impl.hpp), null for the total state
  const double* bgd;
  const float* bgt;
  // device time step; replaces dt when set
  const double* dtp;
};

// primitive state (FP64) of interior-relative layer i of a column whose
// interior cell 0 is at full index q0, through the x1 boundary functions
__device__ __forceinline__ void column_prim(const float* __restrict__ u, int ncell, int q0, int i,
                                            const ColArgs& a, double* W) {
  bool flip = false;
  int ii = bc_map(i, a.nl, a.bc_in, a.bc_out, flip);
  double U[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) U[v] = (double)u[v * ncell + q0 + ii];
  if (flip) U[IVX] = -U[IVX];
  cons_to_prim<double>(U, a.gammad, W);
}

// FP64 primitives of an FP32 conserved state (normal momentum flipped for a
// reflected ghost source)
__device__ __forceinline__ void float_prim(const float* Uf, bool flip, double gammad, double* W) {
  double U[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) U[v] = (double)Uf[v];
  if (flip) U[IVX] = -U[IVX];
  cons_to_prim<double>(U, gammad, W);
}

// FP64 primitives of an FP32 cell of x1 level lvl: on the perturbation
// path the background of that level is added first
// (gno: a gnomonic panel's column, whose cells share cos(theta) = cth)
template <bool kBg>
__device__ __forceinline__ void float_prim_at(const float* Uf, bool flip, const ColArgs& a,
                                              int lvl, double* W, bool gno = false,
                                              double cth = 0.) {
  double U[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) U[v] = (double)Uf[v];
  if (kBg) {
    U[IDN] += a.bgd[lvl];
    U[IPR] += a.bgd[a.n1 + lvl] / (a.gammad - 1.);
  }
  if (flip) U[IVX] = -U[IVX];
  if (gno)
    cons_to_prim_gno<double>(U, a.gammad, cth, W);
  else
    cons_to_prim<double>(U, a.gammad, W);
}

// isentropic refill of the ghost next to boundary cell wb (dir -1: below,
// +1: above; src/hydro/hydro.cpp:231-269, first ghost), with the FP32
// arithmetic of the refill in the RHS kernel, which computes the fluxes of the
// same stage from these ghost values (FP64 on the perturbation path, as
// there).
// (on the perturbation path Ub is the boundary cell's FP32 conserved
// perturbation, refilled as in the RHS kernel)
template <bool kSph = false, bool kBg = false>
__device__ __forceinline__ void refill_ghost(const double* wb, double* wg, const ColArgs& a,
                                             int dir, const float* Ub = nullptr) {
  if (kBg) {
    const BgView V = bg_global_view(a.bgt, BgLayout(a.n1));
    const int lb = dir < 0 ? a.ng : a.ng + a.nl - 1, lg = dir < 0 ? a.ng - 1 : a.ng + a.nl;
    float rg, pg;
    bg_refill_ghost(V, Ub, lb, lg, (float)a.gammad, rg, pg);
    wg[IDN] = a.bgd[lg] + (double)rg;
    wg[IPR] = a.bgd[a.n1 + lg] + (double)pg;
    return;
  }
  float rho, pre;
  if (kSph) {
    const IsentropicRefill R((float)wb[IDN], (float)wb[IPR], (float)a.gammad, -(float)a.grav1,
                             dir);
    R.layer_at((float)(dir < 0 ? a.zin1 : a.zout1), rho, pre);
  } else {
    const IsentropicRefill R((float)wb[IDN], (float)wb[IPR], (float)a.gammad,
                             -(float)a.grav1 * (float)a.dx1, dir);
    R.layer(1, rho, pre);
  }
  wg[IPR] = (double)pre;
  wg[IDN] = (double)rho;
}

template <int N>
__device__ __forceinline__ void column_work(float* __restrict__ u, float* __restrict__ u0,
                                            const float* __restrict__ du,
                                            double* __restrict__ scr, const ColArgs& a, int col) {
  const int ncell = a.n1 * a.n2 * a.n3;
  int j = a.n2 > 1 ? a.ng + col % a.ncol2 : 0;
  int k = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
  int q0 = (k * a.n2 + j) * a.n1 + a.ng;
  const int nl = a.nl;
  const int NN = N * N;
  const int S = NN + N;  // scratch doubles per layer
  const int64_t sstride = a.ncol;

  double sum_a2 = 0.0;
  if (N > 0 && a.scheme) {
    double wm[5], wc[5], wp[5];
    column_prim(u, ncell, q0, 0, a, wc);
    column_prim(u, ncell, q0, -1, a, wm);
    refill_ghost(wc, wm, a, -1);
    double gm1 = a.gammad - 1.;
    double gm1r = 0.5 * (a.gammad + a.gammad) - 1.;
    double Jp[NN], Jc[NN], Jn[NN], Am[NN], Ap[NN];
    flux_jacobian<double, N>(gm1, wm, Jp);
    flux_jacobian<double, N>(gm1, wc, Jc);
    abs_jacobian<double, N>(gm1r, wm, wc, Am);
    double aprev[NN], dprev[N];
    double hv = 0.5 / a.vol;
    for (int i = 0; i < nl; ++i) {
      if (i + 1 < nl) {
        column_prim(u, ncell, q0, i + 1, a, wp);
      } else {
        column_prim(u, ncell, q0, nl, a, wp);
        refill_ghost(wc, wp, a, +1);
      }
      flux_jacobian<double, N>(gm1, wp, Jn);
      abs_jacobian<double, N>(gm1r, wc, wp, Ap);
      double A[NN], B[NN], C[NN];
#pragma unroll
      for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          int o = r * N + c;
          double phi = ((r == 1 && c == 0) || (vic_row(N, r) == IPR && c == 1)) ? a.grav1 : 0.;
          A[o] = (Am[o] * a.area + Ap[o] * a.area + (a.area - a.area) * Jc[o]) * hv - phi;
          if (r == c) A[o] += 1. / a.dt;
          B[o] = -(Am[o] + Jp[o]) * a.area * hv;
          C[o] = -(Ap[o] - Jn[o]) * a.area * hv;
        }
#pragma unroll
      for (int r = 0; r < N; ++r)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          double bnd = c == 1 ? -1. : 1.;
          if (i == 0) A[r * N + c] += B[r * N + c] * bnd;
          if (i == nl - 1) A[r * N + c] += C[r * N + c] * bnd;
        }
      double rhs[N];
      int q = q0 + i;
      rhs[0] = (double)du[IDN * ncell + q] / a.dt;
      rhs[1] = (double)du[IVX * ncell + q] / a.dt;
      if (N == 3) {
        rhs[2] = (double)du[IPR * ncell + q] / a.dt;
      } else {
        rhs[2] = (double)du[IVY * ncell + q] / a.dt;
        rhs[3] = (double)du[IVZ * ncell + q] / a.dt;
        rhs[4] = (double)du[IPR * ncell + q] / a.dt;
      }
      if (i > 0) {
        double tmp[NN], v[N];
        matmul<double, N>(B, aprev, tmp);
#pragma unroll
        for (int e = 0; e < NN; ++e) A[e] -= tmp[e];
        matvec<double, N>(B, dprev, v);
#pragma unroll
        for (int e = 0; e < N; ++e) rhs[e] -= v[e];
      }
      if (N == 3) {
        double inv[9];
        inverse3<double>(A, inv);
        matvec<double, 3>(inv, rhs, dprev);
        matmul<double, 3>(inv, C, aprev);
      } else {
        int indx[5];
        ludcmp5<double>(A, indx);
        double solved[30];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
#pragma unroll
          for (int c = 0; c < 5; ++c) solved[r * 6 + c] = C[r * 5 + c];
          solved[r * 6 + 5] = rhs[r];
        }
        lubksb5<double, 6>(A, indx, solved);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
#pragma unroll
          for (int c = 0; c < 5; ++c) aprev[r * 5 + c] = solved[r * 6 + c];
          dprev[r] = solved[r * 6 + 5];
        }
      }
      double* sp = scr + (int64_t)i * S * sstride + col;
#pragma unroll
      for (int e = 0; e < NN; ++e) sp[e * sstride] = aprev[e];
#pragma unroll
      for (int e = 0; e < N; ++e) sp[(NN + e) * sstride] = dprev[e];
      double ai = wc[IDN] * a.vol;
      sum_a2 += ai * ai;
      // roll the interface data up one layer
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        Jp[e] = Jc[e];
        Jc[e] = Jn[e];
        Am[e] = Ap[e];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        wm[v] = wc[v];
        wc[v] = wp[v];
      }
    }
    // back substitution; delta_i overwrites the scratch slot of layer i
    double b_dry = 0.0;
    {
      int q = q0 + nl - 1;
      b_dry += ((double)du[IDN * ncell + q] - dprev[0]) * 1.0 * a.vol;
    }
    for (int i = nl - 2; i >= 0; --i) {
      double* sp = scr + (int64_t)i * S * sstride + col;
      double Ai[NN], d[N], v[N];
#pragma unroll
      for (int e = 0; e < NN; ++e) Ai[e] = sp[e * sstride];
#pragma unroll
      for (int e = 0; e < N; ++e) d[e] = sp[(NN + e) * sstride];
      matvec<double, N>(Ai, dprev, v);
#pragma unroll
      for (int e = 0; e < N; ++e) {
        dprev[e] = d[e] - v[e];
        sp[(NN + e) * sstride] = dprev[e];
      }
      int q = q0 + i;
      b_dry += ((double)du[IDN * ncell + q] - dprev[0]) * 1.0 * a.vol;
    }
    // redistribution + stage combine
    for (int i = 0; i < nl; ++i) {
      int q = q0 + i;
      double* sp = scr + (int64_t)i * S * sstride + col;
      double d[N];
#pragma unroll
      for (int e = 0; e < N; ++e) d[e] = sp[(NN + e) * sstride];
      double rho = (double)u[IDN * ncell + q];
      double a_i = rho * a.vol;
      double dd[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) dd[v] = (double)du[v * ncell + q];
      double fix = d[0] - dd[IDN];
      dd[IDN] = (dd[IDN] + fix) + rho * a_i * b_dry / sum_a2;
      dd[IVX] = d[1];
      if (N == 3) {
        dd[IPR] = d[2];
      } else {
        dd[IVY] = d[2];
        dd[IVZ] = d[3];
        dd[IPR] = d[4];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        float uc = u[v * ncell + q];
        float ub = a.stage == 0 ? uc : u0[v * ncell + q];
        if (a.stage == 0) u0[v * ncell + q] = uc;
        u[v * ncell + q] = (float)((a.w0 * ub + a.w1 * uc) + a.w2 * dd[v]);
      }
    }
  } else {
    for (int i = 0; i < nl; ++i) {
      int q = q0 + i;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        float uc = u[v * ncell + q];
        float ub = a.stage == 0 ? uc : u0[v * ncell + q];
        if (a.stage == 0) u0[v * ncell + q] = uc;
        u[v * ncell + q] =
            (float)((a.w0 * ub + a.w1 * uc) + a.w2 * (double)du[v * ncell + q]);
      }
    }
  }
}

template <int N>
__global__ void __launch_bounds__(kColThreads)
    fused_column_kernel(float* __restrict__ u, float* __restrict__ u0,
                        const float* __restrict__ du, double* __restrict__ scr, ColArgs a) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncol) return;
  if (a.dtp) a.dt = *a.dtp;
  column_work<N>(u, u0, du, scr, a, col);
}

// ------------------------------------------- cell-parallel column solve ----
//
// One lane per cell: a workgroup holds T / L whole columns, L = lanes per
// column (power of two >= nl).  Primitives and interface |A| are exchanged
// through LDS, every lane assembles its own block row, and the
// block-tridiagonal system is solved by parallel cyclic reduction (log2 nl
// levels; at each level row i eliminates its couplings to rows i -+ s with
// the normalised rows of its neighbours).  The MS-VIC column sums are wave
// reductions, then the lane applies the redistribution and the stage
// combine to its own cell.  Arithmetic in FP64, as in the serial sweep.

// Every thread of the workgroup calls this with the same ColArgs; lane t
// serves layer i = t mod L of column col (col < a.ncol checked here).
// flux_jacobian with the enthalpy term (gamma / (gamma - 1)) p / rho formed
// once through rcp_fast: the same matrix to an ulp or two
template <int N>
__device__ __forceinline__ void flux_jacobian_fast(double gm1, const double* w, double* J) {
  const double ht = (gm1 + 1.) * rcp_fast(gm1) * w[IPR] * rcp_fast(w[IDN]);
  const double v1 = w[IVX], v2 = w[IVY], v3 = w[IVZ];
  const double s2 = (v1 * v1 + v2 * v2) + v3 * v3;
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      const int rv = vic_row(N, r), cv = vic_row(N, c);
      double x;
      if (rv == 4 && cv == 0)
        x = ((gm1 - 1.) * s2 * 0.5 - ht) * v1;
      else if (rv == 4 && cv == 1)
        x = (ht + s2 * 0.5) - gm1 * v1 * v1;
      else
        x = flux_jac(rv, cv, gm1, w[IDN], v1, v2, v3, w[IPR]);
      J[r * N + c] = x;
    }
}

// Block row i of a column's system (vic_assemble_cell, uniform Cartesian
// areas): wm / wp the primitives below / above, Am / Ap the interface |A|
// below / above, Dv the explicit tendency of the cell; reflecting column ends
// fold the ghost coupling into the diagonal; inactive rows are identities.
template <int N>
__device__ __forceinline__ void block_row(const ColArgs& a, int i, int nl, bool act,
                                          const double* wm, const double* wp, const double* Am,
                                          const double* Ap, const float* Dv, double gm1,
                                          double idt, double* A, double* B, double* C,
                                          double* r) {
  constexpr int NN = N * N;
  double Jp[NN], Jn[NN];
  flux_jacobian_fast<N>(gm1, wm, Jp);
  flux_jacobian_fast<N>(gm1, wp, Jn);
  // area * 0.5 / V (the FP64 plain path keeps the reference's order)
  const double ch = a.area * (0.5 * rcp_fast(a.vol));
#pragma unroll
  for (int rr = 0; rr < N; ++rr)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      int o = rr * N + c;
      double phi = ((rr == 1 && c == 0) || (vic_row(N, rr) == IPR && c == 1)) ? a.grav1 : 0.;
      A[o] = (Am[o] + Ap[o]) * ch - phi;
      if (rr == c) A[o] += idt;
      B[o] = -(Am[o] + Jp[o]) * ch;
      C[o] = -(Ap[o] - Jn[o]) * ch;
    }
#pragma unroll
  for (int rr = 0; rr < N; ++rr)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      double bnd = c == 1 ? -1. : 1.;
      if (i == 0) A[rr * N + c] += B[rr * N + c] * bnd;
      if (i == nl - 1) A[rr * N + c] += C[rr * N + c] * bnd;
    }
  if (i == 0)
#pragma unroll
    for (int e = 0; e < NN; ++e) B[e] = 0.;
  if (i == nl - 1)
#pragma unroll
    for (int e = 0; e < NN; ++e) C[e] = 0.;
  r[0] = (double)Dv[IDN] * idt;
  r[1] = (double)Dv[IVX] * idt;
  if (N == 3) {
    r[2] = (double)Dv[IPR] * idt;
  } else {
    r[2] = (double)Dv[IVY] * idt;
    r[3] = (double)Dv[IVZ] * idt;
    r[4] = (double)Dv[IPR] * idt;
  }
  if (!act) {
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      A[e] = (e % (N + 1)) == 0 ? 1. : 0.;
      B[e] = C[e] = 0.;
    }
#pragma unroll
    for (int e = 0; e < N; ++e) r[e] = 0.;
  }
}

// MS-VIC redistribution of one cell (vic_redistribute_impl.h:66-106) from its
// solution d and the column sums, then the SSP-RK3 stage combine and store
__device__ __forceinline__ void finish_cell(float* __restrict__ u, float* __restrict__ u0,
                                            const ColArgs& a, int ncell, int q, int N,
                                            const double* d, double rho, double ai, double s_a2,
                                            double s_b, const float* Dv, const float* Uc,
                                            const float* Ub) {
  double dd[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) dd[v] = (double)Dv[v];
  if (a.scheme) {
    double fix = d[0] - dd[IDN];
    dd[IDN] = (dd[IDN] + fix) + rho * ai * s_b * rcp_fast(s_a2);
    dd[IVX] = d[1];
    if (N == 3) {
      dd[IPR] = d[2];
    } else {
      dd[IVY] = d[2];
      dd[IVZ] = d[3];
      dd[IPR] = d[4];
    }
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    float uc = Uc[v];
    float ub = a.stage == 0 ? uc : Ub[v];
    if (a.stage == 0) u0[v * ncell + q] = uc;
    u[v * ncell + q] = (float)((a.w0 * ub + a.w1 * uc) + a.w2 * dd[v]);
  }
}

// K2 for 3x3 columns of more than 64 cells: two rows per lane (rows 2m and
// 2m + 1 in lane m of the column).  One cyclic-reduction level inside each
// lane removes the even rows: the odd row eliminates the even row below it
// from its own registers and the even row above it from the next lane
// (through LDS); the odd rows, one per lane, are then solved by parallel
// cyclic reduction and the even rows follow by back substitution.  Against
// one row per lane this halves the PCR work and the lanes (a column of up to
// 128 cells is one wavefront, without workgroup barriers).
template <int N, bool kBg = false>
__device__ __forceinline__ void column_pcr2(float* __restrict__ u, float* __restrict__ u0,
                                            const float* __restrict__ du, const ColArgs& a,
                                            int col, double* __restrict__ sh) {
  constexpr int NN = N * N;
  const int T = blockDim.x, t = threadIdx.x;
  const int ncell = a.n1 * a.n2 * a.n3;
  const int nl = a.nl;
  const int half = a.lanes;  // lanes per column
  const int m = t & (half - 1);
  const int i0 = 2 * m, i1 = 2 * m + 1;
  const bool colok = col < a.ncol;
  const bool act0 = colok && i0 < nl, act1 = colok && i1 < nl;
  int qb = 0;  // full index of the column's interior cell 0
  if (colok) {
    int j = a.n2 > 1 ? a.ng + col % a.ncol2 : 0;
    int k = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
    qb = (k * a.n2 + j) * a.n1 + a.ng;
  }
  const int q0 = qb + i0, q1 = qb + i1;

  // ---- every global load of the lane first ----
  float Uc0[5] = {1.f, 0.f, 0.f, 0.f, 1.f}, Uc1[5] = {1.f, 0.f, 0.f, 0.f, 1.f};
  float Ub0[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, Ub1[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float Dv0[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, Dv1[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float Um[5] = {1.f, 0.f, 0.f, 0.f, 1.f}, Up[5] = {1.f, 0.f, 0.f, 0.f, 1.f};
  bool flm = false, flp = false;
  int lm = a.ng, lp = a.ng;  // x1 levels of the ghost sources
  const bool top0 = act0 && i0 == nl - 1, top1 = act1 && i1 == nl - 1;
  if (act0) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      Uc0[v] = u[v * ncell + q0];
      Dv0[v] = du[v * ncell + q0];
      if (a.stage != 0) Ub0[v] = u0[v * ncell + q0];
    }
    if (a.scheme && i0 == 0) {
      int ii = bc_map(-1, nl, a.bc_in, a.bc_out, flm);
      lm = a.ng + ii;
#pragma unroll
      for (int v = 0; v < 5; ++v) Um[v] = u[v * ncell + qb + ii];
    }
  }
  if (act1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      Uc1[v] = u[v * ncell + q1];
      Dv1[v] = du[v * ncell + q1];
      if (a.stage != 0) Ub1[v] = u0[v * ncell + q1];
    }
  }
  if (a.scheme && (top0 || top1)) {
    int ii = bc_map(nl, nl, a.bc_in, a.bc_out, flp);
    lp = a.ng + ii;
#pragma unroll
    for (int v = 0; v < 5; ++v) Up[v] = u[v * ncell + qb + ii];
  }

  double d0[N], d1[N];
#pragma unroll
  for (int e = 0; e < N; ++e) d0[e] = d1[e] = 0.;
  double w0[5], w1[5];
  float_prim_at<kBg>(Uc0, false, a, a.ng + (act0 ? i0 : 0), w0);
  float_prim_at<kBg>(Uc1, false, a, a.ng + (act1 ? i1 : 0), w1);
  double s_a2 = 1., s_b = 0.;
  // every column within one wavefront (half <= 64 lanes, aligned)
  const bool wave = half <= 64;
  if (a.scheme) {
    double Pb0[NN], Pc0[NN], pr0[N];  // the even row, normalised
    // ---- primitives of the neighbours ----
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      sh[v * T + t] = w0[v];
      sh[(5 + v) * T + t] = w1[v];
    }
    column_barrier(wave);
    double wm0[5], wp0[5], wp1[5];  // below row i0, above row i0, above row i1
#pragma unroll
    for (int v = 0; v < 5; ++v) wm0[v] = wp0[v] = wp1[v] = w1[v];
    if (act0 && m > 0) {
#pragma unroll
      for (int v = 0; v < 5; ++v) wm0[v] = sh[(5 + v) * T + t - 1];
    }
    if (act1 && !top1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) wp1[v] = sh[v * T + t + 1];
    }
    {
      // The physical ghosts of the column ends in one pass for the
      // wavefront (a column is one wavefront, so its bottom and top lanes
      // would otherwise take the refill's branch one after the other): lane
      // 0 the ghost below row 0, the top lane the ghost above row nl - 1.
      const bool lo = act0 && m == 0, hi = top0 || top1;
      if (lo || hi) {
        const bool b1 = !lo && top1;  // the boundary row is i1 (else i0)
        float Us[5], Ubd[5];
        double wb[5], g[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          Us[v] = lo ? Um[v] : Up[v];
          Ubd[v] = b1 ? Uc1[v] : Uc0[v];
          wb[v] = b1 ? w1[v] : w0[v];
        }
        float_prim_at<kBg>(Us, lo ? flm : flp, a, lo ? lm : lp, g);
        refill_ghost<false, kBg>(wb, g, a, lo ? -1 : +1, Ubd);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          if (lo)
            wm0[v] = g[v];
          else if (top0)
            wp0[v] = g[v];
          else
            wp1[v] = g[v];
        }
      }
      if (lo && hi) {
        // columns of one or two cells: lane 0 has the top ghost as well
        double g[5];
        float_prim_at<kBg>(Up, flp, a, lp, g);
        refill_ghost<false, kBg>(top0 ? w0 : w1, g, a, +1, top0 ? Uc0 : Uc1);
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          if (top0)
            wp0[v] = g[v];
          else
            wp1[v] = g[v];
        }
      }
    }
    column_barrier(wave);
    // ---- interface |A| of the faces below rows i0, i1 and above i1 ----
    const double gm1 = a.gammad - 1.;
    const double idt = rcp_fast(a.dtp ? *a.dtp : a.dt);
    const double gm1r = 0.5 * (a.gammad + a.gammad) - 1.;
    double Af1[NN], Af2[NN];
    {
      double Af0[NN];
      abs_jacobian_fast<double, N>(gm1r, wm0, w0, Af0);
#pragma unroll
      for (int e = 0; e < NN; ++e) sh[e * T + t] = Af0[e];
      column_barrier(wave);
      abs_jacobian_fast<double, N>(gm1r, w0, wp0, Af1);
      if (act1 && !top1) {
#pragma unroll
        for (int e = 0; e < NN; ++e) Af2[e] = sh[e * T + t + 1];
      } else {
        abs_jacobian_fast<double, N>(gm1r, w1, wp1, Af2);
      }
      column_barrier(wave);
      // ---- even row: block row, normalised to registers and LDS ----
      double A0[NN], B0[NN], C0[NN], r0[N];
      block_row<N>(a, i0, nl, act0, wm0, wp0, Af0, Af1, Dv0, gm1, idt, A0, B0, C0, r0);
      double inv[NN];
      inverse3<double, true>(A0, inv);
      matmul<double, N>(inv, B0, Pb0);
      matmul<double, N>(inv, C0, Pc0);
      matvec<double, N>(inv, r0, pr0);
    }
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      sh[e * T + t] = Pb0[e];
      sh[(NN + e) * T + t] = Pc0[e];
    }
#pragma unroll
    for (int e = 0; e < N; ++e) sh[(2 * NN + e) * T + t] = pr0[e];
    // ---- odd row ----
    double A1[NN], B1[NN], C1[NN], r1[N];
    block_row<N>(a, i1, nl, act1, w0, wp1, Af1, Af2, Dv1, gm1, idt, A1, B1, C1, r1);
    column_barrier(wave);
    {
      double nB[NN], nC[NN], y[NN];
      // the even row below (own registers) couples row i1 to row i1 - 2
      matmul<double, N>(B1, Pb0, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) nB[e] = -y[e];
      matmul<double, N>(B1, Pc0, y);
#pragma unroll
      for (int e = 0; e < NN; ++e) A1[e] -= y[e];
      matvec<double, N>(B1, pr0, y);
#pragma unroll
      for (int e = 0; e < N; ++e) r1[e] -= y[e];
      // the even row above (next lane) couples it to row i1 + 2
#pragma unroll
      for (int e = 0; e < NN; ++e) nC[e] = 0.;
      if (m + 1 < half) {
        double Pb[NN], Pc[NN], pr[N];
#pragma unroll
        for (int e = 0; e < NN; ++e) {
          Pb[e] = sh[e * T + t + 1];
          Pc[e] = sh[(NN + e) * T + t + 1];
        }
#pragma unroll
        for (int e = 0; e < N; ++e) pr[e] = sh[(2 * NN + e) * T + t + 1];
        matmul<double, N>(C1, Pc, y);
#pragma unroll
        for (int e = 0; e < NN; ++e) nC[e] = -y[e];
        matmul<double, N>(C1, Pb, y);
#pragma unroll
        for (int e = 0; e < NN; ++e) A1[e] -= y[e];
        matvec<double, N>(C1, pr, y);
#pragma unroll
        for (int e = 0; e < N; ++e) r1[e] -= y[e];
      }
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        B1[e] = nB[e];
        C1[e] = nC[e];
      }
    }
    column_barrier(wave);
    // ---- the odd rows by parallel cyclic reduction (lane m = row 2m + 1) ----
    pcr_solve<N, true>(A1, B1, C1, r1, d1, sh, T, t, m, nl / 2, act1, wave);
    // ---- even rows by back substitution ----
#pragma unroll
    for (int e = 0; e < N; ++e) sh[e * T + t] = d1[e];
    column_barrier(wave);
    double xb[N];
#pragma unroll
    for (int e = 0; e < N; ++e) xb[e] = m > 0 ? sh[e * T + t - 1] : 0.;
#pragma unroll
    for (int rr = 0; rr < N; ++rr) {
      double v = pr0[rr];
#pragma unroll
      for (int c = 0; c < N; ++c) v -= Pb0[rr * N + c] * xb[c] + Pc0[rr * N + c] * d1[c];
      d0[rr] = v;
    }
    // ---- MS-VIC column sums ----
    const double ai0 = w0[IDN] * a.vol, ai1 = w1[IDN] * a.vol;
    double sums[2] = {(act0 ? ai0 * ai0 : 0.) + (act1 ? ai1 * ai1 : 0.),
                      (act0 ? ((double)Dv0[IDN] - d0[0]) * 1.0 * a.vol : 0.) +
                          (act1 ? ((double)Dv1[IDN] - d1[0]) * 1.0 * a.vol : 0.)};
    column_sums<2>(sums, sh, t, a.lanes, a.lshift);
    s_a2 = sums[0], s_b = sums[1];
  }
  if (act0)
    finish_cell(u, u0, a, ncell, q0, N, d0, w0[IDN], w0[IDN] * a.vol, s_a2, s_b, Dv0, Uc0, Ub0);
  if (act1)
    finish_cell(u, u0, a, ncell, q1, N, d1, w1[IDN], w1[IDN] * a.vol, s_a2, s_b, Dv1, Uc1, Ub1);
}

// kSph: spherical-polar blocks, per-layer face areas and volumes
// (vic_assemble_cell with area_i != area_i+1) and the uneven ghost refill
template <int N, int kSph = 0, bool kBg = false>
__device__ __forceinline__ void column_pcr(float* __restrict__ u, float* __restrict__ u0,
                                           const float* __restrict__ du, const ColArgs& a,
                                           int col, double* __restrict__ sh) {
  constexpr int NN = N * N;
  const int T = blockDim.x, t = threadIdx.x;
  const int ncell = a.n1 * a.n2 * a.n3;
  const int nl = a.nl;
  const int i = t & (a.lanes - 1);
  const bool colok = col < a.ncol;
  const bool act = colok && i < nl;
  int q = 0;
  if (colok) {
    int j = a.n2 > 1 ? a.ng + col % a.ncol2 : 0;
    int k = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
    q = (k * a.n2 + j) * a.n1 + a.ng + i;
  }
  const int q0 = q - i;

  // ---- every global load of the lane first: one memory round trip ----
  float Uc[5] = {1.f, 0.f, 0.f, 0.f, 1.f}, Ub[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  float Um[5] = {1.f, 0.f, 0.f, 0.f, 1.f}, Up[5] = {1.f, 0.f, 0.f, 0.f, 1.f};
  float Dv[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  bool flm = false, flp = false;
  int lm = a.ng, lp = a.ng;  // x1 levels of the ghost sources
  if (act) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      Uc[v] = u[v * ncell + q];
      Dv[v] = du[v * ncell + q];
      if (a.stage != 0) Ub[v] = u0[v * ncell + q];
    }
    // the x1 ghost sources of the end lanes, through the boundary functions
    if (a.scheme && i == 0) {
      int ii = bc_map(-1, nl, a.bc_in, a.bc_out, flm);
      lm = a.ng + ii;
#pragma unroll
      for (int v = 0; v < 5; ++v) Um[v] = u[v * ncell + q0 + ii];
    }
    if (a.scheme && i == nl - 1) {
      int ii = bc_map(nl, nl, a.bc_in, a.bc_out, flp);
      lp = a.ng + ii;
#pragma unroll
      for (int v = 0; v < 5; ++v) Up[v] = u[v * ncell + q0 + ii];
    }
  }

  double dd[5];
  // every column inside one wavefront: no workgroup barriers
  const bool wave = a.lanes == 64;
  // a gnomonic panel's column: one cos(theta) for all its cells
  constexpr bool gno = kSph == 2;
  const int cj = a.n2 > 1 ? a.ng + col % a.ncol2 : 0, ck = a.n3 > 1 ? a.ng + col / a.ncol2 : 0;
  const double cth = gno ? gno_cos_cell(a.geo, ck, cj) : 0.;
  if (a.scheme) {
    // ---- primitives of the cell and of its neighbours ----
    double wc[5], wm[5], wp[5];
    float_prim_at<kBg>(Uc, false, a, a.ng + (act ? i : 0), wc, gno, cth);
#pragma unroll
    for (int v = 0; v < 5; ++v) sh[v * T + t] = wc[v];
    column_barrier(wave);
#pragma unroll
    for (int v = 0; v < 5; ++v) wm[v] = wp[v] = wc[v];
    if (act) {
      if (i > 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) wm[v] = sh[v * T + t - 1];
      } else {
        float_prim_at<kBg>(Um, flm, a, lm, wm, gno, cth);
        refill_ghost<(kSph != 0), kBg>(wc, wm, a, -1, Uc);
      }
      if (i < nl - 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) wp[v] = sh[v * T + t + 1];
      } else {
        float_prim_at<kBg>(Up, flp, a, lp, wp, gno, cth);
        refill_ghost<(kSph != 0), kBg>(wc, wp, a, +1, Uc);
      }
    }
    column_barrier(wave);
    // ---- interface |A|: lane i owns the face below its cell ----
    const double gm1 = a.gammad - 1.;
    const double idt = rcp_fast(a.dtp ? *a.dtp : a.dt);
    const double gm1r = 0.5 * (a.gammad + a.gammad) - 1.;
    double Am[NN], Ap[NN];
    abs_jacobian_fast<double, N>(gm1r, wm, wc, Am);
#pragma unroll
    for (int e = 0; e < NN; ++e) sh[e * T + t] = Am[e];
    column_barrier(wave);
    if (act && i < nl - 1) {
#pragma unroll
      for (int e = 0; e < NN; ++e) Ap[e] = sh[e * T + t + 1];
    } else {
      abs_jacobian_fast<double, N>(gm1r, wc, wp, Ap);
    }
    column_barrier(wave);
    // ---- block row i (vic_assemble_cell; uniform Cartesian areas) ----
    double A[NN], B[NN], C[NN], r[N];
    // face areas below / above the layer and its volume
    double ar_i = a.area, ar_p = a.area, vol_i = a.vol;
    if (gno) {
      const GnoCell gc = gno_cell(a.geo, ck, cj, a.ng + i);
      ar_i = gc.a1m, ar_p = gc.a1p, vol_i = gc.vol;
    } else if (kSph) {
      const SphCell gc = sph_cell(a.geo, ck, cj, a.ng + i);
      ar_i = gc.a1m, ar_p = gc.a1p, vol_i = gc.vol;
    }
    {
      double Jp[NN], Jn[NN], Jc[NN];
      flux_jacobian<double, N>(gm1, wm, Jp);
      flux_jacobian<double, N>(gm1, wp, Jn);
      if (kSph) flux_jacobian<double, N>(gm1, wc, Jc);
      const double hv = 0.5 * rcp_fast(vol_i);
#pragma unroll
      for (int rr = 0; rr < N; ++rr)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          int o = rr * N + c;
          double phi = ((rr == 1 && c == 0) || (vic_row(N, rr) == IPR && c == 1)) ? a.grav1 : 0.;
          if (kSph)
            A[o] = (Am[o] * ar_i + Ap[o] * ar_p + (ar_p - ar_i) * Jc[o]) * hv - phi;
          else
            A[o] = (Am[o] * ar_i + Ap[o] * ar_p) * hv - phi;
          if (rr == c) A[o] += idt;
          B[o] = -(Am[o] + Jp[o]) * ar_i * hv;
          C[o] = -(Ap[o] - Jn[o]) * ar_p * hv;
        }
      // reflecting column ends fold the ghost coupling into the diagonal
#pragma unroll
      for (int rr = 0; rr < N; ++rr)
#pragma unroll
        for (int c = 0; c < N; ++c) {
          double bnd = c == 1 ? -1. : 1.;
          if (i == 0) A[rr * N + c] += B[rr * N + c] * bnd;
          if (i == nl - 1) A[rr * N + c] += C[rr * N + c] * bnd;
        }
      if (i == 0)
#pragma unroll
        for (int e = 0; e < NN; ++e) B[e] = 0.;
      if (i == nl - 1)
#pragma unroll
        for (int e = 0; e < NN; ++e) C[e] = 0.;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) dd[v] = (double)Dv[v];
    r[0] = dd[IDN] * idt;
    r[1] = dd[IVX] * idt;
    if (N == 3) {
      r[2] = dd[IPR] * idt;
    } else {
      r[2] = dd[IVY] * idt;
      r[3] = dd[IVZ] * idt;
      r[4] = dd[IPR] * idt;
    }
    if (!act) {
#pragma unroll
      for (int e = 0; e < NN; ++e) {
        A[e] = (e % (N + 1)) == 0 ? 1. : 0.;
        B[e] = C[e] = 0.;
      }
#pragma unroll
      for (int e = 0; e < N; ++e) r[e] = 0.;
    }
    // ---- parallel cyclic reduction, then delta = A^-1 r ----
    double d[N];
    pcr_solve<N, true>(A, B, C, r, d, sh, T, t, i, nl, act, wave);
    // ---- MS-VIC column sums ----
    double ai = wc[IDN] * vol_i;
    double sums[2] = {act ? ai * ai : 0., act ? (dd[IDN] - d[0]) * 1.0 * vol_i : 0.};
    column_sums<2>(sums, sh, t, a.lanes, a.lshift);
    const double s_a2 = sums[0], s_b = sums[1];
    double rho = wc[IDN];
    double fix = d[0] - dd[IDN];
    dd[IDN] = (dd[IDN] + fix) + rho * ai * s_b * rcp_fast(s_a2);
    dd[IVX] = d[1];
    if (N == 3) {
      dd[IPR] = d[2];
    } else {
      dd[IVY] = d[2];
      dd[IVZ] = d[3];
      dd[IPR] = d[4];
    }
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) dd[v] = (double)Dv[v];
  }
  // ---- stage combine ----
  if (act) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      float uc = Uc[v];
      float ub = a.stage == 0 ? uc : Ub[v];
      if (a.stage == 0) u0[v * ncell + q] = uc;
      u[v * ncell + q] = (float)((a.w0 * ub + a.w1 * uc) + a.w2 * dd[v]);
    }
  }
}

// kBg: perturbation state about the hydrostatic background a.bgd
template <int N, bool kBg = false>
__global__ void __launch_bounds__(PcrShape<N>::kThreads)
    fused_column_pcr_kernel(float* __restrict__ u, float* __restrict__ u0,
                            const float* __restrict__ du, ColArgs a) {
  extern __shared__ double sh[];
  int col = blockIdx.x * (blockDim.x >> a.lshift) + (threadIdx.x >> a.lshift);
  column_pcr<N, false, kBg>(u, u0, du, a, col, sh);
}

// K2 of spherical-polar blocks
template <int N>
__global__ void __launch_bounds__(PcrShape<N>::kThreads)
    fused_column_pcr_sph_kernel(float* __restrict__ u, float* __restrict__ u0,
                                const float* __restrict__ du, ColArgs a) {
  extern __shared__ double sh[];
  int col = blockIdx.x * (blockDim.x >> a.lshift) + (threadIdx.x >> a.lshift);
  column_pcr<N, 1>(u, u0, du, a, col, sh);
}

// K2 of gnomonic panels
template <int N>
__global__ void __launch_bounds__(PcrShape<N>::kThreads)
    fused_column_pcr_gno_kernel(float* __restrict__ u, float* __restrict__ u0,
                                const float* __restrict__ du, ColArgs a) {
  extern __shared__ double sh[];
  int col = blockIdx.x * (blockDim.x >> a.lshift) + (threadIdx.x >> a.lshift);
  column_pcr<N, 2>(u, u0, du, a, col, sh);
}

template <bool kBg = false>
__global__ void __launch_bounds__(PcrShape<3>::kThreads, 2)
    fused_column_pcr2_kernel(float* __restrict__ u, float* __restrict__ u0,
                             const float* __restrict__ du, ColArgs a) {
  extern __shared__ double sh[];
  int col = blockIdx.x * (blockDim.x >> a.lshift) + (threadIdx.x >> a.lshift);
  column_pcr2<3, kBg>(u, u0, du, a, col, sh);
}

// ------------------------------------------------------- batched members ----

// One member of a batch: its block geometry / options, where its fields start
// in the concatenated arrays, and where its work items start in the
// persistent kernels' work lists.
struct BatchMember {
  FusedArgs a;
  ColArgs c;
  int64_t off;      // element offset of the member's [5][n3][n2][n1] field
  int64_t scr_off;  // double offset of the member's column scratch
  int tile0;        // first K1 work item
  int col0;         // first K2 work item of the lane-per-column kernel
                    // (multiple of the wavefront size)
  int grp0;         // first K2 work item of the cell-parallel kernel (one
                    // workgroup of whole columns)
  int grp2;         // first work item of the two-rows-per-lane column kernel
};

enum { kListTiles = 0, kListCols = 1, kListGroups = 2, kListGroups2 = 3 };

__device__ __forceinline__ int find_member(const BatchMember* m, int nm, int item, int list) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    // members outside a list start where the next member of the list
    // starts; among equal starts the last member is taken, which is the one
    // that owns the items
    int start = list == kListTiles  ? m[mid].tile0
                : list == kListCols ? m[mid].col0
                : list == kListGroups ? m[mid].grp0
                                      : m[mid].grp2;
    if (start <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// persistent K1 over the (member, tile) work list; kSph: every member is a
// spherical-polar block (merged face pass with geometric sources)
template <int kSph, class P = MarchGeneric>
__global__ void __launch_bounds__(kRhsThreads)
    batch_rhs_kernel(const float* __restrict__ u, float* __restrict__ du,
                     const BatchMember* __restrict__ mem, int nm, int total) {
  extern __shared__ float smem[];
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    int m = find_member(mem, nm, item, kListTiles);
    // (curvilinear members read their arguments in place, Cartesian ones
    // from a copy: either way without a copy in scratch memory)
    RhsArgs<kSph> a = mem[m].a;
    rhs_tile<kSph ? 1 : -1, kSph, false, P>(u + mem[m].off, du + mem[m].off, a,
                                            item - mem[m].tile0, smem, a.dt);
    __syncthreads();
  }
}

using BatchRhsKernel = void (*)(const float*, float*, const BatchMember*, int, int);

// K1 of a batch: the instance of the members' coordinate system, and for
// Cartesian members the compiled face-option instance that every member
// matches (else the generic one)
static BatchRhsKernel batch_rhs_for(const snapx_config* members, int nm) {
  if (members[0].coord == SNAPX_COORD_GNOMONIC_EQUIANGLE) return batch_rhs_kernel<2>;
  if (members[0].coord != SNAPX_COORD_CARTESIAN) return batch_rhs_kernel<1>;
  int pol = face_policy(&members[0]);
  for (int m = 1; m < nm && pol != kFacesGeneric; ++m)
    if (face_policy(&members[m]) != pol) pol = kFacesGeneric;
  switch (pol) {
    case kFacesLmRoRo: return batch_rhs_kernel<0, FacesLmRoRo>;
    case kFacesLm: return batch_rhs_kernel<0, FacesLm>;
    case kFacesRo: return batch_rhs_kernel<0, FacesRo>;
    case kFacesHl: return batch_rhs_kernel<0, FacesHl>;
    default: return batch_rhs_kernel<0>;
  }
}

// persistent K2 over the (member, column) work list; member column ranges are
// padded to whole wavefronts so that a wavefront serves one member
template <int N>
__global__ void __launch_bounds__(kColThreads)
    batch_column_kernel(float* __restrict__ u, float* __restrict__ u0,
                        const float* __restrict__ du, double* __restrict__ scr,
                        const BatchMember* __restrict__ mem, int nm, int total, int stage) {
  for (int item = blockIdx.x * blockDim.x + threadIdx.x; item < total;
       item += gridDim.x * blockDim.x) {
    int m = find_member(mem, nm, item, kListCols);
    ColArgs c = mem[m].c;
    int col = item - mem[m].col0;
    if (col >= c.ncol) continue;
    static const double W[3][3] = {
        {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};
    c.stage = stage;
    c.w0 = W[stage][0], c.w1 = W[stage][1], c.w2 = W[stage][2];
    int64_t o = mem[m].off;
    column_work<N>(u + o, u0 + o, du + o, scr + mem[m].scr_off, c, col);
  }
}

// persistent cell-parallel K2 over the (member, column group) work list
template <int N, int kSph>
__global__ void __launch_bounds__(PcrShape<N>::kThreads)
    batch_column_pcr_kernel(float* __restrict__ u, float* __restrict__ u0,
                            const float* __restrict__ du, const BatchMember* __restrict__ mem,
                            int nm, int total, int stage) {
  extern __shared__ double sh[];
  static const double W[3][3] = {
      {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    int m = find_member(mem, nm, item, kListGroups);
    ColArgs c = mem[m].c;
    c.stage = stage;
    c.w0 = W[stage][0], c.w1 = W[stage][1], c.w2 = W[stage][2];
    int col = (item - mem[m].grp0) * (blockDim.x >> c.lshift) + (threadIdx.x >> c.lshift);
    int64_t o = mem[m].off;
    column_pcr<N, kSph>(u + o, u0 + o, du + o, c, col, sh);
    __syncthreads();  // LDS is reused by the next item
  }
}

// persistent two-rows-per-lane K2 (3x3) over its (member, column group) list
__global__ void __launch_bounds__(PcrShape<3>::kThreads, 2)
    batch_column_pcr2_kernel(float* __restrict__ u, float* __restrict__ u0,
                             const float* __restrict__ du, const BatchMember* __restrict__ mem,
                             int nm, int total, int stage) {
  extern __shared__ double sh[];
  static const double W[3][3] = {
      {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};
  for (int item = blockIdx.x; item < total; item += gridDim.x) {
    int m = find_member(mem, nm, item, kListGroups2);
    ColArgs c = mem[m].c;
    c.stage = stage;
    c.w0 = W[stage][0], c.w1 = W[stage][1], c.w2 = W[stage][2];
    int col = (item - mem[m].grp2) * (blockDim.x >> c.lshift) + (threadIdx.x >> c.lshift);
    int64_t o = mem[m].off;
    column_pcr2<3>(u + o, u0 + o, du + o, c, col, sh);
    __syncthreads();  // LDS is reused by the next item
  }
}

// ----------------------------------------------------------- host side ----

struct Tile {
  int tx1, tx2, tx3;
};

// Tile shape: x1 extent as long as possible (coalesced loads, halo ratio),
// then enough tiles to fill the 256 CUs; the LDS footprint stays below 64 KB
// so that two workgroups fit one CU.
static Tile choose_tile(const Grid& g) {
  Tile t;
  if (g.n3 > 1) {
    t = {16, 4, 4};
  } else if (g.n2 > 1) {
    t = {32, 8, 1};
    int64_t ntile = (int64_t)((g.nx1 + 31) / 32) * ((g.nx2 + 7) / 8);
    if (ntile < 512) t = {16, 4, 1};
  } else {
    t = {256, 1, 1};
  }
  t.tx1 = std::min(t.tx1, g.nx1);
  t.tx2 = g.n2 > 1 ? std::min(t.tx2, g.nx2) : 1;
  t.tx3 = g.n3 > 1 ? std::min(t.tx3, g.nx3) : 1;
  return t;
}

static size_t rhs_face_count(const FusedArgs& a, bool all) {
  size_t f1 = (size_t)(a.tx1 + 1) * a.tx2 * a.tx3;
  size_t f2 = a.h2 ? (size_t)a.tx1 * (a.tx2 + 1) * a.tx3 : 0;
  size_t f3 = a.h3 ? (size_t)a.tx1 * a.tx2 * (a.tx3 + 1) : 0;
  return all ? f1 + f2 + f3 : std::max(f1, std::max(f2, f3));
}

static size_t rhs_smem_bytes(const FusedArgs& a) {
  size_t H = (size_t)(a.tx1 + 6) * (a.tx2 + 2 * a.h2) * (a.tx3 + 2 * a.h3);
  // curvilinear tiles keep the x1 face pressures after the fluxes
  size_t fp = a.geo.coord != SNAPX_COORD_CARTESIAN ? (size_t)(a.tx1 + 1) * a.tx2 * a.tx3 : 0;
  // the perturbation path keeps the total density as a sixth plane
  size_t bgp = a.bgt ? H : 0;
  return sizeof(float) * (5 * (H + rhs_face_count(a, a.merged)) + fp + bgp);
}

// The option of a block that the fused path does not take, or null.
static const char* fused_unsupported(const snapx_config* c) {
  if (c->coord != SNAPX_COORD_CARTESIAN) {
    // the curvilinear K2 is the cell-parallel column kernel only
    const int N = ((c->implicit_scheme >> 3) & 1) ? 5 : 3;
    if (c->nx1 > (N == 3 ? PcrShape<3>::kThreads : PcrShape<5>::kThreads))
      return "fused path: curvilinear columns of at most 256 (3x3) / 128 (5x5) cells";
  }
  if (c->nghost != 3) return "fused path: nghost must be 3";
  // the reference advances custom ghost cells with the stage combine and the
  // body forces; the fused kernels never update ghost cells within a step
  for (int b = 0; b < 6; ++b)
    if (c->bc[b] == SNAPX_BC_CUSTOM)
      return "fused path: custom boundaries are not supported (use the plain path)";
  if (c->nx1 <= 1) return "fused path: x1 must be active";
  if (c->limiter) return "fused path: the EOS limiter is not supported (use the plain path)";
  if (c->non_hydrostatic != 1.0)
    return "fused path: non-hydrostatic < 1 is not supported (use the plain path)";
  for (int d = 0; d < 3; ++d)
    if (c->disable_flux[d]) return "fused path: disable-flux is not supported";
  return nullptr;
}

static int check_fused(const snapx_config* c) {
  if (int e = validate_config(c)) return e;
  if (const char* m = fused_unsupported(c)) {
    set_error(m);
    return SNAPX_EINVAL;
  }
  return 0;
}

static size_t col_count(const snapx_config* c) {
  size_t n2 = c->nx2 > 1 ? c->nx2 : 1, n3 = c->nx3 > 1 ? c->nx3 : 1;
  return n2 * n3;
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

static size_t scratch_bytes(const snapx_config* c) {
  int N = ((c->implicit_scheme >> 3) & 1) ? 5 : 3;
  return c->implicit_scheme ? sizeof(double) * (size_t)(N * N + N) * c->nx1 * col_count(c) : 0;
}

static FusedArgs fill_fused_args(const snapx_config* cfg, Tile t, double dt) {
  Grid g = make_grid(cfg);
  FusedArgs a;
  t.tx1 = std::min(t.tx1, g.nx1);
  t.tx2 = g.n2 > 1 ? std::min(t.tx2, g.nx2) : 1;
  t.tx3 = g.n3 > 1 ? std::min(t.tx3, g.nx3) : 1;
  a.n1 = g.n1, a.n2 = g.n2, a.n3 = g.n3, a.ng = g.ng;
  a.nx1 = cfg->nx1, a.nx2 = cfg->nx2, a.nx3 = cfg->nx3;
  a.tx1 = t.tx1, a.tx2 = t.tx2, a.tx3 = t.tx3;
  a.h2 = g.n2 > 1 ? 3 : 0;
  a.h3 = g.n3 > 1 ? 3 : 0;
  a.nt1 = (a.nx1 + a.tx1 - 1) / a.tx1;
  a.nt2 = g.n2 > 1 ? (a.nx2 + a.tx2 - 1) / a.tx2 : 1;
  a.nt3 = g.n3 > 1 ? (a.nx3 + a.tx3 - 1) / a.tx3 : 1;
  for (int r = 0; r < 2; ++r) {
    a.recon_type[r] = cfg->recon_type[r];
    a.recon_scale[r] = cfg->recon_scale[r];
    a.recon_shock[r] = cfg->recon_shock[r];
  }
  for (int d = 0; d < 3; ++d) a.riemann[d] = cfg->riemann[d];
  for (int b = 0; b < 6; ++b) a.bc[b] = cfg->bc[b];
  a.gammad = (float)cfg->gammad;
  a.dx1 = (float)g.dx1, a.dx2 = (float)g.dx2, a.dx3 = (float)g.dx3;
  a.dt = (float)dt;
  a.grav1 = (float)cfg->grav1, a.grav2 = (float)cfg->grav2, a.grav3 = (float)cfg->grav3;
  a.geo = make_geometry(cfg);
  a.bgt = nullptr;
  a.bgd = nullptr;
  a.dtp = nullptr;
  const bool sph = cfg->coord == SNAPX_COORD_SPHERICAL_POLAR;
  // ghost-layer distances of the refill: dx1v of the layers crossed
  double zi = 0., zo = 0.;
  for (int n = 0; n < 3; ++n) {
    zi += sph ? sph_dx1v(a.geo, g.ng - n - 1) : g.dx1;
    zo += sph ? sph_dx1v(a.geo, g.n1 - g.ng - 1 + n) : g.dx1;
    a.zin[n] = (float)zi, a.zout[n] = (float)zo;
  }
  // one pass over all faces when they number at most two per thread;
  // curvilinear tiles always take it (their sources read every face)
  a.merged = cfg->coord != SNAPX_COORD_CARTESIAN ||
             rhs_face_count(a, true) <= 2 * (size_t)kRhsThreads;
  return a;
}

static int make_fused_args(const snapx_config* cfg, Tile t, double dt, FusedArgs* out) {
  FusedArgs a = fill_fused_args(cfg, t, dt);
  SNAPX_CHECK_ARG(a.tx1 >= 1 && a.tx2 >= 1 && a.tx3 >= 1, "fused path: empty tile");
  SNAPX_CHECK_ARG((int64_t)a.tx1 * a.tx2 * a.tx3 <= kRhsThreads * kMaxCellsPerThread,
                  "fused path: tile too large");
  SNAPX_CHECK_ARG(rhs_smem_bytes(a) <= 64 * 1024, "fused path: tile needs more than 64 KB LDS");
  *out = a;
  return 0;
}

// 2.5-D K1 for 3-D blocks (fused_rhs_march_kernel): x1-x2 tiles of at most
// kRhsThreads cells (32 x 8 where the block allows) marching through slabs of
// 16, 8 or 4 slices along x3: the deepest slab that gives at least
// kMarchGroups (tile, slab) workgroups, else 4 slices.  Blocks with fewer
// than kMarchMinGroups workgroups at 4 slices take the 16 x 4 x 2 tiles.
// K1 per launch (us; slab 16 / 8 / 4 / tiled) on blocks nx1 x nx2 x nx3:
// 128x256x256 470 / 478 / - / -, 64x256x256 237 / 243 / 254 / -,
// 128x128x128 172 / 128 / 133 / 171, 96x128x128 168 / 109 / 116 / 238,
// 64x128x128 152 / 94 / 72 / 177, 128x96x96 - / 104 / 100 / -,
// 256x64x64 - / 94 / 71 / -, 128x64x64 151 / 82 / 53 / 95,
// 64x64x64 150 / 80 / 46 / 52, 32x64x64 152 / 78 / 43 / 33.
constexpr int kMarchSlabs[3] = {16, 8, 4};
constexpr int kMarchGroups = 768;
constexpr int kMarchMinGroups = 256;

static bool march_ok(const snapx_config* cfg) {
  return cfg->nx3 > 1 && cfg->nx2 > 1 && cfg->coord == SNAPX_COORD_CARTESIAN;
}

static size_t march_smem_bytes(const FusedArgs& a) {
  size_t HS = (size_t)(a.tx1 + 6) * (a.tx2 + 6);
  size_t ftot = (size_t)(a.tx1 + 1) * a.tx2 + (size_t)a.tx1 * (a.tx2 + 1);
  // perturbation path: a sixth plane per slice and the tile's background
  // tables (5 per level, 10 per x1 face)
  const size_t bgt = a.bgt ? HS + kBgLevelRows * (size_t)(a.tx1 + 6) +
                                 kBgFaceRows * (size_t)(a.tx1 + 1)
                           : 0;
  return sizeof(float) * (5 * HS + 5 * ftot + bgt);
}

static int make_march_args(const snapx_config* cfg, double dt, FusedArgs* out) {
  Tile t;
  t.tx1 = std::min(32, cfg->nx1);
  t.tx2 = std::min(cfg->nx2, kRhsThreads / t.tx1);
  // the deepest slab with kMarchGroups workgroups, else the shallowest
  const int ntile = ((cfg->nx1 + t.tx1 - 1) / t.tx1) * ((cfg->nx2 + t.tx2 - 1) / t.tx2);
  int slab = kMarchSlabs[2];
  for (int k = 0; k < 3; ++k)
    if ((int64_t)ntile * ((cfg->nx3 + kMarchSlabs[k] - 1) / kMarchSlabs[k]) >= kMarchGroups) {
      slab = kMarchSlabs[k];
      break;
    }
  t.tx3 = std::min(slab, cfg->nx3);
  FusedArgs a = fill_fused_args(cfg, t, dt);
  a.merged = 0;
  SNAPX_CHECK_ARG(a.tx1 >= 1 && a.tx2 >= 1 && a.tx3 >= 1 && a.tx1 * a.tx2 <= kRhsThreads,
                  "fused path: bad x1-x2 tile");
  // the kernel keeps the halo'd x1 positions of its plane cells in 8 bits
  SNAPX_CHECK_ARG(a.tx1 + 6 <= 255, "fused path: x1 tile too long for the marching kernel");
  SNAPX_CHECK_ARG(march_smem_bytes(a) <= 64 * 1024, "fused path: tile needs more than 64 KB LDS");
  *out = a;
  return 0;
}

using MarchKernel = void (*)(const float*, float*, FusedArgs);

// Compiled instances of the marching K1 (WENO5 density with CP5 velocities
// and pressure, shock off, scale off, on both axis groups): LMARS in x1 with
// Roe in x2 / x3, and one solver in every direction.  Other options run the
// generic instance.
static MarchKernel march_kernel_for(const snapx_config* c, bool bg) {
  const int pol = face_policy(c);
  if (bg)
    // perturbation path: the 3-D workload's solvers, else the generic instance
    return pol == kFacesLmRoRo ? fused_rhs_march_kernel<FacesLmRoRo, true>
                               : fused_rhs_march_kernel<MarchGeneric, true>;
  switch (pol) {
    case kFacesLmRoRo: return fused_rhs_march_kernel<FacesLmRoRo>;
    case kFacesLm: return fused_rhs_march_kernel<FacesLm>;
    case kFacesRo: return fused_rhs_march_kernel<FacesRo>;
    case kFacesHl: return fused_rhs_march_kernel<FacesHl>;
    default: return fused_rhs_march_kernel<MarchGeneric>;
  }
}

// The tiled K1 (merged or per-direction face passes, or the perturbation
// path's merged tiles) for a block's face options.
template <int kMerged>
static MarchKernel rhs_kernel_for(int pol) {
  switch (pol) {
    case kFacesLmRoRo: return fused_rhs_kernel<kMerged, FacesLmRoRo>;
    case kFacesLm: return fused_rhs_kernel<kMerged, FacesLm>;
    case kFacesRo: return fused_rhs_kernel<kMerged, FacesRo>;
    case kFacesHl: return fused_rhs_kernel<kMerged, FacesHl>;
    default: return fused_rhs_kernel<kMerged>;
  }
}
static MarchKernel rhs_bg_kernel_for(int pol) {
  switch (pol) {
    case kFacesLmRoRo: return fused_rhs_bg_kernel<FacesLmRoRo>;
    case kFacesLm: return fused_rhs_bg_kernel<FacesLm>;
    default: return fused_rhs_bg_kernel<>;
  }
}

static int make_col_args(const snapx_config* cfg, double dt, ColArgs* out) {
  Grid g = make_grid(cfg);
  ColArgs c;
  c.n1 = g.n1, c.n2 = g.n2, c.n3 = g.n3, c.ng = g.ng;
  c.nl = cfg->nx1;
  c.ncol2 = g.n2 > 1 ? cfg->nx2 : 1;
  c.ncol = (int)col_count(cfg);
  c.bc_in = cfg->bc[0], c.bc_out = cfg->bc[1];
  c.scheme = cfg->implicit_scheme;
  c.stage = 0;
  c.gammad = cfg->gammad;
  c.dx1 = g.dx1;
  c.area = g.dx3 * g.dx2;
  c.vol = g.dx3 * (g.dx2 * g.dx1);
  c.dt = dt;
  c.grav1 = cfg->grav1;
  c.w0 = 0., c.w1 = 1., c.w2 = 1.;
  c.lanes = 64, c.lshift = 6;
  while (c.lanes < c.nl) c.lanes <<= 1, c.lshift += 1;
  c.geo = make_geometry(cfg);
  c.bgd = nullptr;
  c.bgt = nullptr;
  c.dtp = nullptr;
  const bool sph = cfg->coord == SNAPX_COORD_SPHERICAL_POLAR;
  c.zin1 = sph ? sph_dx1v(c.geo, g.ng - 1) : g.dx1;
  c.zout1 = sph ? sph_dx1v(c.geo, g.n1 - g.ng - 1) : g.dx1;
  // 3x3 columns of more than 32 cells: two rows per lane (column_pcr2;
  // Cartesian blocks), 32 lanes or more per column
  c.rows2 = (cfg->coord == SNAPX_COORD_CARTESIAN && !((cfg->implicit_scheme >> 3) & 1) &&
             c.nl > 32) ? 1 : 0;
  if (c.rows2) c.lanes >>= 1, c.lshift -= 1;
  SNAPX_CHECK_ARG(!c.scheme || c.grav1 != 0.,
                  "[ImplicitHydro] forcing does not have const-gravity");
  *out = c;
  return 0;
}

// The cell-parallel column kernel holds whole columns in one workgroup:
// usable when the lanes per column fit the workgroup of that block size.
static bool pcr_ok(const ColArgs& c, int N) {
  return c.lanes <= (N == 3 ? PcrShape<3>::kThreads : PcrShape<5>::kThreads);
}
static int pcr_threads(int N) { return N == 3 ? PcrShape<3>::kThreads : PcrShape<5>::kThreads; }
static size_t pcr_smem(int N) {
  return sizeof(double) * (size_t)pcr_threads(N) *
         (N == 3 ? PcrShape<3>::kPub : PcrShape<5>::kPub);
}
static int pcr_groups(const ColArgs& c, int N) {
  int cpw = pcr_threads(N) / c.lanes;
  return (c.ncol + cpw - 1) / cpw;
}
// Workgroup size of the single-block column kernel: whole columns, at most
// pcr_threads(N) lanes, and no more columns per workgroup than leaves at least
// 512 workgroups (two per CU), so that a block of few columns spreads over
// every CU and each CU's LDS serves fewer wavefronts.
static int pcr_block(const ColArgs& c, int N) {
  int cpw = pcr_threads(N) / c.lanes;
  while (cpw > 1 && (c.ncol + cpw - 1) / cpw < 512) cpw >>= 1;
  return cpw * c.lanes;
}

// K1 of one block: the kernel (marching tiles for 3-D blocks with enough
// (tile, slab) workgroups, curvilinear, perturbation-path or 1-D / 2-D /
// small 3-D tiles), its arguments and launch shape.
struct RhsPlan {
  FusedArgs a;
  bool march, sph, gno;
  const snapx_config* cfg;
  int nblk;
  size_t smem;
};

// bg: device tables of the hydrostatic background (null: total state); dtp:
// the time step in device memory (replaces dt)
static int rhs_plan(const snapx_config* cfg, double dt, const void* bg, const double* dtp,
                    RhsPlan* p) {
  Grid g = make_grid(cfg);
  bool march = march_ok(cfg);
  const bool sph = cfg->coord != SNAPX_COORD_CARTESIAN;  // curvilinear kernels
  const bool gno = cfg->coord == SNAPX_COORD_GNOMONIC_EQUIANGLE;
  FusedArgs a;
  if (march) {
    if (int e = make_march_args(cfg, dt, &a)) return e;
    // the marching kernel needs enough (tile, slab) workgroups to fill the
    // GPU; below kMarchMinGroups the merged 16x4x2 tiles win
    if (a.nt1 * a.nt2 * a.nt3 < kMarchMinGroups) march = false;
  }
  if (!march) {
    Tile t = choose_tile(g);
    // 3-D tiles: 16x4x2, every face of the tile in one LDS pass
    if (g.n3 > 1) t = {16, 4, 2};
    // the perturbation path runs merged tiles only
    if (bg && g.n3 == 1 && g.n2 > 1) t = {16, 4, 1};
    if (int e = make_fused_args(cfg, t, dt, &a)) return e;
  }
  if (bg) {
    a.bgd = reinterpret_cast<const double*>(bg);
    a.bgt = reinterpret_cast<const float*>(a.bgd + 2 * g.n1);
    SNAPX_CHECK_ARG(march || a.merged, "perturbation path: tile without a merged face pass");
  }
  a.dtp = dtp;
  size_t smem = march ? march_smem_bytes(a) : rhs_smem_bytes(a);
  SNAPX_CHECK_ARG(smem <= 64 * 1024, "fused path: tile needs more than 64 KB LDS");
  p->a = a;
  p->march = march;
  p->sph = sph;
  p->gno = gno;
  p->cfg = cfg;
  p->nblk = a.nt1 * a.nt2 * a.nt3;
  p->smem = smem;
  return 0;
}

// du <- the explicit tendency of u (dt times the flux divergence and the
// sources, interior cells); ra / rb: optional events stamped with the start
// and end of the dispatch
static void rhs_launch(const RhsPlan& p, const float* u, float* du, hipStream_t st,
                       hipEvent_t ra = nullptr, hipEvent_t rb = nullptr) {
  const uint32_t smem = (uint32_t)p.smem;
  const dim3 grid(p.nblk), block(kRhsThreads);
  if (p.sph)
    hipExtLaunchKernelGGL(p.gno ? fused_rhs_gno_kernel : fused_rhs_sph_kernel, grid, block, smem,
                          st, ra, rb, 0u, u, du, p.a);
  else if (p.march)
    hipExtLaunchKernelGGL(march_kernel_for(p.cfg, p.a.bgt != nullptr), grid, block, smem, st, ra,
                          rb, 0u, u, du, p.a);
  else if (p.a.bgt)
    hipExtLaunchKernelGGL(rhs_bg_kernel_for(face_policy(p.cfg)), grid, block, smem, st, ra, rb, 0u,
                          u, du, p.a);
  else if (p.a.merged)
    hipExtLaunchKernelGGL(rhs_kernel_for<1>(face_policy(p.cfg)), grid, block, smem, st, ra, rb, 0u,
                          u, du, p.a);
  else
    hipExtLaunchKernelGGL(rhs_kernel_for<0>(face_policy(p.cfg)), grid, block, smem, st, ra, rb, 0u,
                          u, du, p.a);
}

bool fused_rhs_ok(const snapx_config* c) { return fused_unsupported(c) == nullptr; }

int fused_explicit_tendency(const snapx_config* c, double dt, const float* u, float* du,
                            hipStream_t st) {
  RhsPlan rp;
  if (int e = rhs_plan(c, dt, nullptr, nullptr, &rp)) return e;
  rhs_launch(rp, u, du, st);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

struct EventPair {
  hipEvent_t a, b;
};

// One fused RK3 step; with ev != nullptr the 7 launches are bracketed by the
// event pairs ev[0..6] (K1, K2 per stage, then the boundary kernel).
// bg: device tables of the hydrostatic background (bg_impl.hpp); u is then
// the perturbation state and the stored ghosts are not refreshed.  dtp: the
// time step in device memory (replaces dt).
static int fused_step(const snapx_config* cfg, double dt, float* u, void* workspace,
                      hipStream_t st, EventPair* ev, bool bc = true, const void* bg = nullptr,
                      const double* dtp = nullptr) {
  Grid g = make_grid(cfg);
  size_t field = align256(sizeof(float) * 5 * (size_t)g.ncell);
  char* base = reinterpret_cast<char*>(workspace);
  float* du = reinterpret_cast<float*>(base);
  float* u0 = reinterpret_cast<float*>(base + field);
  double* scr = reinterpret_cast<double*>(base + 2 * field);

  RhsPlan rp;
  if (int e = rhs_plan(cfg, dt, bg, dtp, &rp)) return e;
  const bool sph = rp.sph, gno = rp.gno;
  ColArgs c;
  if (int e = make_col_args(cfg, dt, &c)) return e;
  c.bgd = rp.a.bgd;
  c.bgt = rp.a.bgt;
  c.dtp = dtp;
  int N = ((cfg->implicit_scheme >> 3) & 1) ? 5 : 3;
  int cblk = (c.ncol + kColThreads - 1) / kColThreads;

  static const double W[3][3] = {
      {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};
  static const char* const kRhsRange[3] = {"snapx fused K1 rhs stage 0",
                                           "snapx fused K1 rhs stage 1",
                                           "snapx fused K1 rhs stage 2"};
  static const char* const kColRange[3] = {"snapx fused K2 column stage 0",
                                           "snapx fused K2 column stage 1",
                                           "snapx fused K2 column stage 2"};
  TraceRange step_range(bg ? "snapx fused step (perturbation)" : "snapx fused step");
  // with a timing record, the events carry the start and end time stamps of
  // the kernel dispatch itself (hipExtLaunchKernel), not extra stream markers
  for (int s = 0; s < 3; ++s) {
    hipEvent_t ra = ev ? ev[2 * s].a : nullptr, rb = ev ? ev[2 * s].b : nullptr;
    roctxRangePushA(kRhsRange[s]);
    rhs_launch(rp, (const float*)u, du, st, ra, rb);
    roctxRangePop();
    SNAPX_LAUNCH_CHECK();
    TraceRange col_range(kColRange[s]);
    c.stage = s;
    c.w0 = W[s][0], c.w1 = W[s][1], c.w2 = W[s][2];
    hipEvent_t ca = ev ? ev[2 * s + 1].a : nullptr, cb = ev ? ev[2 * s + 1].b : nullptr;
    if (pcr_ok(c, N)) {
      const int threads = pcr_block(c, N), cpw = threads / c.lanes;
      dim3 grid((c.ncol + cpw - 1) / cpw), block(threads);
      const uint32_t smem = (uint32_t)(sizeof(double) * threads *
                                       (N == 3 ? PcrShape<3>::kPub : PcrShape<5>::kPub));
      if (sph && N == 3)
        hipExtLaunchKernelGGL(gno ? fused_column_pcr_gno_kernel<3> : fused_column_pcr_sph_kernel<3>,
                              grid, block, smem, st, ca, cb, 0u, u, u0, (const float*)du, c);
      else if (sph)
        hipExtLaunchKernelGGL(gno ? fused_column_pcr_gno_kernel<5> : fused_column_pcr_sph_kernel<5>,
                              grid, block, smem, st, ca, cb, 0u, u, u0, (const float*)du, c);
      else if (bg && N == 3 && c.rows2)
        hipExtLaunchKernelGGL(fused_column_pcr2_kernel<true>, grid, block, smem, st, ca, cb, 0u, u,
                              u0, (const float*)du, c);
      else if (bg && N == 3)
        hipExtLaunchKernelGGL((fused_column_pcr_kernel<3, true>), grid, block, smem, st, ca, cb, 0u,
                              u, u0, (const float*)du, c);
      else if (bg)
        hipExtLaunchKernelGGL((fused_column_pcr_kernel<5, true>), grid, block, smem, st, ca, cb, 0u,
                              u, u0, (const float*)du, c);
      else if (N == 3 && c.rows2)
        hipExtLaunchKernelGGL(fused_column_pcr2_kernel<false>, grid, block, smem, st, ca, cb, 0u, u,
                              u0, (const float*)du, c);
      else if (N == 3)
        hipExtLaunchKernelGGL(fused_column_pcr_kernel<3>, grid, block, smem, st, ca, cb, 0u, u,
                              u0, (const float*)du, c);
      else
        hipExtLaunchKernelGGL(fused_column_pcr_kernel<5>, grid, block, smem, st, ca, cb, 0u, u,
                              u0, (const float*)du, c);
    } else if (N == 3) {
      hipExtLaunchKernelGGL(fused_column_kernel<3>, dim3(cblk), dim3(kColThreads), 0u, st, ca, cb,
                            0u, u, u0, (const float*)du, scr, c);
    } else {
      hipExtLaunchKernelGGL(fused_column_kernel<5>, dim3(cblk), dim3(kColThreads), 0u, st, ca, cb,
                            0u, u, u0, (const float*)du, scr, c);
    }
    SNAPX_LAUNCH_CHECK();
  }
  if (ev) SNAPX_HIP(hipEventRecord(ev[6].a, st));
  if (bc) {
    TraceRange bc_range("snapx fused boundary");
    if (int e = launch_bc_f32(cfg, u, st)) return e;
  }
  if (ev) SNAPX_HIP(hipEventRecord(ev[6].b, st));
  return 0;
}

}  // namespace snapx

using namespace snapx;

extern "C" size_t snapx_fused_workspace_bytes(const snapx_config* cfg) {
  if (check_fused(cfg)) return 0;
  Grid g = make_grid(cfg);
  size_t field = align256(sizeof(float) * 5 * (size_t)g.ncell);
  return 2 * field + align256(scratch_bytes(cfg));
}

extern "C" int snapx_fused_rk3_step(const snapx_config* cfg, double dt, float* u, void* workspace,
                                    size_t workspace_bytes, void* stream) {
  if (int e = check_fused(cfg)) return e;
  SNAPX_CHECK_ARG(u && workspace, "snapx_fused_rk3_step: null pointer");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_fused_workspace_bytes(cfg),
                  "snapx_fused_rk3_step: workspace too small");
  return fused_step(cfg, dt, u, workspace, as_stream(stream), nullptr);
}

extern "C" int snapx_fused_rk3_step_dt(const snapx_config* cfg, const double* dt, float* u,
                                       void* workspace, size_t workspace_bytes, void* stream) {
  if (int e = check_fused(cfg)) return e;
  SNAPX_CHECK_ARG(dt && u && workspace, "snapx_fused_rk3_step_dt: null pointer");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_fused_workspace_bytes(cfg),
                  "snapx_fused_rk3_step_dt: workspace too small");
  return fused_step(cfg, 0., u, workspace, as_stream(stream), nullptr, true, nullptr, dt);
}

extern "C" int snapx_fused_rk3_steps(const snapx_config* cfg, double dt, int nsteps, float* u,
                                     void* workspace, size_t workspace_bytes,
                                     snapx_timing* timing, void* stream) {
  if (int e = check_fused(cfg)) return e;
  SNAPX_CHECK_ARG(u && workspace && nsteps >= 0, "snapx_fused_rk3_steps: bad arguments");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_fused_workspace_bytes(cfg),
                  "snapx_fused_rk3_steps: workspace too small");
  hipStream_t st = as_stream(stream);
  if (!timing) {
    // the fused kernels never read ghost cells, so the stored ghosts are
    // refreshed once, after the last step
    for (int n = 0; n < nsteps; ++n)
      if (int e = fused_step(cfg, dt, u, workspace, st, nullptr, n == nsteps - 1)) return e;
    return 0;
  }
  EventPair ev[7];
  for (auto& e : ev) {
    SNAPX_HIP(hipEventCreate(&e.a));
    SNAPX_HIP(hipEventCreate(&e.b));
  }
  int status = 0;
  for (int n = 0; n < nsteps && !status; ++n) {
    status = fused_step(cfg, dt, u, workspace, st, ev);
    if (status) break;
    if (hipEventSynchronize(ev[6].b) != hipSuccess) {
      set_error("snapx_fused_rk3_steps: event synchronisation failed");
      status = SNAPX_EHIP;
      break;
    }
    for (int k = 0; k < 7; ++k) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, ev[k].a, ev[k].b);
      if (k == 6) {
        timing->bc_ms += ms;
        timing->bc_launches += 1;
      } else if (k % 2 == 0) {
        timing->rhs_ms += ms;
        timing->rhs_launches += 1;
      } else {
        timing->column_ms += ms;
        timing->column_launches += 1;
      }
    }
  }
  for (auto& e : ev) {
    hipEventDestroy(e.a);
    hipEventDestroy(e.b);
  }
  return status;
}

// ------------------------------ perturbation path (hydrostatic background) ----

namespace snapx {

static int check_bg(const snapx_config* c) {
  if (int e = check_fused(c)) return e;
  SNAPX_CHECK_ARG(c->coord == SNAPX_COORD_CARTESIAN, "perturbation path: Cartesian blocks only");
  SNAPX_CHECK_ARG(c->recon_type[0] == SNAPX_INTERP_WENO5 && !c->recon_shock[0],
                  "perturbation path: vertical reconstruction must be weno5 with shock: false");
  SNAPX_CHECK_ARG(!c->recon_scale[0] && !c->recon_scale[1],
                  "perturbation path: reconstruction scale must be off");
  SNAPX_CHECK_ARG(c->grav2 == 0.0 && c->grav3 == 0.0,
                  "perturbation path: the background varies with x1 only (grav2 = grav3 = 0)");
  ColArgs ca;
  if (int e = make_col_args(c, 1.0, &ca)) return e;
  const int N = ((c->implicit_scheme >> 3) & 1) ? 5 : 3;
  SNAPX_CHECK_ARG(pcr_ok(ca, N),
                  "perturbation path: columns of at most 512 (3x3) / 128 (5x5) cells");
  return 0;
}

template <typename T>
__global__ void bg_split_kernel(const T* __restrict__ u, const double* __restrict__ bgd,
                                float* __restrict__ up, int n1, int ncell, double igm1) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    const int l = q % n1;
    up[IDN * ncell + q] = (float)((double)u[IDN * ncell + q] - bgd[l]);
    up[IVX * ncell + q] = (float)u[IVX * ncell + q];
    up[IVY * ncell + q] = (float)u[IVY * ncell + q];
    up[IVZ * ncell + q] = (float)u[IVZ * ncell + q];
    up[IPR * ncell + q] = (float)((double)u[IPR * ncell + q] - bgd[n1 + l] * igm1);
  }
}

template <typename T>
__global__ void bg_join_kernel(const float* __restrict__ up, const double* __restrict__ bgd,
                               T* __restrict__ u, int n1, int ncell, double igm1) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < ncell; q += gridDim.x * blockDim.x) {
    const int l = q % n1;
    u[IDN * ncell + q] = (T)(bgd[l] + (double)up[IDN * ncell + q]);
    u[IVX * ncell + q] = (T)up[IVX * ncell + q];
    u[IVY * ncell + q] = (T)up[IVY * ncell + q];
    u[IVZ * ncell + q] = (T)up[IVZ * ncell + q];
    u[IPR * ncell + q] = (T)(bgd[n1 + l] * igm1 + (double)up[IPR * ncell + q]);
  }
}

}  // namespace snapx

extern "C" size_t snapx_fused_background_bytes(const snapx_config* cfg) {
  if (check_bg(cfg)) return 0;
  return bg_bytes(make_grid(cfg).n1);
}

extern "C" int snapx_fused_background(const snapx_config* cfg, const double* rho_bar,
                                      const double* p_bar, void* bg, size_t bg_bytes_,
                                      void* stream) {
  if (int e = check_bg(cfg)) return e;
  SNAPX_CHECK_ARG(rho_bar && p_bar && bg, "snapx_fused_background: null pointer");
  const int n1 = make_grid(cfg).n1;
  SNAPX_CHECK_ARG(bg_bytes_ >= bg_bytes(n1), "snapx_fused_background: buffer too small");
  for (int l = 0; l < n1; ++l)
    SNAPX_CHECK_ARG(rho_bar[l] > 0. && p_bar[l] > 0. && std::isfinite(rho_bar[l]) &&
                        std::isfinite(p_bar[l]),
                    "snapx_fused_background: background density and pressure must be positive");
  std::vector<double> d64;
  std::vector<float> f32;
  bg_build_tables(cfg, rho_bar, p_bar, d64, f32);
  hipStream_t st = as_stream(stream);
  char* dst = reinterpret_cast<char*>(bg);
  SNAPX_HIP(hipMemcpyAsync(dst, d64.data(), sizeof(double) * d64.size(), hipMemcpyHostToDevice, st));
  SNAPX_HIP(hipMemcpyAsync(dst + sizeof(double) * d64.size(), f32.data(),
                           sizeof(float) * f32.size(), hipMemcpyHostToDevice, st));
  // the host tables go out of scope here
  SNAPX_HIP(hipStreamSynchronize(st));
  return 0;
}

extern "C" int snapx_fused_split(const snapx_config* cfg, const void* bg, const void* u, int dtype,
                                 float* u_pert, void* stream) {
  if (int e = check_bg(cfg)) return e;
  SNAPX_CHECK_ARG(bg && u && u_pert, "snapx_fused_split: null pointer");
  SNAPX_CHECK_ARG(dtype == SNAPX_F32 || dtype == SNAPX_F64, "snapx_fused_split: dtype");
  Grid g = make_grid(cfg);
  const double igm1 = 1. / (cfg->gammad - 1.);
  const double* bgd = reinterpret_cast<const double*>(bg);
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(bg_split_kernel<double>, blocks_for(g.ncell, 256), dim3(256), 0, st,
                       reinterpret_cast<const double*>(u), bgd, u_pert, g.n1, g.ncell, igm1);
  else
    hipLaunchKernelGGL(bg_split_kernel<float>, blocks_for(g.ncell, 256), dim3(256), 0, st,
                       reinterpret_cast<const float*>(u), bgd, u_pert, g.n1, g.ncell, igm1);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_fused_join(const snapx_config* cfg, const void* bg, const float* u_pert,
                                void* u, int dtype, void* stream) {
  if (int e = check_bg(cfg)) return e;
  SNAPX_CHECK_ARG(bg && u && u_pert, "snapx_fused_join: null pointer");
  SNAPX_CHECK_ARG(dtype == SNAPX_F32 || dtype == SNAPX_F64, "snapx_fused_join: dtype");
  Grid g = make_grid(cfg);
  const double igm1 = 1. / (cfg->gammad - 1.);
  const double* bgd = reinterpret_cast<const double*>(bg);
  hipStream_t st = as_stream(stream);
  if (dtype == SNAPX_F64)
    hipLaunchKernelGGL(bg_join_kernel<double>, blocks_for(g.ncell, 256), dim3(256), 0, st, u_pert,
                       bgd, reinterpret_cast<double*>(u), g.n1, g.ncell, igm1);
  else
    hipLaunchKernelGGL(bg_join_kernel<float>, blocks_for(g.ncell, 256), dim3(256), 0, st, u_pert,
                       bgd, reinterpret_cast<float*>(u), g.n1, g.ncell, igm1);
  SNAPX_LAUNCH_CHECK();
  // the stored ghosts of the perturbation are not maintained: the total
  // state gets the reference's boundary pass
  return snapx_apply_bc(cfg, u, 5, SNAPX_CONSERVED, dtype, stream);
}

extern "C" int snapx_fused_rk3_steps_bg(const snapx_config* cfg, const void* bg, double dt,
                                        int nsteps, float* u_pert, void* workspace,
                                        size_t workspace_bytes, snapx_timing* timing,
                                        void* stream) {
  if (int e = check_bg(cfg)) return e;
  SNAPX_CHECK_ARG(bg && u_pert && workspace && nsteps >= 0, "snapx_fused_rk3_steps_bg: bad arguments");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_fused_workspace_bytes(cfg),
                  "snapx_fused_rk3_steps_bg: workspace too small");
  hipStream_t st = as_stream(stream);
  if (!timing) {
    for (int n = 0; n < nsteps; ++n)
      if (int e = fused_step(cfg, dt, u_pert, workspace, st, nullptr, false, bg)) return e;
    return 0;
  }
  EventPair ev[7];
  for (auto& e : ev) {
    SNAPX_HIP(hipEventCreate(&e.a));
    SNAPX_HIP(hipEventCreate(&e.b));
  }
  int status = 0;
  for (int n = 0; n < nsteps && !status; ++n) {
    status = fused_step(cfg, dt, u_pert, workspace, st, ev, false, bg);
    if (status) break;
    if (hipEventSynchronize(ev[5].b) != hipSuccess) {
      set_error("snapx_fused_rk3_steps_bg: event synchronisation failed");
      status = SNAPX_EHIP;
      break;
    }
    for (int k = 0; k < 6; ++k) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[k].a, ev[k].b);
      if (k % 2 == 0) {
        timing->rhs_ms += ms;
        timing->rhs_launches += 1;
      } else {
        timing->column_ms += ms;
        timing->column_launches += 1;
      }
    }
  }
  for (auto& e : ev) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  return status;
}

// ------------------------------------- fused steps without allocation ----

namespace snapx {

// Workspace of snapx_fused_rk3_steps_static: one array per GPU, held from
// the first load of this library's code object on that GPU until the process
// ends.  Every call on a GPU shares it (the last call's contents stay); no
// part of it has to start at zero, because every kernel of a fused step
// writes each word it later reads within the same step (K1 writes the
// tendency of every interior cell, stage 0 of K2 writes the stage-0 state,
// the lane-per-column solve writes its column scratch in the forward sweep).
static __device__ __align__(256) char scratch[SNAPX_STATIC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxDevices];

// Device address of the current GPU's copy of `scratch` (the array's name in
// host code is only a stand-in for the symbol).  Looked up once per device.
static int static_scratch(void** out) {
  int dev = 0;
  SNAPX_HIP(hipGetDevice(&dev));
  SNAPX_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "snapx_fused_rk3_steps_static: device index");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SNAPX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SNAPX_CHECK_ARG(p != nullptr, "snapx_fused_rk3_steps_static: no device address for scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace snapx

extern "C" int snapx_fused_rk3_steps_static(const snapx_config* cfg, double dt, int nsteps,
                                            float* u, void* stream) {
  if (int e = check_fused(cfg)) return e;
  SNAPX_CHECK_ARG(u && nsteps >= 0, "snapx_fused_rk3_steps_static: bad arguments");
  const size_t need = snapx_fused_workspace_bytes(cfg);
  if (need > (size_t)SNAPX_STATIC_SCRATCH_BYTES) {
    set_error("snapx_fused_rk3_steps_static: the block needs " + std::to_string(need) +
              " workspace bytes, more than the static " +
              std::to_string((size_t)SNAPX_STATIC_SCRATCH_BYTES) +
              " (use snapx_fused_rk3_steps with a caller workspace)");
    return SNAPX_ESPACE;
  }
  void* ws = nullptr;
  if (int e = static_scratch(&ws)) return e;
  hipStream_t st = as_stream(stream);
  for (int n = 0; n < nsteps; ++n)
    if (int e = fused_step(cfg, dt, u, ws, st, nullptr, n == nsteps - 1)) return e;
  return 0;
}

// ---------------------------------------------------- batched C ABI ----

namespace snapx {

static size_t batch_table_bytes(int nm) { return align256(sizeof(BatchMember) * (size_t)nm); }

static int batch_layout(const snapx_config* members, int nm, double dt, int tx1, int tx2, int tx3,
                        std::vector<BatchMember>* table, snapx_batch_info* info) {
  SNAPX_CHECK_ARG(members && nm >= 1, "batch: no members");
  int scheme = members[0].implicit_scheme;
  const bool sph = members[0].coord != SNAPX_COORD_CARTESIAN;  // curvilinear
  bool three_d = false, two_d = false;
  for (int m = 0; m < nm; ++m) {
    if (int e = check_fused(&members[m])) return e;
    SNAPX_CHECK_ARG(members[m].coord == members[0].coord,
                    "batch: every member needs the same coordinate system");
    SNAPX_CHECK_ARG(members[m].implicit_scheme == scheme,
                    "batch: every member needs the same implicit scheme");
    three_d |= members[m].nx3 > 1;
    two_d |= members[m].nx2 > 1;
  }
  Tile def = three_d ? Tile{16, 4, 4} : (two_d ? Tile{16, 4, 1} : Tile{256, 1, 1});
  Tile t = {tx1 > 0 ? tx1 : def.tx1, tx2 > 0 ? tx2 : def.tx2, tx3 > 0 ? tx3 : def.tx3};
  table->resize(nm);
  int64_t off = 0, soff = 0;
  int tiles = 0, cols = 0, groups = 0, groups2 = 0;
  bool pcr = true;
  size_t smem = 0;
  for (int m = 0; m < nm; ++m) {
    BatchMember& b = (*table)[m];
    if (int e = make_fused_args(&members[m], t, dt, &b.a)) return e;
    if (int e = make_col_args(&members[m], dt, &b.c)) return e;
    smem = std::max(smem, rhs_smem_bytes(b.a));
    b.off = off;
    b.scr_off = soff;
    b.tile0 = tiles;
    b.col0 = cols;
    b.grp0 = groups;
    b.grp2 = groups2;
    Grid g = make_grid(&members[m]);
    off += 5 * (int64_t)g.ncell;
    soff += (int64_t)scratch_bytes(&members[m]) / sizeof(double);
    tiles += b.a.nt1 * b.a.nt2 * b.a.nt3;
    cols += (b.c.ncol + kColThreads - 1) / kColThreads * kColThreads;
    int N = ((scheme >> 3) & 1) ? 5 : 3;
    // spherical-polar columns have the cell-parallel kernel only (one row per lane)
    SNAPX_CHECK_ARG(!sph || (pcr_ok(b.c, N) && !b.c.rows2),
                    "batch: spherical-polar member column too tall for the column kernel");
    pcr &= pcr_ok(b.c, N);
    if (pcr_ok(b.c, N)) (b.c.rows2 ? groups2 : groups) += pcr_groups(b.c, N);
  }
  info->nmembers = nm;
  info->tile_x1 = t.tx1, info->tile_x2 = t.tx2, info->tile_x3 = t.tx3;
  info->total_tiles = tiles;
  info->total_cols = cols;
  info->col_groups = pcr ? groups : 0;
  info->col_groups2 = pcr ? groups2 : 0;
  info->total_elems = off;
  info->scratch_doubles = soff;
  info->rhs_smem = smem;
  info->vic_size = ((scheme >> 3) & 1) ? 5 : 3;
  info->dt = dt;
  return 0;
}

}  // namespace snapx

extern "C" size_t snapx_batch_workspace_bytes(const snapx_config* members, int nmembers) {
  std::vector<BatchMember> table;
  snapx_batch_info info;
  if (batch_layout(members, nmembers, 1.0, 0, 0, 0, &table, &info)) return 0;
  size_t field = align256(sizeof(float) * (size_t)info.total_elems);
  return batch_table_bytes(nmembers) + 2 * field + align256(sizeof(double) * info.scratch_doubles);
}

extern "C" int snapx_batch_plan(const snapx_config* members, int nmembers, double dt,
                                int tile_x1, int tile_x2, int tile_x3, void* workspace,
                                size_t workspace_bytes, snapx_batch_info* info) {
  SNAPX_CHECK_ARG(workspace && info, "snapx_batch_plan: null pointer");
  std::vector<BatchMember> table;
  if (int e = batch_layout(members, nmembers, dt, tile_x1, tile_x2, tile_x3, &table, info))
    return e;
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_batch_workspace_bytes(members, nmembers),
                  "snapx_batch_plan: workspace too small");
  SNAPX_HIP(hipMemcpy(workspace, table.data(), sizeof(BatchMember) * table.size(),
                      hipMemcpyHostToDevice));
  return 0;
}

extern "C" int snapx_batch_rk3_steps(const snapx_config* members, const snapx_batch_info* info,
                                     int nsteps, float* u, void* workspace,
                                     size_t workspace_bytes, snapx_timing* timing, void* stream) {
  SNAPX_CHECK_ARG(members && info && u && workspace && nsteps >= 0,
                  "snapx_batch_rk3_steps: bad arguments");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_batch_workspace_bytes(members, info->nmembers),
                  "snapx_batch_rk3_steps: workspace too small");
  hipStream_t st = as_stream(stream);
  int nm = info->nmembers;
  const bool sph = members[0].coord != SNAPX_COORD_CARTESIAN;  // curvilinear
  const bool gno = members[0].coord == SNAPX_COORD_GNOMONIC_EQUIANGLE;
  char* base = reinterpret_cast<char*>(workspace);
  const BatchMember* table = reinterpret_cast<const BatchMember*>(base);
  size_t field = align256(sizeof(float) * (size_t)info->total_elems);
  float* du = reinterpret_cast<float*>(base + batch_table_bytes(nm));
  float* u0 = reinterpret_cast<float*>(base + batch_table_bytes(nm) + field);
  double* scr = reinterpret_cast<double*>(base + batch_table_bytes(nm) + 2 * field);
  // persistent grids: two workgroups per CU for K1, enough wavefronts for K2
  int rhs_grid = std::min(info->total_tiles, 512);
  const BatchRhsKernel rhs_kernel = batch_rhs_for(members, nm);
  int col_grid = std::min((info->total_cols + kColThreads - 1) / kColThreads, 2048);
  EventPair ev[2];
  if (timing)
    for (auto& e : ev) {
      SNAPX_HIP(hipEventCreate(&e.a));
      SNAPX_HIP(hipEventCreate(&e.b));
    }
  for (int n = 0; n < nsteps; ++n) {
    TraceRange step_range("snapx batch step");
    for (int s = 0; s < 3; ++s) {
      if (timing) SNAPX_HIP(hipEventRecord(ev[0].a, st));
      hipLaunchKernelGGL(rhs_kernel, dim3(rhs_grid), dim3(kRhsThreads), info->rhs_smem, st, u, du,
                         table, nm, info->total_tiles);
      SNAPX_LAUNCH_CHECK();
      if (timing) {
        SNAPX_HIP(hipEventRecord(ev[0].b, st));
        SNAPX_HIP(hipEventRecord(ev[1].a, st));
      }
      if (info->col_groups > 0 || info->col_groups2 > 0) {
        int N = info->vic_size;
        if (info->col_groups > 0) {
          dim3 grid(std::min(info->col_groups, 4096)), block(pcr_threads(N));
          auto k3 = gno ? batch_column_pcr_kernel<3, 2>
                        : (sph ? batch_column_pcr_kernel<3, 1> : batch_column_pcr_kernel<3, 0>);
          auto k5 = gno ? batch_column_pcr_kernel<5, 2>
                        : (sph ? batch_column_pcr_kernel<5, 1> : batch_column_pcr_kernel<5, 0>);
          hipLaunchKernelGGL(N == 3 ? k3 : k5, grid, block, pcr_smem(N), st, u, u0, du, table, nm,
                             info->col_groups, s);
        }
        if (info->col_groups2 > 0)
          hipLaunchKernelGGL(batch_column_pcr2_kernel, dim3(std::min(info->col_groups2, 4096)),
                             dim3(pcr_threads(3)), pcr_smem(3), st, u, u0, du, table, nm,
                             info->col_groups2, s);
      } else if (info->vic_size == 3) {
        hipLaunchKernelGGL(batch_column_kernel<3>, dim3(col_grid), dim3(kColThreads), 0, st, u,
                           u0, du, scr, table, nm, info->total_cols, s);
      } else {
        hipLaunchKernelGGL(batch_column_kernel<5>, dim3(col_grid), dim3(kColThreads), 0, st, u,
                           u0, du, scr, table, nm, info->total_cols, s);
      }
      SNAPX_LAUNCH_CHECK();
      if (timing) {
        SNAPX_HIP(hipEventRecord(ev[1].b, st));
        SNAPX_HIP(hipEventSynchronize(ev[1].b));
        float a = 0.f, b = 0.f;
        hipEventElapsedTime(&a, ev[0].a, ev[0].b);
        hipEventElapsedTime(&b, ev[1].a, ev[1].b);
        timing->rhs_ms += a;
        timing->column_ms += b;
        timing->rhs_launches += 1;
        timing->column_launches += 1;
      }
    }
  }
  if (timing)
    for (auto& e : ev) {
      hipEventDestroy(e.a);
      hipEventDestroy(e.b);
    }
  // refresh the stored ghosts of every member once
  if (nsteps > 0) {
    int64_t off = 0;
    for (int m = 0; m < nm; ++m) {
      if (int e = launch_bc_f32(&members[m], u + off, st)) return e;
      off += 5 * (int64_t)make_grid(&members[m]).ncell;
    }
  }
  return 0;
}

// ------------------------------------------------------------ HIP graphs ----

namespace snapx {

// Captures whatever enqueue(stream) launches on a private capture stream
// into an instantiated graph.  The capture stream has the default priority.
template <typename F>
static int capture_graph(F enqueue, void** graph_out) {
  SNAPX_CHECK_ARG(graph_out, "graph: null output pointer");
  hipStream_t cs;
  SNAPX_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipStreamDestroy(cs);
    set_error("graph: stream capture could not start");
    return SNAPX_EHIP;
  }
  int status = enqueue(cs);
  hipGraph_t g = nullptr;
  hipError_t ec = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (status) {
    if (g) (void)hipGraphDestroy(g);
    return status;
  }
  if (ec != hipSuccess || !g) {
    set_error("graph: stream capture failed");
    return SNAPX_EHIP;
  }
  hipGraphExec_t exec = nullptr;
  hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    set_error("graph: instantiation failed");
    return SNAPX_EHIP;
  }
  *graph_out = exec;
  return 0;
}

}  // namespace snapx

namespace snapx {

// per-member time steps from device memory into the member table
__global__ void batch_set_dt_kernel(BatchMember* __restrict__ mem, const double* __restrict__ dt,
                                    int nm) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const double d = dt[m];
  mem[m].a.dt = (float)d;
  mem[m].c.dt = d;
}

}  // namespace snapx

extern "C" int snapx_batch_set_dt(const snapx_batch_info* info, const double* dt,
                                  void* workspace, size_t workspace_bytes, void* stream) {
  SNAPX_CHECK_ARG(info && dt && workspace, "snapx_batch_set_dt: null pointer");
  SNAPX_CHECK_ARG(info->nmembers >= 1 && workspace_bytes >= batch_table_bytes(info->nmembers),
                  "snapx_batch_set_dt: workspace too small");
  hipLaunchKernelGGL(batch_set_dt_kernel, dim3((info->nmembers + 63) / 64), dim3(64), 0,
                     as_stream(stream), reinterpret_cast<BatchMember*>(workspace), dt,
                     info->nmembers);
  SNAPX_LAUNCH_CHECK();
  return 0;
}

extern "C" int snapx_fused_graph_create(const snapx_config* cfg, double dt, int nsteps, float* u,
                                        void* workspace, size_t workspace_bytes,
                                        void** graph) {
  if (int e = check_fused(cfg)) return e;
  SNAPX_CHECK_ARG(u && workspace && nsteps >= 1, "snapx_fused_graph_create: bad arguments");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_fused_workspace_bytes(cfg),
                  "snapx_fused_graph_create: workspace too small");
  return capture_graph(
      [&](hipStream_t cs) {
        return snapx_fused_rk3_steps(cfg, dt, nsteps, u, workspace, workspace_bytes, nullptr,
                                     cs);
      },
      graph);
}

extern "C" int snapx_batch_graph_create(const snapx_config* members,
                                        const snapx_batch_info* info, int nsteps, float* u,
                                        void* workspace, size_t workspace_bytes, void** graph) {
  SNAPX_CHECK_ARG(members && info && u && workspace && nsteps >= 1,
                  "snapx_batch_graph_create: bad arguments");
  return capture_graph(
      [&](hipStream_t cs) {
        return snapx_batch_rk3_steps(members, info, nsteps, u, workspace, workspace_bytes,
                                     nullptr, cs);
      },
      graph);
}

extern "C" int snapx_graph_launch(void* graph, void* stream) {
  SNAPX_CHECK_ARG(graph, "snapx_graph_launch: null graph");
  SNAPX_HIP(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph), as_stream(stream)));
  return 0;
}

extern "C" int snapx_graph_destroy(void* graph) {
  if (graph) SNAPX_HIP(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph)));
  return 0;
}

// ------------------------------------------------------ tile autotuner ----

namespace snapx {

// Candidate K1 tiles: x1 extents that keep rows coalesced, then x2 / x3
// extents, within the kernel limits (<= 1024 cells, <= 64 KB LDS).
static std::vector<Tile> tile_candidates(bool two_d, bool three_d) {
  std::vector<Tile> out;
  const int t1s[] = {16, 32, 64, 128, 256};
  const int t2s[] = {1, 2, 4, 8, 16};
  const int t3s[] = {1, 2, 4, 8};
  for (int t1 : t1s)
    for (int t2 : t2s)
      for (int t3 : t3s) {
        if (!two_d && !three_d && (t2 > 1 || t3 > 1)) continue;
        if (two_d && !three_d && (t2 == 1 || t3 > 1)) continue;
        if (three_d && (t2 == 1 || t3 == 1)) continue;
        if (t1 * t2 * t3 > kRhsThreads * kMaxCellsPerThread) continue;
        out.push_back({t1, t2, t3});
      }
  return out;
}

}  // namespace snapx

extern "C" int snapx_batch_autotune(const snapx_config* members, int nmembers, double dt,
                                    const float* u, void* workspace, size_t workspace_bytes,
                                    int trials, void* stream, snapx_batch_info* info,
                                    double* best_ms) {
  SNAPX_CHECK_ARG(members && nmembers >= 1 && u && workspace && info && trials >= 1,
                  "snapx_batch_autotune: bad arguments");
  SNAPX_CHECK_ARG(workspace_bytes >= snapx_batch_workspace_bytes(members, nmembers),
                  "snapx_batch_autotune: workspace too small");
  hipStream_t st = as_stream(stream);
  bool two_d = false, three_d = false;
  for (int m = 0; m < nmembers; ++m) {
    three_d |= members[m].nx3 > 1;
    two_d |= members[m].nx2 > 1;
  }
  char* base = reinterpret_cast<char*>(workspace);
  const BatchMember* table = reinterpret_cast<const BatchMember*>(base);
  float* du = reinterpret_cast<float*>(base + batch_table_bytes(nmembers));
  hipEvent_t e0, e1;
  SNAPX_HIP(hipEventCreate(&e0));
  SNAPX_HIP(hipEventCreate(&e1));
  double best = 1e300;
  Tile best_t = {0, 0, 0};
  int status = 0;
  for (Tile t : tile_candidates(two_d, three_d)) {
    std::vector<BatchMember> tab;
    snapx_batch_info inf;
    // candidates that do not fit some member (LDS) are skipped
    if (batch_layout(members, nmembers, dt, t.tx1, t.tx2, t.tx3, &tab, &inf)) continue;
    if (hipMemcpy(workspace, tab.data(), sizeof(BatchMember) * tab.size(),
                  hipMemcpyHostToDevice) != hipSuccess) {
      set_error("snapx_batch_autotune: table copy failed");
      status = SNAPX_EHIP;
      break;
    }
    int grid = std::min(inf.total_tiles, 512);
    // K1 reads u and writes du only: the state is not changed
    auto rhs = batch_rhs_for(members, nmembers);
    hipLaunchKernelGGL(rhs, dim3(grid), dim3(kRhsThreads), inf.rhs_smem, st, u, du, table,
                       nmembers, inf.total_tiles);
    (void)hipEventRecord(e0, st);
    for (int k = 0; k < trials; ++k)
      hipLaunchKernelGGL(rhs, dim3(grid), dim3(kRhsThreads), inf.rhs_smem, st, u, du, table,
                         nmembers, inf.total_tiles);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
      set_error("snapx_batch_autotune: trial launch failed");
      status = SNAPX_EHIP;
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms / trials < best) best = ms / trials, best_t = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (status) return status;
  SNAPX_CHECK_ARG(best_t.tx1 > 0, "snapx_batch_autotune: no tile fits the members");
  if (best_ms) *best_ms = best;
  return snapx_batch_plan(members, nmembers, dt, best_t.tx1, best_t.tx2, best_t.tx3, workspace,
                          workspace_bytes, info);
}
