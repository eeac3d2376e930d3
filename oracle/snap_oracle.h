/*
 * snap_oracle.h -- CPU restatement (FP64) of snapy's finite-volume RHS and
 * vertically implicit step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library may include, link
 * or call this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU timing
 * baseline.
 *
 * Layout follows the reference: a field is [nvar][n3][n2][n1] (x1 fastest,
 * x1 is the vertical axis), index ((v*n3 + k)*n2 + j)*n1 + i.
 * dim follows the reference tensor-axis convention: 3 = x1, 2 = x2, 1 = x3
 * (src/hydro/hydro_forward.cpp:16).
 *
 * Each function cites the reference file:line it restates.  See the header
 * comment of snap_oracle.c for what pins each function.
 */
#ifndef SNAP_ORACLE_H
#define SNAP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OC_IDN = 0, OC_IVX = 1, OC_IVY = 2, OC_IVZ = 3, OC_IPR = 4, OC_ICY = 5 };
enum { OC_ILT = 0, OC_IRT = 1 };

/* interpolation types (src/recon/interpolation.cpp:6-40) */
enum {
  OC_INTERP_DC = 0,
  OC_INTERP_PLM = 1,
  OC_INTERP_CP3 = 2,
  OC_INTERP_CP5 = 3,
  OC_INTERP_WENO3 = 4,
  OC_INTERP_WENO5 = 5
};

/* riemann solver types (src/riemann/riemann_solver.cpp:47-68) */
enum { OC_RIEMANN_ROE = 0, OC_RIEMANN_LMARS = 1, OC_RIEMANN_HLLC = 2 };

/* boundary types (src/bc/bc_func.cpp:7-63) */
enum {
  OC_BC_NONE = -1,
  OC_BC_REFLECTING = 0,
  OC_BC_OUTFLOW = 1,
  OC_BC_PERIODIC = 2,
  OC_BC_CUSTOM = 3 /* custom_inner / custom_outer (src/bc/bc_func.cpp:4-5): no-ops */
};

/* variable types for boundary functions (src/snap.h:62-71) */
enum { OC_PRIMITIVE = 0, OC_CONSERVED = 1, OC_SCALAR = 2 };

typedef struct {
  int nx1, nx2, nx3, nghost;
  double x1min, x1max, x2min, x2max, x3min, x3max;
  double gammad;
  /* reconstruction: [0] = vertical (x1), [1] = horizontal (x2/x3) */
  int recon_type[2], recon_scale[2], recon_shock[2];
  /* riemann solver per direction: [0]=x1, [1]=x2, [2]=x3 */
  int riemann[3];
  int limiter;
  double density_floor, pressure_floor, temperature_floor, weight;
  double grav1, grav2, grav3, non_hydrostatic;
  int implicit_scheme;
  int disable_flux[3];
  /* external BCs: x1-inner, x1-outer, x2-inner, x2-outer, x3-inner, x3-outer */
  int bc[6];
  /* geometry.type: OC_COORD_* */
  int coord;
} oc_config;

#define OC_COORD_CARTESIAN 0
#define OC_COORD_SPHERICAL_POLAR 1
#define OC_COORD_GNOMONIC_EQUIANGLE 2

/* ---- grid helpers ---- */
int oc_nc1(const oc_config* c);
/* gnomonic face frames (src/coord/gnomonic_equiangle.cpp:236-360) */
void oc_gno_prim2local(const oc_config* c, double* w, int dim);
void oc_gno_flux2global(const oc_config* c, double* f, int dim);
int oc_nc2(const oc_config* c);
int oc_nc3(const oc_config* c);
/* Cartesian face widths dx?f (length nc?) (src/coord/cartesian.cpp:8-47) */
void oc_cartesian_widths(const oc_config* c, double* dx1f, double* dx2f,
                         double* dx3f);

/* ---- scalar interpolants (src/recon/interp_simple.hpp) ---- */
double oc_interp_cp3(double phim1, double phi, double phip1);
double oc_interp_cp5(double phim2, double phim1, double phi, double phip1,
                     double phip2);
double oc_interp_weno3(double phim1, double phi, double phip1);
double oc_interp_weno5(double phim2, double phim1, double phi, double phip1,
                       double phip2);
double oc_interp_plm(double phim1, double phi, double phip1);

double oc_interp_window(int type, int right, int scale, const double* phi);

/* ---- EOS (src/eos/ideal_gas.cpp, src/eos/equation_of_state.cpp) ---- */
void oc_cons2prim(const oc_config* c, double* u, double* w, int64_t ncell);
void oc_prim2cons(const oc_config* c, double* w, double* u, int64_t ncell);

/* ---- reconstruction (src/recon/reconstruct.cpp:75-154) ---- */
void oc_reconstruct(const double* w, double* wlr, int nvar, int n3, int n2,
                    int n1, int dim, int type, int scale, int shock,
                    int has_parent, int limiter, double dfloor,
                    double pfloor);

/* ---- Riemann solvers (src/riemann/{lmars,hllc,roe}_impl.h) ---- */
void oc_riemann(int solver, double* wl, double* wr, double* flx, double* fp,
                int nvar, int n3, int n2, int n1, int dim, double gammad);

/* ---- divergence (src/coord/coordinate.cpp:352-396) ---- */
void oc_divergence(const double* f1, const double* f2, const double* f3,
                   double* div, int nvar, int n3, int n2, int n1, int nghost,
                   const double* dx1f, const double* dx2f, const double* dx3f);

/* ---- coordinate forward: divergence + spherical-polar sources
   (src/coord/coordinate.cpp:428-434, src/coord/spherical_polar.cpp:212-279);
   fp1 may be NULL; div covers every cell ---- */
void oc_coord_forward(const oc_config* c, const double* w, const double* f1,
                      const double* f2, const double* f3, const double* fp1,
                      double* div);
/* geometry arrays: 0 x1v, 1 x2v, 2 dx1v, 3 volume, 4/5/6 face areas,
   7/8/9 CFL widths */
void oc_geometry(const oc_config* c, int what, double* out);

/* ---- gravity boundary revisions (src/hydro/hydro.cpp:217-269) ---- */
void oc_revise_x1_ghost(double* w, int nvar, int n3, int n2, int n1,
                        int nghost, double grav1, double gammad,
                        const double* dx1v);
void oc_revise_x1_lr(double* wlr, int nvar, int n3, int n2, int n1,
                     int nghost);

/* ---- vertically implicit correction (src/implicit) ---- */
void oc_implicit_correction(const oc_config* c, double dt, double* du,
                            const double* w);
/* pieces of the MS-VIC redistribution for an arbitrary species count ny
   (src/implicit/vic_redistribute_impl.h:26-106); delta is [nlayer][N],
   a is [nlayer][N*N] (row-major) */
void oc_vic_backward_reduce(int N, double* du, const double* w, double* a,
                            double* delta, const double* vol, double* scratch,
                            int il, int iu, int dir, int ny, int stride1,
                            int stride2);
void oc_vic_redistribute_cell(int N, double* du, const double* w,
                              double* mass_fix, const double* delta,
                              const double* vol, const double* scratch, int i,
                              int dir, int ny, int stride1, int stride2);
/* dense helpers exposed for brute-force checks */
void oc_inverse3(const double* m, double* inv);
/* flux Jacobian dF1/dU at primitive state w (flux_decomposition_impl.h:107-128) */
void oc_flux_jacobian5(double gm1, const double* w, double* J);
/* Roe-averaged primitive state and sound speed of an interface
   (flux_decomposition_impl.h:16-70) */
void oc_roe_state(double gammad, const double* wl, const double* wr, double* prim,
                  double* cs);
/* R, R^-1, signed eigenvalues and |A| = R|Lambda|R^-1 at the Roe state
   (flux_decomposition_impl.h:72-105, vic_assemble_partial_impl.h:70-71) */
void oc_roe_eigensystem(double gammad, const double* wl, const double* wr, double* R,
                        double* Ri, double* lam, double* absA);
/* block coefficients of one column: w points at the first interior cell,
   w[v*stride1 + i] valid for i = -1..nl; area has nl+1 faces, vol nl cells
   (vic_assemble_partial_impl.h:28-111, vic_assemble_full_impl.h:15-106) */
void oc_vic_assemble(int N, const double* w, long stride1, int nl, double gammad,
                     const double* area, const double* vol, double dt, double grav,
                     double* a, double* b, double* c);
void oc_vic_column_matrices(const oc_config* c, double dt, const double* w,
                            int k, int j, double* a, double* b, double* cc);

/* ---- hydro RHS (src/hydro/hydro_forward.cpp:14-217) ---- */
void oc_hydro_forward(const oc_config* c, double dt, double* u, double* w,
                      double* du, double* f1, double* f2, double* f3);

/* ---- boundary conditions (src/bc/bc_func.cpp:7-63) ---- */
void oc_apply_bc(const oc_config* c, double* var, int nvar, int var_type);

/* ---- time stepping (src/mesh/meshblock.cpp:511-544,577-786) ---- */
double oc_max_time_step(const oc_config* c, const double* w, double cfl);
void oc_rk3_step(const oc_config* c, double dt, double* u, double* w);
/* initialisation from primitives: BCs on w, W->U, BCs on u
   (src/mesh/meshblock.cpp:412-510) */
void oc_initialize(const oc_config* c, double* w, double* u);

/* planted mistakes, for the tests that show each independent check fails
   on a plausible error (0 = none) */
extern int oc_plant;
enum {
  OC_PLANT_NONE = 0,
  OC_PLANT_C_JAC_I = 1,        /* c_i uses J_i in place of J_{i+1} */
  OC_PLANT_NO_BOTTOM_FOLD = 2, /* reflecting fold skipped at the column bottom */
  OC_PLANT_PHI_PLACE = 3,      /* energy row of Phi on the density column */
  OC_PLANT_B_ABS_PLUS = 4,     /* b_i uses |A| of the upper face */
  OC_PLANT_RINV_SIGN = 5,      /* sign of one left-eigenvector entry */
  OC_PLANT_FLUXJAC = 6,        /* v1*v1 -> v1*v2 in one flux-Jacobian entry */
  OC_PLANT_ROE_VELOCITY = 7,   /* arithmetic in place of sqrt(rho)-weighted Roe velocity */
  OC_PLANT_ROE_UPWIND_SWAP = 8, /* Roe supersonic switches take the downwind flux */
  OC_PLANT_ROE_LLF_SPEED = 9,  /* LLF speed without the factor 1/2 */
  OC_PLANT_GNO_FLUX_COS = 10,  /* gnomonic x2-face flux2global with +cos in T32 */
  OC_PLANT_GNO_XOV_SIN = 11,   /* x2 metric source with the cell sine for the face sine */
  OC_PLANT_GNO_NO_RAISE = 12,  /* cons2prim without the contravariant raise */
  OC_PLANT_GNO_ANG3 = 13       /* dx3f_ang with y1 * y1 for y1 * y2 */
};

void oc_set_threads(int n);
int oc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
