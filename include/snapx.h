/*
 * snapx.h -- C ABI of the MI355X-native finite-volume RHS and vertically
 * implicit step (libsnapx.so).
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative SNAPX_E* code;
 *    snapx_last_error() then holds a message (thread-local).  No exception
 *    crosses the ABI.
 *  - All array arguments are device pointers (hipMalloc / torch device
 *    memory), laid out like the reference tensors: a field is
 *    [nvar][nc3][nc2][nc1], x1 (vertical) fastest; an interface-state array
 *    is [2][nvar][nc3][nc2][nc1] with [0] = left (ILT), [1] = right (IRT).
 *    nc? = nx? + 2*nghost on active axes, 1 otherwise
 *    (src/coord/coordinate.hpp:58-60).
 *  - dim is the reference tensor axis: 3 = x1, 2 = x2, 1 = x3
 *    (src/hydro/hydro_forward.cpp:16).
 *  - dtype selects float (SNAPX_F32) or double (SNAPX_F64) element storage.
 *  - stream is a hipStream_t passed as void* (0 = default stream); every
 *    launch goes to it and no entry point synchronises the device unless it
 *    says so.
 *  - The caller owns all memory.  Entry points that need scratch take a
 *    workspace pointer and its size; *_workspace_bytes() returns the size.
 *  - Entry points are reentrant: there is no global mutable state apart from
 *    the thread-local error string, and the static per-GPU workspace of
 *    snapx_fused_rk3_steps_static (the one entry point that allocates
 *    nothing; its calls on one GPU must not overlap).
 *  - Geometry (snapx_config.coord) is Cartesian with uniform face widths
 *    (x?max - x?min)/nx? (src/coord/cartesian.cpp:8-47), or spherical-polar
 *    (r, theta, phi) on the plain FP64/FP32 path: faces uniform in each
 *    coordinate as torch::linspace places them (src/coord/coordinate.cpp:
 *    201-221), with the centres, face areas, volumes and source terms of
 *    src/coord/spherical_polar.cpp, or one gnomonic-equiangle panel
 *    (src/coord/gnomonic_equiangle.cpp) with external boundaries.  The fused
 *    path takes curvilinear blocks with columns of at most 256 (3x3 blocks)
 *    / 128 (5x5) cells; so do batches whose members all share one coordinate
 *    system.  The adjoint takes Cartesian and spherical-polar blocks
 *    (spherical columns of at most 256 / 128 cells), not panels.
 */
#ifndef SNAPX_H
#define SNAPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNAPX_VERSION 1

/* status codes */
#define SNAPX_OK 0
#define SNAPX_EINVAL -1   /* invalid argument / unsupported option */
#define SNAPX_EHIP -2     /* HIP runtime error */
#define SNAPX_ESPACE -3   /* workspace too small */

/* element types */
#define SNAPX_F32 0
#define SNAPX_F64 1

/* interpolation types (src/recon/interpolation.cpp:6-40) */
#define SNAPX_INTERP_DC 0
#define SNAPX_INTERP_PLM 1
#define SNAPX_INTERP_CP3 2
#define SNAPX_INTERP_CP5 3
#define SNAPX_INTERP_WENO3 4
#define SNAPX_INTERP_WENO5 5

/* Riemann solvers (src/riemann/riemann_solver.cpp:47-68) */
#define SNAPX_RIEMANN_ROE 0
#define SNAPX_RIEMANN_LMARS 1
#define SNAPX_RIEMANN_HLLC 2

/* external boundary types (src/bc/bc_func.cpp:7-63) */
#define SNAPX_BC_NONE (-1)
#define SNAPX_BC_REFLECTING 0
#define SNAPX_BC_OUTFLOW 1
#define SNAPX_BC_PERIODIC 2
/* custom_inner / custom_outer: no-ops, the caller owns those ghost cells
   (src/bc/bc_func.cpp:4-5); plain path only (the reference's stage combine
   and body forces still advance those ghosts, which the fused kernels do not
   reproduce) */
#define SNAPX_BC_CUSTOM 3

/* coordinate systems (src/coord/coordinate.cpp:437-452) */
#define SNAPX_COORD_CARTESIAN 0
#define SNAPX_COORD_SPHERICAL_POLAR 1
/* one panel of the cubed sphere with external boundaries: x1 = r, x2 and x3
   equiangular panel coordinates in radians with nx2 = nx3
   (src/coord/gnomonic_equiangle.cpp); plain path, limiter off */
#define SNAPX_COORD_GNOMONIC_EQUIANGLE 2

/* variable kinds for boundary functions (src/snap.h:62-71) */
#define SNAPX_PRIMITIVE 0
#define SNAPX_CONSERVED 1
#define SNAPX_SCALAR 2

/*
 * Block configuration: the hot-path YAML options of one MeshBlock
 * (dynamics.*, forcing.const-gravity.*, integration.implicit-scheme,
 * geometry.*, boundary-condition.external.*).  Plain data, no pointers.
 */
typedef struct snapx_config {
  int nx1, nx2, nx3, nghost;
  double x1min, x1max, x2min, x2max, x3min, x3max;
  double gammad;
  /* reconstruction: [0] = dynamics.reconstruct.vertical (x1),
                     [1] = dynamics.reconstruct.horizontal (x2, x3) */
  int recon_type[2], recon_scale[2], recon_shock[2];
  /* Riemann solver per direction: [0] = x1, [1] = x2, [2] = x3 */
  int riemann[3];
  int limiter;
  double density_floor, pressure_floor, temperature_floor, weight;
  double grav1, grav2, grav3, non_hydrostatic;
  int implicit_scheme; /* 0 none, 1 vic-partial (3x3), 9 vic-full (5x5) */
  int disable_flux[3];
  /* x1-inner, x1-outer, x2-inner, x2-outer, x3-inner, x3-outer */
  int bc[6];
  /* geometry.type: SNAPX_COORD_*; spherical-polar needs x2 in [0, pi];
     gnomonic-equiangle needs nx2 = nx3 and |x2|, |x3| < pi / 2 with ghosts */
  int coord;
} snapx_config;

const char* snapx_last_error(void);
int snapx_version(void);
/* number of HIP devices visible (0 on a host without GPUs) */
int snapx_device_count(void);

/* ---------------------------------------------------------------------- */
/* Module kernels of the plain path (one call = one reference module call) */
/* ---------------------------------------------------------------------- */

/* IdealGasImpl::_cons2prim (src/eos/ideal_gas.cpp:90-109) with the limiters
   of src/eos/equation_of_state.cpp:110-179; replaces the dispatch stub
   ideal_gas_cons2prim (src/eos/eos_dispatch.hpp:9-13).  u may be modified by
   the conserved limiter (as in the reference). */
int snapx_ideal_gas_cons2prim(const snapx_config* cfg, void* u, void* w,
                              int dtype, void* stream);

/* IdealGasImpl::_prim2cons (src/eos/ideal_gas.cpp:67-88), "W->U". */
int snapx_ideal_gas_prim2cons(const snapx_config* cfg, void* w, void* u,
                              int dtype, void* stream);

/* ReconstructImpl::forward (src/recon/reconstruct.cpp:75-154): w
   [nvar][n3][n2][n1] -> wlr [2][nvar][n3][n2][n1], including the dummy-face
   fill and limiter clamps; replaces call_weno5 / call_poly5 / call_weno3 /
   call_poly3 (src/recon/recon_dispatch.hpp:9-18).  revise_x1_lr = 1 applies
   HydroImpl::_revise_x1inner_lr/_revise_x1outer_lr
   (src/hydro/hydro.cpp:217-229) at the physical x1 faces (dim = 3 only). */
int snapx_reconstruct(const void* w, void* wlr, int nvar, int n3, int n2,
                      int n1, int dim, int interp_type, int scale, int shock,
                      int has_parent, int limiter, double density_floor,
                      double pressure_floor, int revise_x1_lr, int nghost,
                      int dtype, void* stream);

/* RiemannSolverImpl::forward for LMARS / HLLC / Roe
   (src/riemann/lmars.cpp:30-103, src/riemann/hllc.cpp:30-106,
   src/riemann/roe.cpp:14-59) over every face of the block, ideal gas;
   replaces call_lmars / call_hllc / call_roe
   (src/riemann/riemann_dispatch.hpp:10-23).  face_pressure may be NULL. */
int snapx_riemann(int solver, const void* wl, const void* wr, void* flx,
                  void* face_pressure, int nvar, int n3, int n2, int n1,
                  int dim, double gammad, int dtype, void* stream);

/* CoordinateImpl::divergence (src/coord/coordinate.cpp:352-396), Cartesian:
   div = sum_d (A F_{+} - A F_{-}) / V on the interior ranges of each active
   axis.  f1/f2/f3 may be NULL for inactive or disabled axes. */
int snapx_divergence(const snapx_config* cfg, const void* f1, const void* f2,
                     const void* f3, void* div, int dtype, void* stream);

/* CoordinateImpl::forward (src/coord/coordinate.cpp:428-434) and
   SphericalPolarImpl::forward (src/coord/spherical_polar.cpp:190-279): the
   flux divergence plus, in spherical-polar geometry, the geometric source
   terms of the momenta, using the primitives w and the x1 face pressure fp1
   ([nc3][nc2][nc1], may be NULL: then 2 p / r stands in).  div is written
   on every cell, ghosts included, as the reference's module returns it. */
int snapx_coord_forward(const snapx_config* cfg, const void* w, const void* f1,
                        const void* f2, const void* f3, const void* fp1,
                        void* div, int dtype, void* stream);

/* HydroImpl::_revise_x1inner_ghost / _revise_x1outer_ghost
   (src/hydro/hydro.cpp:231-269): isentropic hydrostatic refill of the x1
   ghost cells of w (density and pressure). */
int snapx_revise_x1_ghost(const snapx_config* cfg, void* w, int dtype,
                          void* stream);

/* External boundary functions (src/bc/bc_func.cpp:7-63) applied in the order
   of src/mesh/meshblock.cpp:717-722. */
int snapx_apply_bc(const snapx_config* cfg, void* var, int nvar, int var_kind,
                   int dtype, void* stream);

/* Stage combine of the SSP-RK3 integrator, u = w0*u0 + w1*u1 + w2*du
   (src/mesh/meshblock.cpp:675). */
int snapx_rk_combine(void* u, const void* u0, const void* u1, const void* du,
                     double w0, double w1, double w2, int64_t n, int dtype,
                     void* stream);

/* ImplicitHydroImpl::forward (src/implicit/implicit_hydro.cpp:99-189) via
   HydroImpl::_apply_implicit_correction (src/hydro/hydro.cpp:186-215):
   block-tridiagonal assembly, block Thomas sweep and MS-VIC redistribution
   over every interior column; replaces vic_assemble_* / vic_solve_* /
   vic_redistribute_* (src/implicit/implicit_dispatch.hpp:9-19).  du is
   modified in place; w must hold the revised x1 ghosts.  The sweep runs in
   FP64 whatever dtype is. */
size_t snapx_implicit_workspace_bytes(const snapx_config* cfg);
int snapx_implicit_correction(const snapx_config* cfg, double dt, void* du,
                              const void* w, void* workspace,
                              size_t workspace_bytes, int dtype, void* stream);

/* ---------------------------------------------------------------------- */
/* Composed plain path                                                     */
/* ---------------------------------------------------------------------- */

/* HydroImpl::forward (src/hydro/hydro_forward.cpp:14-217): u -> du (dt
   factor included); w (hydro_w) is written, ghosts revised as in the
   reference.  f1/f2/f3 (optional, may be NULL) receive the face fluxes. */
size_t snapx_hydro_workspace_bytes(const snapx_config* cfg, int dtype);
int snapx_hydro_forward(const snapx_config* cfg, double dt, void* u, void* w,
                        void* du, void* f1, void* f2, void* f3,
                        void* workspace, size_t workspace_bytes, int dtype,
                        void* stream);

/* One SSP-RK3 step of MeshBlockImpl::advance_local (three stages of
   hydro forward, stage combine, external boundaries;
   src/mesh/meshblock.cpp:577-786) on the plain path.  u is advanced in
   place; w receives the last stage's primitives. */
size_t snapx_rk3_workspace_bytes(const snapx_config* cfg, int dtype);
int snapx_rk3_step(const snapx_config* cfg, double dt, void* u, void* w,
                   void* workspace, size_t workspace_bytes, int dtype,
                   void* stream);

/* snapx_rk3_step with the time step read from device memory (*dt, a
   double) by the kernels that use it, so that a CFL step computed on the
   device (snapx_max_time_step) needs no host round trip.  Same results as
   snapx_rk3_step with that dt. */
int snapx_rk3_step_dt(const snapx_config* cfg, const double* dt, void* u,
                      void* w, void* workspace, size_t workspace_bytes,
                      int dtype, void* stream);

/* HydroImpl::max_time_step (src/hydro/hydro.cpp:118-180) times cfl
   (src/mesh/meshblock.cpp:511-530): writes the CFL time step to the device
   double *dt_out (no host synchronisation).  A NaN candidate (a negative
   pressure or density in w) gives NaN, as the reference's tensor minimum
   does. */
int snapx_max_time_step(const snapx_config* cfg, const void* w, double cfl,
                        double* dt_out, int dtype, void* stream);

/* Interior positivity check of MeshBlockImpl::check_redo
   (src/mesh/meshblock.cpp:974-1028) on the conserved state u: writes to the
   device int *flag SNAPX_REDO_DENSITY when min(rho) <= 0 and
   SNAPX_REDO_PRESSURE when min(E) <= 0 and the cons2prim pressure (with the
   configured limiter) has min <= 0, both as the reference's NaN-propagating
   minima decide them, plus SNAPX_STATE_NONFINITE when any interior conserved
   value is NaN or infinite (information only; the reference does not redo on
   it).  No host synchronisation and no memset: the flag is cleared by a
   kernel, so the call can be captured in a HIP graph.  u is not modified. */
#define SNAPX_REDO_DENSITY 1
#define SNAPX_REDO_PRESSURE 2
#define SNAPX_STATE_NONFINITE 4
int snapx_check_state(const snapx_config* cfg, const void* u, int* flag, int dtype,
                      void* stream);

/* ---------------------------------------------------------------------- */
/* Fused FP32 path (no reference counterpart; same results as the plain    */
/* path to FP32 rounding)                                                  */
/* ---------------------------------------------------------------------- */

/* One SSP-RK3 step with FP32 state: per stage one fused kernel (cons->prim,
   recon, Riemann, divergence, gravity; interface states stay on chip) and
   one column kernel (implicit solve in FP64, stage combine, boundaries).
   u is float [5][nc3][nc2][nc1], advanced in place.  Requires nvar = 5,
   nghost = 3 and the WENO5/CP5 family. */
size_t snapx_fused_workspace_bytes(const snapx_config* cfg);
int snapx_fused_rk3_step(const snapx_config* cfg, double dt, float* u,
                         void* workspace, size_t workspace_bytes,
                         void* stream);

/* snapx_fused_rk3_step with the time step read from device memory (*dt, a
   double, rounded to FP32 where the FP32 kernels use it, as the host dt is). */
int snapx_fused_rk3_step_dt(const snapx_config* cfg, const double* dt,
                            float* u, void* workspace, size_t workspace_bytes,
                            void* stream);

/* Per-kernel device time accumulated by snapx_fused_rk3_steps. */
typedef struct snapx_timing {
  double rhs_ms;    /* fused RHS kernels (K1) */
  double column_ms; /* implicit column + stage combine kernels (K2) */
  double bc_ms;     /* end-of-step boundary kernels */
  int64_t rhs_launches, column_launches, bc_launches;
} snapx_timing;

/* nsteps fused SSP-RK3 steps with a fixed dt.  With timing == NULL the
   launches are only enqueued; with a timing record every kernel launch is
   bracketed by HIP events on the stream, the stream is synchronised after
   each step and the device times are accumulated into *timing. */
int snapx_fused_rk3_steps(const snapx_config* cfg, double dt, int nsteps,
                          float* u, void* workspace, size_t workspace_bytes,
                          snapx_timing* timing, void* stream);

/* ---------------------------------------------------------------------- */
/* FP32 perturbation path about an FP64 hydrostatic background (no         */
/* reference counterpart: the reference steps the total state in one       */
/* precision)                                                              */
/* ---------------------------------------------------------------------- */

/* The state is carried in FP32 as u' = u - ubar(x1) about a background at
   rest that varies with x1 only, ubar = (rho_bar, 0, 0, 0, p_bar /
   (gamma - 1)), given in FP64.  The fused kernels then difference pressure,
   energy and density perturbations where the total state would cancel
   (the hydrostatic pressure gradient against gravity, the jumps of the Roe
   dissipation), with the background's own flux divergence and gravity as an
   FP64 residual per level; in exact arithmetic the step is the fused step of
   the total state.  Requires a Cartesian block, vertical reconstruction
   weno5 with shock: false, scale off, grav2 = grav3 = 0, and columns of at
   most 512 (3x3 blocks) / 128 (5x5) cells.

   snapx_fused_background fills the device tables bg (bytes from
   snapx_fused_background_bytes) from the host arrays rho_bar[nc1],
   p_bar[nc1] (x1 cell levels, ghosts included; synchronises the stream).
   snapx_fused_split writes u' = u - ubar (u float or double per dtype, every
   cell); snapx_fused_join writes u = ubar + u' and applies the external
   boundaries to u.  snapx_fused_rk3_steps_bg advances u' nsteps fused steps
   (workspace as snapx_fused_workspace_bytes; the stored ghosts of u' are not
   refreshed: join restores them on the total state). */
size_t snapx_fused_background_bytes(const snapx_config* cfg);
int snapx_fused_background(const snapx_config* cfg, const double* rho_bar,
                           const double* p_bar, void* bg, size_t bg_bytes,
                           void* stream);
int snapx_fused_split(const snapx_config* cfg, const void* bg, const void* u,
                      int dtype, float* u_pert, void* stream);
int snapx_fused_join(const snapx_config* cfg, const void* bg,
                     const float* u_pert, void* u, int dtype, void* stream);
int snapx_fused_rk3_steps_bg(const snapx_config* cfg, const void* bg,
                             double dt, int nsteps, float* u_pert,
                             void* workspace, size_t workspace_bytes,
                             snapx_timing* timing, void* stream);

/* No reference counterpart.  For callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc): snapx_fused_rk3_steps (timing NULL) with its workspace in a
   static device array of SNAPX_STATIC_SCRATCH_BYTES per GPU, which each GPU
   the library runs on holds until the process ends.  Calls no hipMalloc,
   hipFree or allocating library routine.  Blocks whose workspace
   (snapx_fused_workspace_bytes) exceeds the array are rejected with
   SNAPX_ESPACE.  Calls on one GPU share the array: calls that overlap on one
   GPU (two streams, two threads) race and must be serialised by the caller. */
#define SNAPX_STATIC_SCRATCH_BYTES (64u << 20)
int snapx_fused_rk3_steps_static(const snapx_config* cfg, double dt, int nsteps,
                                 float* u, void* stream);

/* ---------------------------------------------------------------------- */
/* Adjoint (no reference counterpart)                                      */
/* ---------------------------------------------------------------------- */

/* Discrete adjoint of one FP64 plain-path SSP-RK3 step
   (snapx_rk3_step): on entry lambda (double [5][nc3][nc2][nc1]) holds
   dL/du_out; on exit it holds dL/du_in for the step input u_in.  The
   forward stages are recomputed from u_in.  Requires limiter off,
   non-hydrostatic = 1 and no disable-flux; Cartesian or spherical-polar
   geometry (on spherical-polar blocks the adjoint includes the per-cell
   areas and volumes, the geometric sources with the x1 face pressure, the
   per-layer column geometry and the ghost distances). */
size_t snapx_adjoint_workspace_bytes(const snapx_config* cfg);
int snapx_adjoint_rk3_step(const snapx_config* cfg, double dt,
                           const double* u_in, double* lambda,
                           void* workspace, size_t workspace_bytes,
                           void* stream);

/* snapx_adjoint_rk3_step in two halves, so that a checkpointed trajectory
   replays each segment once: snapx_adjoint_record recomputes the step from
   u_in into a tape of snapx_adjoint_tape_bytes (stage inputs, primitives,
   explicit tendencies, VIC blocks) and, when u_out is not NULL, writes the
   step's output (the replay's next input; it agrees with snapx_rk3_step to
   rounding: the recompute solves the columns by cyclic reduction);
   snapx_adjoint_sweep then maps lambda = dL/du_out to dL/du_in from that
   tape (and may overwrite the tape).  The work space
   (snapx_adjoint_work_bytes) is scratch shared by both; a step's workspace
   for snapx_adjoint_rk3_step is one work space followed by one tape.  Record
   followed by sweep gives the bits of snapx_adjoint_rk3_step.  No reference
   counterpart. */
size_t snapx_adjoint_tape_bytes(const snapx_config* cfg);
size_t snapx_adjoint_work_bytes(const snapx_config* cfg);
int snapx_adjoint_record(const snapx_config* cfg, double dt,
                         const double* u_in, double* u_out, void* tape,
                         size_t tape_bytes, void* work, size_t work_bytes,
                         void* stream);
int snapx_adjoint_sweep(const snapx_config* cfg, double dt, void* tape,
                        size_t tape_bytes, double* lambda, void* work,
                        size_t work_bytes, void* stream);

/* The same adjoint with FP32 storage: u_in, u_out and lambda are float
   [5][nc3][nc2][nc1], and the tape holds FP32 stage inputs, primitives and
   explicit tendencies.  Arithmetic in registers is FP64; the VIC blocks,
   the column solves (forward and transposed) and their tape entries stay
   FP64.  The recompute runs the explicit part on the FP32 plain path
   (snapx_hydro_forward, SNAPX_F32), so the stage states follow an FP32
   trajectory with FP64 column solves.  Tape and work space are about half
   the FP64 sizes.  No reference counterpart. */
size_t snapx_adjoint_f32_tape_bytes(const snapx_config* cfg);
size_t snapx_adjoint_f32_work_bytes(const snapx_config* cfg);
size_t snapx_adjoint_f32_workspace_bytes(const snapx_config* cfg);
int snapx_adjoint_f32_rk3_step(const snapx_config* cfg, double dt,
                               const float* u_in, float* lambda,
                               void* workspace, size_t workspace_bytes,
                               void* stream);
int snapx_adjoint_f32_record(const snapx_config* cfg, double dt,
                             const float* u_in, float* u_out, void* tape,
                             size_t tape_bytes, void* work, size_t work_bytes,
                             void* stream);
int snapx_adjoint_f32_sweep(const snapx_config* cfg, double dt, void* tape,
                            size_t tape_bytes, float* lambda, void* work,
                            size_t work_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched members (ensembles / blocks of uneven size; no reference        */
/* counterpart: the reference runs one block per thread of                 */
/* src/mesh/mesh.cpp:37-122)                                               */
/* ---------------------------------------------------------------------- */

/* Plan of a batch, filled by snapx_batch_plan (plain data). */
typedef struct snapx_batch_info {
  int nmembers;
  int tile_x1, tile_x2, tile_x3; /* K1 tile shape shared by all members */
  int total_tiles, total_cols;   /* work-list lengths of the persistent kernels */
  int vic_size;                  /* 3 or 5 */
  int64_t total_elems;           /* floats of the concatenated state */
  int64_t scratch_doubles;
  size_t rhs_smem;
  double dt;
  int col_groups; /* work items of the cell-parallel column kernel, one row
                     per lane (col_groups + col_groups2 = 0: the lane-per-column
                     kernel serves the batch) */
  int col_groups2; /* work items of the cell-parallel column kernel with two
                      rows per lane (3x3 columns of more than 64 cells) */
} snapx_batch_info;

/* Members are independent blocks, each with its own snapx_config (same
   implicit scheme); their FP32 states are concatenated in member order, each
   [5][nc3][nc2][nc1].  snapx_batch_plan writes the member table into the
   workspace (synchronous copy) for a fixed dt and tile shape (0 = default);
   snapx_batch_rk3_steps then advances every member nsteps SSP-RK3 steps with
   one persistent launch per kernel and stage for the whole batch (it only
   enqueues kernels, so a caller may capture it in a HIP graph when timing is
   NULL). */
size_t snapx_batch_workspace_bytes(const snapx_config* members, int nmembers);
int snapx_batch_plan(const snapx_config* members, int nmembers, double dt,
                     int tile_x1, int tile_x2, int tile_x3, void* workspace,
                     size_t workspace_bytes, snapx_batch_info* info);
int snapx_batch_rk3_steps(const snapx_config* members,
                          const snapx_batch_info* info, int nsteps, float* u,
                          void* workspace, size_t workspace_bytes,
                          snapx_timing* timing, void* stream);

/* Replaces the planned dt of every member with dt[m] read from device memory
   (one double per member) by a kernel on the stream: the batched steps (and
   graphs of them) enqueued after it use the new time steps; info->dt keeps
   the planned value. */
int snapx_batch_set_dt(const snapx_batch_info* info, const double* dt,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Tile autotuner of the batched path (no reference counterpart).  Times
   `trials` launches of the RHS kernel (it reads u, writes only workspace) for
   every candidate tile that fits the members, then plans the batch with the
   fastest one (as snapx_batch_plan; *best_ms = its average launch time). */
int snapx_batch_autotune(const snapx_config* members, int nmembers, double dt,
                         const float* u, void* workspace, size_t workspace_bytes,
                         int trials, void* stream, snapx_batch_info* info,
                         double* best_ms);

/* ---------------------------------------------------------------------- */
/* CFL time loop with redo on the device (examples/shock.cpp:59-79 around  */
/* MeshBlockImpl::max_time_step and check_redo,                            */
/* src/mesh/meshblock.cpp:511-530,974-1028)                                */
/* ---------------------------------------------------------------------- */

/* Loop state in device memory.  The caller fills the settings, zeroes the
   rest and copies the record to the device; each attempt of the loop then
   runs as kernels that read and update it, so any number of attempts is
   enqueued without a host synchronisation and can be captured in one HIP
   graph (no memset nodes).  One attempt is one pass of the reference loop:
     stop when status != 0, nlim >= 0 && cycle >= nlim, or time >= tlim;
     ++cycle;  dt = 2^-redo * cfl * HydroImpl::max_time_step(w);
     one SSP-RK3 step with dt;
     verdict = snapx_check_state(u);
     if (verdict & redo_mask): ++redo; if redo > max_redo: status = -1 (u
       keeps the failed step, as the reference returns before restoring);
       else u = the step's input, --cycle;
     else: redo = 0, time += dt.
   Attempts after the loop has stopped change nothing (their step is
   discarded: u and w are restored). */
typedef struct snapx_step_control {
  /* settings */
  double cfl;        /* Courant number */
  double tlim;       /* the loop stops once time >= tlim */
  int64_t nlim;      /* ... or once nlim >= 0 and cycle >= nlim */
  int32_t max_redo;  /* more redos in a row than this end the loop */
  int32_t redo_mask; /* verdict bits that ask for a redo; the reference's:
                        SNAPX_REDO_DENSITY | SNAPX_REDO_PRESSURE */
  /* loop state, updated on the device */
  double time;       /* current time */
  double dt;         /* time step of the latest attempt */
  double dt_raw;     /* HydroImpl::max_time_step of the latest attempt */
  int64_t cycle;     /* MeshBlock::cycle */
  int32_t redo;      /* current_redo */
  int32_t status;    /* 0 running, 1 stopped on tlim / nlim, -1 more than
                        max_redo redos in a row */
  int32_t verdict;   /* snapx_check_state bits of the latest attempt */
  int32_t active;    /* 1 when the latest attempt was a pass of the loop */
  int32_t steps;     /* accepted steps */
  int32_t redos;     /* redos in total */
  int32_t restore;   /* kernel scratch */
  int32_t reserved2;
} snapx_step_control;

/* Where an attempt's CFL step reads the primitives (dt_source): */
#define SNAPX_DT_FROM_W 0 /* w as the previous attempt left it, the last
                             stage's primitives (the reference) */
#define SNAPX_DT_FROM_U 1 /* the primitives of the attempt's input u */

/* nattempts attempts on the plain path (snapx_rk3_step_dt).  With
   SNAPX_DT_FROM_W, w holds the primitives of the first attempt's CFL step
   (after initialisation, the initial primitives).  dt_log, when not NULL,
   receives one double per attempt: its dt, or 0 for an attempt after the
   loop stopped.  ctrl, u, w, dt_log and the workspace are device memory. */
size_t snapx_rk3_attempts_workspace_bytes(const snapx_config* cfg, int dtype);
int snapx_rk3_attempts(const snapx_config* cfg, snapx_step_control* ctrl,
                       int nattempts, int dt_source, void* u, void* w,
                       double* dt_log, void* workspace, size_t workspace_bytes,
                       int dtype, void* stream);

/* The same loop on the fused FP32 path (snapx_fused_rk3_step_dt).  The fused
   step keeps no primitives, so the CFL step reads those of the attempt's
   input (SNAPX_DT_FROM_U). */
size_t snapx_fused_attempts_workspace_bytes(const snapx_config* cfg);
int snapx_fused_attempts(const snapx_config* cfg, snapx_step_control* ctrl,
                         int nattempts, float* u, double* dt_log,
                         void* workspace, size_t workspace_bytes,
                         void* stream);

/* ---------------------------------------------------------------------- */
/* HIP graphs (no reference counterpart)                                   */
/* ---------------------------------------------------------------------- */

/* Capture nsteps fused steps (snapx_fused_rk3_steps) or batched steps
   (snapx_batch_rk3_steps) on the given buffers into an executable graph;
   snapx_graph_launch replays it on a stream.  The caller keeps u and the
   workspace alive and unchanged in place while the graph exists. */
int snapx_fused_graph_create(const snapx_config* cfg, double dt, int nsteps,
                             float* u, void* workspace, size_t workspace_bytes,
                             void** graph);
int snapx_batch_graph_create(const snapx_config* members,
                             const snapx_batch_info* info, int nsteps, float* u,
                             void* workspace, size_t workspace_bytes,
                             void** graph);
int snapx_graph_launch(void* graph, void* stream);
int snapx_graph_destroy(void* graph);

#ifdef __cplusplus
}
#endif
#endif /* SNAPX_H */
