// C++ host modules mirroring the reference's hot-path module API
// (namespace snap): options structs with from_yaml, and the EquationOfState,
// Reconstruct, RiemannSolver, Coordinate, ImplicitHydro, Hydro and MeshBlock
// modules.  Tensors (libtorch, HIP device memory) are the data carrier; every
// module forwards to the C ABI of libsnapx (include/snapx.h) on torch's
// current HIP stream.  There is no CPU path: device tensors are required.
#pragma once

#include <torch/torch.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "mini_yaml.hpp"
#include "snapx.h"

// fluent option accessors like the reference's ADD_ARG (src/add_arg.h:9-25)
#define SNAP_ARG(T, name)                                       \
 public:                                                        \
  inline auto name(const T& new_##name)->decltype(*this) {      \
    this->name##_ = new_##name;                                 \
    return *this;                                               \
  }                                                             \
  inline const T& name() const noexcept { return this->name##_; } \
  inline T& name() noexcept { return this->name##_; }           \
                                                                \
 protected:                                                     \
  T name##_

namespace snap {

enum { IDN = 0, IVX = 1, IVY = 2, IVZ = 3, IPR = 4, ICY = 5 };
enum { ILT = 0, IRT = 1 };
enum { kPrimitive = 0, kConserved = 1, kScalar = 2 };

using Variables = std::map<std::string, torch::Tensor>;

// ---------------------------------------------------------------- options

// src/eos/equation_of_state.hpp + from_yaml (src/eos/equation_of_state.cpp:26-80)
struct EquationOfStateOptionsImpl {
  static std::shared_ptr<EquationOfStateOptionsImpl> create() {
    return std::make_shared<EquationOfStateOptionsImpl>();
  }
  static std::shared_ptr<EquationOfStateOptionsImpl> from_yaml(std::string const& filename,
                                                               bool verbose = false);
  static std::shared_ptr<EquationOfStateOptionsImpl> from_yaml(yaml::Node const& root,
                                                               bool verbose = false);
  SNAP_ARG(std::string, type) = "ideal-gas";
  SNAP_ARG(double, gammad) = 1.4;
  SNAP_ARG(double, weight) = 29.e-3;
  SNAP_ARG(double, density_floor) = 1.e-10;
  SNAP_ARG(double, pressure_floor) = 1.e-10;
  SNAP_ARG(double, temperature_floor) = 20.;
  SNAP_ARG(bool, limiter) = false;
  SNAP_ARG(bool, verbose) = false;
};
using EquationOfStateOptions = std::shared_ptr<EquationOfStateOptionsImpl>;

// src/recon/interpolation.hpp:16-31
struct InterpOptionsImpl {
  static std::shared_ptr<InterpOptionsImpl> create() {
    return std::make_shared<InterpOptionsImpl>();
  }
  InterpOptionsImpl() = default;
  explicit InterpOptionsImpl(std::string name) { type(name); }
  SNAP_ARG(std::string, type) = "dc";
  SNAP_ARG(bool, scale) = false;
};
using InterpOptions = std::shared_ptr<InterpOptionsImpl>;

// src/recon/reconstruct.hpp:23-58 + from_yaml (src/recon/reconstruct.cpp:13-35)
struct ReconstructOptionsImpl {
  static std::shared_ptr<ReconstructOptionsImpl> create() {
    auto op = std::make_shared<ReconstructOptionsImpl>();
    op->interp() = InterpOptionsImpl::create();
    return op;
  }
  static std::shared_ptr<ReconstructOptionsImpl> from_yaml(std::string const& filename,
                                                           std::string const& section);
  static std::shared_ptr<ReconstructOptionsImpl> from_yaml(yaml::Node const& node,
                                                           std::string const& section);
  SNAP_ARG(bool, is_boundary_lower) = false;
  SNAP_ARG(bool, is_boundary_upper) = false;
  SNAP_ARG(bool, shock) = true;
  SNAP_ARG(double, density_floor) = 1.e-10;
  SNAP_ARG(double, pressure_floor) = 1.e-10;
  SNAP_ARG(bool, limiter) = false;
  SNAP_ARG(InterpOptions, interp) = nullptr;
};
using ReconstructOptions = std::shared_ptr<ReconstructOptionsImpl>;

// src/riemann/riemann_solver.hpp:26-38 + from_yaml
// (src/riemann/riemann_solver.cpp:10-28).  type_x1/x2/x3 extend the
// reference with one solver per direction (empty = use type).
struct RiemannSolverOptionsImpl {
  static std::shared_ptr<RiemannSolverOptionsImpl> create() {
    return std::make_shared<RiemannSolverOptionsImpl>();
  }
  static std::shared_ptr<RiemannSolverOptionsImpl> from_yaml(std::string const& filename,
                                                             std::string const& section);
  static std::shared_ptr<RiemannSolverOptionsImpl> from_yaml(yaml::Node const& node);
  std::string type_of(int axis) const;  // axis 1, 2, 3
  SNAP_ARG(std::string, type) = "hllc";
  SNAP_ARG(std::string, dir) = "omni";
  SNAP_ARG(std::string, type_x1) = "";
  SNAP_ARG(std::string, type_x2) = "";
  SNAP_ARG(std::string, type_x3) = "";
};
using RiemannSolverOptions = std::shared_ptr<RiemannSolverOptionsImpl>;

// src/forcing/forcing.hpp:25-53 + from_yaml (src/forcing/const_gravity.cpp:10-24)
struct ConstGravityOptionsImpl {
  static std::shared_ptr<ConstGravityOptionsImpl> create() {
    return std::make_shared<ConstGravityOptionsImpl>();
  }
  static std::shared_ptr<ConstGravityOptionsImpl> from_yaml(yaml::Node const& forcing);
  SNAP_ARG(double, non_hydrostatic) = 1.;
  SNAP_ARG(double, grav1) = 0.;
  SNAP_ARG(double, grav2) = 0.;
  SNAP_ARG(double, grav3) = 0.;
};
using ConstGravityOptions = std::shared_ptr<ConstGravityOptionsImpl>;

// src/implicit/implicit_hydro.hpp:16-42 + from_yaml (implicit_hydro.cpp:20-48)
struct ImplicitOptionsImpl {
  static std::shared_ptr<ImplicitOptionsImpl> create() {
    return std::make_shared<ImplicitOptionsImpl>();
  }
  static std::shared_ptr<ImplicitOptionsImpl> from_yaml(std::string const& filename,
                                                        bool verbose = false);
  static std::shared_ptr<ImplicitOptionsImpl> from_yaml(yaml::Node const& node);
  int size() const { return ((scheme() >> 3) & 1) ? 5 : 3; }
  std::string type() const;
  SNAP_ARG(int, scheme) = 0;
};
using ImplicitOptions = std::shared_ptr<ImplicitOptionsImpl>;

// src/coord/coordinate.hpp (Cartesian subset) + from_yaml
// (src/coord/coordinate.cpp:58-166, one block)
struct CoordinateOptionsImpl {
  static std::shared_ptr<CoordinateOptionsImpl> create() {
    return std::make_shared<CoordinateOptionsImpl>();
  }
  static std::shared_ptr<CoordinateOptionsImpl> from_yaml(std::string const& filename);
  static std::shared_ptr<CoordinateOptionsImpl> from_yaml(yaml::Node const& root);
  int nc1() const { return nx1() > 1 ? nx1() + 2 * nghost() : 1; }
  int nc2() const { return nx2() > 1 ? nx2() + 2 * nghost() : 1; }
  int nc3() const { return nx3() > 1 ? nx3() + 2 * nghost() : 1; }
  SNAP_ARG(std::string, type) = "cartesian";
  SNAP_ARG(double, x1min) = 0.;
  SNAP_ARG(double, x1max) = 1.;
  SNAP_ARG(double, x2min) = 0.;
  SNAP_ARG(double, x2max) = 1.;
  SNAP_ARG(double, x3min) = 0.;
  SNAP_ARG(double, x3max) = 1.;
  SNAP_ARG(int, nx1) = 1;
  SNAP_ARG(int, nx2) = 1;
  SNAP_ARG(int, nx3) = 1;
  SNAP_ARG(int, nghost) = 1;
};
using CoordinateOptions = std::shared_ptr<CoordinateOptionsImpl>;

// src/hydro/hydro.hpp:24-72 + from_yaml (src/hydro/hydro_options.cpp:14-102)
struct HydroOptionsImpl {
  static std::shared_ptr<HydroOptionsImpl> create() {
    return std::make_shared<HydroOptionsImpl>();
  }
  static std::shared_ptr<HydroOptionsImpl> from_yaml(std::string const& filename,
                                                     bool verbose = false);
  SNAP_ARG(bool, verbose) = false;
  SNAP_ARG(bool, disable_flux_x1) = false;
  SNAP_ARG(bool, disable_flux_x2) = false;
  SNAP_ARG(bool, disable_flux_x3) = false;
  SNAP_ARG(ConstGravityOptions, grav) = nullptr;
  SNAP_ARG(EquationOfStateOptions, eos) = nullptr;
  SNAP_ARG(ReconstructOptions, recon1) = nullptr;
  SNAP_ARG(ReconstructOptions, recon23) = nullptr;
  SNAP_ARG(RiemannSolverOptions, riemann) = nullptr;
  SNAP_ARG(ImplicitOptions, icorr) = nullptr;
};
using HydroOptions = std::shared_ptr<HydroOptionsImpl>;

// One entry of the "outputs" list (src/output/output_type.cpp:32-83).  This
// library writes the "restart" type; "netcdf" and "pnetcdf" entries keep their
// schedule (file numbers and next times, which restart files record) but
// write nothing.
struct OutputOptionsImpl {
  static std::shared_ptr<OutputOptionsImpl> create() {
    return std::make_shared<OutputOptionsImpl>();
  }
  static std::shared_ptr<OutputOptionsImpl> from_yaml(yaml::Node const& node, int fid);
  SNAP_ARG(int, fid) = 0;
  SNAP_ARG(double, dt) = 0.;
  SNAP_ARG(std::string, file_type) = "";
  SNAP_ARG(std::vector<std::string>, variables);
  SNAP_ARG(bool, combine) = true;
};
using OutputOptions = std::shared_ptr<OutputOptionsImpl>;

// the schedule of one output (src/output/output_type.hpp:100-112)
struct OutputType {
  OutputOptions options;
  int file_number = 0;
  double next_time = 0.;
};

// MeshBlock options of the hot path (src/mesh/meshblock_options.cpp:12-195):
// geometry, hydro, integration, outputs and external boundary functions.
struct MeshBlockOptionsImpl {
  static std::shared_ptr<MeshBlockOptionsImpl> create() {
    return std::make_shared<MeshBlockOptionsImpl>();
  }
  static std::shared_ptr<MeshBlockOptionsImpl> from_yaml(std::string const& filename,
                                                         bool verbose = false);
  SNAP_ARG(CoordinateOptions, coord) = nullptr;
  SNAP_ARG(HydroOptions, hydro) = nullptr;
  SNAP_ARG(std::string, integrator) = "rk3";
  SNAP_ARG(double, cfl) = 0.9;
  SNAP_ARG(int, nlim) = -1;
  SNAP_ARG(double, tlim) = 1.e9;
  // integration.max-redo: redo attempts with dt halved before check_redo
  // gives up (the count of harp::IntegratorOptions::max_redo)
  SNAP_ARG(int, max_redo) = 5;
  // x1-inner, x1-outer, x2-inner, x2-outer, x3-inner, x3-outer ("" = none)
  SNAP_ARG(std::vector<std::string>, bfuncs);
  // file names of outputs: <output_dir>/<basename>.… (basename = the YAML
  // file's name without its extension)
  SNAP_ARG(std::string, basename) = "";
  SNAP_ARG(std::string, output_dir) = ".";
  SNAP_ARG(std::vector<OutputOptions>, outputs);
};
using MeshBlockOptions = std::shared_ptr<MeshBlockOptionsImpl>;

// C-ABI configuration of a block
snapx_config make_config(CoordinateOptions const& coord, HydroOptions const& hydro,
                         std::vector<std::string> const& bfuncs);

// ---------------------------------------------------------------- modules

// IdealGasImpl (src/eos/ideal_gas.cpp:30-109)
class IdealGasImpl {
 public:
  IdealGasImpl(EquationOfStateOptions options, CoordinateOptions coord);
  EquationOfStateOptions options;
  int nvar() const { return 5; }
  torch::Tensor forward(torch::Tensor cons, c10::optional<torch::Tensor> out = c10::nullopt);
  torch::Tensor compute(std::string ab, std::vector<torch::Tensor> const& args);

 private:
  snapx_config cfg_;
};
using EquationOfState = std::shared_ptr<IdealGasImpl>;
using IdealGas = std::shared_ptr<IdealGasImpl>;

// ReconstructImpl (src/recon/reconstruct.hpp:60-120)
class ReconstructImpl {
 public:
  ReconstructImpl(ReconstructOptions options, EquationOfStateOptions eos = nullptr,
                  bool has_parent = true, int nghost = 3);
  ReconstructOptions options;
  int stencils() const;
  // w -> [wl, wr]; dim is the tensor axis (3 = x1, 2 = x2, 1 = x3)
  torch::Tensor forward(torch::Tensor w, int dim);

 private:
  EquationOfStateOptions eos_;
  bool has_parent_;
  int nghost_;
};
using Reconstruct = std::shared_ptr<ReconstructImpl>;

// RiemannSolverImpl and the LMARS / HLLC / Roe solvers
// (src/riemann/riemann_solver.hpp:40-170)
class RiemannSolverImpl {
 public:
  RiemannSolverImpl(RiemannSolverOptions options, double gammad, std::string type);
  RiemannSolverOptions options;
  const std::string& type() const { return type_; }
  torch::Tensor forward(torch::Tensor wl, torch::Tensor wr, int dim, torch::Tensor flx,
                        torch::Tensor face_pressure = torch::Tensor());

 private:
  double gammad_;
  std::string type_;
};
using RiemannSolver = std::shared_ptr<RiemannSolverImpl>;
RiemannSolver create_riemann_solver(RiemannSolverOptions const& opts, double gammad,
                                    std::string const& type = "");

// CoordinateImpl, Cartesian (src/coord/coordinate.cpp:194-434)
class CoordinateImpl {
 public:
  explicit CoordinateImpl(CoordinateOptions options);
  CoordinateOptions options;
  int il() const { return options->nx1() > 1 ? options->nghost() : 0; }
  int iu() const { return options->nx1() > 1 ? options->nghost() + options->nx1() - 1 : 0; }
  int jl() const { return options->nx2() > 1 ? options->nghost() : 0; }
  int ju() const { return options->nx2() > 1 ? options->nghost() + options->nx2() - 1 : 0; }
  int kl() const { return options->nx3() > 1 ? options->nghost() : 0; }
  int ku() const { return options->nx3() > 1 ? options->nghost() + options->nx3() - 1 : 0; }
  // FP64, host: faces, centres (volume-weighted on spherical-polar blocks)
  // and face widths
  torch::Tensor x1f, x2f, x3f, x1v, x2v, x3v, dx1f, dx2f, dx3f;
  torch::Tensor divergence(torch::Tensor flux1, torch::Tensor flux2, torch::Tensor flux3) const;
  torch::Tensor forward(torch::Tensor prim, torch::Tensor flux1, torch::Tensor flux2,
                        torch::Tensor flux3, torch::Tensor face_pressure1 = torch::Tensor());
};
using Coordinate = std::shared_ptr<CoordinateImpl>;

// ImplicitHydroImpl (src/implicit/implicit_hydro.cpp:50-189)
class ImplicitHydroImpl {
 public:
  ImplicitHydroImpl(ImplicitOptions options, snapx_config cfg);
  ImplicitOptions options;
  // modifies du in place and returns the correction du_new - du_old
  torch::Tensor forward(torch::Tensor du, torch::Tensor w, torch::Tensor gamma, double dt);

 private:
  snapx_config cfg_;
  torch::Tensor ws_;
};
using ImplicitHydro = std::shared_ptr<ImplicitHydroImpl>;

// HydroImpl (src/hydro/hydro.hpp:74-160)
class HydroImpl {
 public:
  HydroImpl(HydroOptions options, CoordinateOptions coord,
            std::vector<std::string> bfuncs = {});
  HydroOptions options;
  EquationOfState peos;
  Reconstruct precon1, precon23;
  RiemannSolver priemann;
  ImplicitHydro picorr;
  Coordinate pcoord;
  const snapx_config& config() const { return cfg_; }

  double max_time_step(torch::Tensor hydro_w) const;
  // dt, conserved state u, other["hydro_w"] (written in place) -> du
  torch::Tensor forward(double dt, torch::Tensor u, Variables const& other);
  torch::Tensor flux1() const { return F1_; }
  torch::Tensor flux2() const { return F2_; }
  torch::Tensor flux3() const { return F3_; }

 private:
  snapx_config cfg_;
  torch::Tensor F1_, F2_, F3_, ws_;
  void ensure_buffers(torch::Tensor const& u);
};
using Hydro = std::shared_ptr<HydroImpl>;

// MeshBlockImpl, stage driver of one block (src/mesh/meshblock.cpp:412-786)
// with SSP-RK3 and the external boundaries; "fused" selects the FP32 fused
// stage (hydro_u must then be float32).
class MeshBlockImpl {
 public:
  explicit MeshBlockImpl(MeshBlockOptions options);
  MeshBlockOptions options;
  Hydro phydro;
  int nstages() const { return 3; }
  // hydro_w (primitive) -> BCs, hydro_u = W->U, BCs (meshblock.cpp:412-510);
  // returns the start time 0.  With restart_file, the variables, cycle and
  // output schedule come from that file instead (meshblock.cpp:1030-1094) and
  // the file's time is returned.
  double initialize(Variables& vars, char const* restart_file = nullptr);
  // meshblock.cpp:813-830: every output whose next time has come (all of them
  // when final_write) is written; its next time advances by its dt
  void make_outputs(Variables const& vars, double current_time, bool final_write = false);
  // the final outputs (meshblock.cpp:909-911)
  void finalize(Variables const& vars, double time);
  std::vector<std::shared_ptr<OutputType>> output_types;
  double max_time_step(Variables const& vars);
  // one RK stage on vars["hydro_u"] (meshblock.cpp:577-786)
  void forward(Variables& vars, double dt, int stage);
  // full SSP-RK3 step, plain path (dtype of hydro_u) or the FP32 fused path
  void step(Variables& vars, double dt, bool fused = false);
  void apply_bc(torch::Tensor var, int var_type);
  // MeshBlockImpl::check_redo (meshblock.cpp:974-1028): 0 = good (redo count
  // reset), 1 = negative density or pressure, hydro_u restored to the step's
  // input and the next max_time_step halved once more, -1 = max_redo exceeded
  int check_redo(Variables& vars);
  int current_redo = 0;
  int cycle = 0;

 private:
  double init_from_restart(Variables& vars, std::string const& fname);
  // RestartOutput::write_output_file (src/output/restart.cpp:25-86)
  void write_restart(OutputType const& out, Variables const& vars, double time,
                     bool final_write);
  snapx_config cfg_;
  torch::Tensor u0_, u1_, fused_ws_, flag_;
};
using MeshBlock = std::shared_ptr<MeshBlockImpl>;

}  // namespace snap
