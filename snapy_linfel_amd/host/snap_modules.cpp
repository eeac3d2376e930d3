#include "snap_modules.hpp"

#include <c10/hip/HIPStream.h>

#include <cmath>
#include <filesystem>
#include <iostream>

#include "restart.hpp"

namespace snap {

namespace {

void* stream_of(torch::Tensor const& t) {
  return reinterpret_cast<void*>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

void check(int status, const char* what) {
  TORCH_CHECK(status == 0, "[snapx] ", what, " failed: ", snapx_last_error());
}

int dtype_code(torch::Tensor const& t) {
  if (t.scalar_type() == torch::kFloat64) return SNAPX_F64;
  TORCH_CHECK(t.scalar_type() == torch::kFloat32, "unsupported dtype ", t.scalar_type());
  return SNAPX_F32;
}

void require_device(torch::Tensor const& t, const char* name) {
  TORCH_CHECK(t.defined(), name, " is undefined");
  TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor (the MI355X path has no CPU fallback)");
  TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

int interp_code(std::string const& type) {
  if (type == "dc") return SNAPX_INTERP_DC;
  if (type == "plm") return SNAPX_INTERP_PLM;
  if (type == "cp3") return SNAPX_INTERP_CP3;
  if (type == "cp5") return SNAPX_INTERP_CP5;
  if (type == "weno3") return SNAPX_INTERP_WENO3;
  if (type == "weno5") return SNAPX_INTERP_WENO5;
  TORCH_CHECK(type != "ppm", "Interp: ppm is not supported (dc, plm, cp3, cp5, weno3, weno5)");
  TORCH_CHECK(false, "Interp: unknown type " + type);
  return -1;
}

int riemann_code(std::string const& type) {
  if (type == "roe") return SNAPX_RIEMANN_ROE;
  if (type == "lmars") return SNAPX_RIEMANN_LMARS;
  if (type == "hllc") return SNAPX_RIEMANN_HLLC;
  TORCH_CHECK(false, "RiemannSolver: unknown type " + type);
  return -1;
}

int bc_code(std::string const& name) {
  if (name.empty()) return SNAPX_BC_NONE;
  if (name == "reflecting") return SNAPX_BC_REFLECTING;
  if (name == "outflow") return SNAPX_BC_OUTFLOW;
  if (name == "periodic") return SNAPX_BC_PERIODIC;
  if (name == "custom") return SNAPX_BC_CUSTOM;
  TORCH_CHECK(false, "Boundary function '", name, "' is not defined.");
  return -1;
}

}  // namespace

// ================================================================ options

EquationOfStateOptions EquationOfStateOptionsImpl::from_yaml(std::string const& filename,
                                                             bool verbose) {
  return from_yaml(yaml::LoadFile(filename), verbose);
}

EquationOfStateOptions EquationOfStateOptionsImpl::from_yaml(yaml::Node const& config,
                                                             bool verbose) {
  auto op = create();
  if (!config["dynamics"]) return op;
  if (!config["dynamics"]["equation-of-state"]) return op;
  auto node = config["dynamics"]["equation-of-state"];
  op->verbose() = node["verbose"].as<bool>(verbose);
  op->type() = node["type"].as<std::string>("moist-mixture");
  op->gammad() = node["gammad"].as<double>(1.4);
  op->weight() = node["weight"].as<double>(29.e-3);
  op->density_floor() = node["density-floor"].as<double>(1.e-6);
  op->pressure_floor() = node["pressure-floor"].as<double>(1.e-3);
  op->temperature_floor() = node["temperature-floor"].as<double>(20.);
  op->limiter() = node["limiter"].as<bool>(false);
  return op;
}

ReconstructOptions ReconstructOptionsImpl::from_yaml(std::string const& filename,
                                                     std::string const& section) {
  auto config = yaml::LoadFile(filename);
  auto op = create();
  if (!config["dynamics"]) return op;
  if (!config["dynamics"]["reconstruct"]) return op;
  return from_yaml(config["dynamics"]["reconstruct"], section);
}

ReconstructOptions ReconstructOptionsImpl::from_yaml(yaml::Node const& node,
                                                     std::string const& section) {
  auto op = create();
  if (!node[section]) return op;
  op->shock() = node[section]["shock"].as<bool>(false);
  op->interp() = InterpOptionsImpl::create();
  op->interp()->type() = node[section]["type"].as<std::string>("dc");
  op->interp()->scale() = node[section]["scale"].as<bool>(false);
  return op;
}

RiemannSolverOptions RiemannSolverOptionsImpl::from_yaml(std::string const& filename,
                                                         std::string const& section) {
  auto op = create();
  auto config = yaml::LoadFile(filename);
  if (!config[section]) return op;
  if (!config[section]["riemann-solver"]) return op;
  // the reference reads the "dynamics" section whatever section names
  // (src/riemann/riemann_solver.cpp:15-24)
  return from_yaml(config["dynamics"]["riemann-solver"]);
}

RiemannSolverOptions RiemannSolverOptionsImpl::from_yaml(yaml::Node const& node) {
  auto op = create();
  op->type() = node["type"].as<std::string>("roe");
  op->dir() = node["dir"].as<std::string>("omni");
  op->type_x1() = node["type-x1"].as<std::string>("");
  op->type_x2() = node["type-x2"].as<std::string>("");
  op->type_x3() = node["type-x3"].as<std::string>("");
  return op;
}

std::string RiemannSolverOptionsImpl::type_of(int axis) const {
  const std::string& t = axis == 1 ? type_x1() : (axis == 2 ? type_x2() : type_x3());
  return t.empty() ? type() : t;
}

ConstGravityOptions ConstGravityOptionsImpl::from_yaml(yaml::Node const& forcing) {
  if (!forcing["const-gravity"]) return nullptr;
  auto node = forcing["const-gravity"];
  auto op = create();
  op->grav1() = node["grav1"].as<double>(0.);
  op->grav2() = node["grav2"].as<double>(0.);
  op->grav3() = node["grav3"].as<double>(0.);
  op->non_hydrostatic() = node["non-hydrostatic"].as<double>(1.);
  TORCH_CHECK(op->non_hydrostatic() >= 0. && op->non_hydrostatic() <= 1.);
  return op;
}

ImplicitOptions ImplicitOptionsImpl::from_yaml(std::string const& filename, bool) {
  auto config = yaml::LoadFile(filename);
  if (!config["integration"]) return nullptr;
  if (!config["integration"]["implicit-scheme"]) return nullptr;
  return from_yaml(config["integration"]["implicit-scheme"]);
}

ImplicitOptions ImplicitOptionsImpl::from_yaml(yaml::Node const& node) {
  auto op = create();
  op->scheme(node.as<int>());
  return op;
}

std::string ImplicitOptionsImpl::type() const {
  switch (scheme()) {
    case 0:
      return "none";
    case 1:
      return "vic-partial";
    case 9:
      return "vic-full";
    default:
      TORCH_CHECK(false, "Unsupported implicit scheme");
  }
  return "";
}

CoordinateOptions CoordinateOptionsImpl::from_yaml(std::string const& filename) {
  return from_yaml(yaml::LoadFile(filename));
}

CoordinateOptions CoordinateOptionsImpl::from_yaml(yaml::Node const& config) {
  auto op = create();
  auto node = config["geometry"];
  if (!node) return op;
  op->type(node["type"].as<std::string>("cartesian"));
  TORCH_CHECK(op->type() == "cartesian" || op->type() == "spherical-polar" ||
                  op->type() == "gnomonic-equiangle",
              "geometry.type '", op->type(),
              "' is not supported (cartesian, spherical-polar and gnomonic-equiangle only)");
  if (node["bounds"]) {
    // src/coord/coordinate.cpp:72-101: x2/x3 bounds may be given in units of pi
    auto b = node["bounds"];
    auto bound = [&](const char* key, double fallback) {
      std::string pi_key = std::string(key) + "_pi";
      if (b[pi_key].defined()) return b[pi_key].as<double>() * M_PI;
      return b[key].as<double>(fallback);
    };
    op->x1min() = b["x1min"].as<double>(0.0);
    op->x2min() = bound("x2min", 0.0);
    op->x3min() = bound("x3min", 0.0);
    op->x1max() = b["x1max"].as<double>(1.0);
    op->x2max() = bound("x2max", 1.0);
    op->x3max() = bound("x3max", 1.0);
  }
  if (!node["cells"]) return op;
  op->nx1() = node["cells"]["nx1"].as<int>(1);
  op->nx2() = node["cells"]["nx2"].as<int>(1);
  op->nx3() = node["cells"]["nx3"].as<int>(1);
  op->nghost() = node["cells"]["nghost"].as<int>(1);
  TORCH_CHECK(!(op->nx1() > 1 && op->nx1() < op->nghost()),
              "Number of x1 grids must be greater than the ghost zone size");
  TORCH_CHECK(!(op->nx2() > 1 && op->nx2() < op->nghost()),
              "Number of x2 grids must be greater than the ghost zone size");
  TORCH_CHECK(!(op->nx3() > 1 && op->nx3() < op->nghost()),
              "Number of x3 grids must be greater than the ghost zone size");
  return op;
}

HydroOptions HydroOptionsImpl::from_yaml(std::string const& filename, bool verbose) {
  auto config = yaml::LoadFile(filename);
  auto op = create();
  op->eos() = EquationOfStateOptionsImpl::from_yaml(config, verbose);
  op->recon1() = ReconstructOptionsImpl::from_yaml(filename, "vertical");
  op->recon23() = ReconstructOptionsImpl::from_yaml(filename, "horizontal");
  op->riemann() = RiemannSolverOptionsImpl::from_yaml(filename, "dynamics");
  op->icorr() = ImplicitOptionsImpl::from_yaml(filename);
  auto dyn = config["dynamics"];
  if (dyn) {
    op->verbose() = dyn["verbose"].as<bool>(verbose);
    op->disable_flux_x1() = dyn["disable-flux-x1"].as<bool>(false);
    op->disable_flux_x2() = dyn["disable-flux-x2"].as<bool>(false);
    op->disable_flux_x3() = dyn["disable-flux-x3"].as<bool>(false);
  }
  auto forcing = config["forcing"];
  if (!forcing) return op;
  op->grav() = ConstGravityOptionsImpl::from_yaml(forcing);
  if (op->grav()) {
    if (op->disable_flux_x1()) op->grav()->grav1(0.);
    if (op->disable_flux_x2()) op->grav()->grav2(0.);
    if (op->disable_flux_x3()) op->grav()->grav3(0.);
  }
  return op;
}

OutputOptions OutputOptionsImpl::from_yaml(yaml::Node const& node, int fid) {
  auto op = create();
  op->fid() = fid;
  op->dt() = node["dt"].as<double>(0.);
  TORCH_CHECK(node["type"].defined(), "OutputOptions::from_yaml: output file type must be specified");
  op->file_type() = node["type"].as<std::string>();
  auto vars = node["variables"];
  if (vars.defined())
    for (size_t i = 0; i < vars.size(); ++i) op->variables().push_back(vars[i].as<std::string>());
  op->combine() = node["combine"].as<bool>(true);
  return op;
}

MeshBlockOptions MeshBlockOptionsImpl::from_yaml(std::string const& filename, bool verbose) {
  auto config = yaml::LoadFile(filename);
  auto op = create();
  // the YAML file's name without directory and extension
  // (src/mesh/meshblock_options.cpp:19-25)
  std::string stem = std::filesystem::path(filename).filename().string();
  op->basename() = stem.substr(0, stem.find_last_of('.'));
  auto outs = config["outputs"];
  if (outs.defined())
    for (size_t i = 0; i < outs.size(); ++i)
      op->outputs().push_back(OutputOptionsImpl::from_yaml(outs[i], (int)i));
  op->coord() = CoordinateOptionsImpl::from_yaml(config);
  op->hydro() = HydroOptionsImpl::from_yaml(filename, verbose);
  if (config["integration"]) {
    auto in = config["integration"];
    op->integrator() = in["type"].as<std::string>("rk3");
    op->cfl() = in["cfl"].as<double>(0.9);
    op->nlim() = in["nlim"].as<int>(-1);
    op->tlim() = in["tlim"].as<double>(1.e9);
    op->max_redo() = in["max-redo"].as<int>(5);
  }
  TORCH_CHECK(op->integrator() == "rk3",
              "integration.type '", op->integrator(), "' is not supported (rk3 only)");
  auto& bf = op->bfuncs();
  bf.assign(6, "");
  auto ext = config["boundary-condition"]["external"];
  if (ext) {
    // src/mesh/meshblock_options.cpp:87-191 (defaults "reflecting")
    const char* keys[6] = {"x1-inner", "x1-outer", "x2-inner", "x2-outer", "x3-inner", "x3-outer"};
    int nx[3] = {op->coord()->nx1(), op->coord()->nx2(), op->coord()->nx3()};
    for (int i = 0; i < 6; ++i)
      if (nx[i / 2] > 1) bf[i] = ext[keys[i]].as<std::string>("reflecting");
  }
  return op;
}

snapx_config make_config(CoordinateOptions const& coord, HydroOptions const& hydro,
                         std::vector<std::string> const& bfuncs) {
  snapx_config c{};
  c.nx1 = coord->nx1();
  c.nx2 = coord->nx2();
  c.nx3 = coord->nx3();
  c.nghost = coord->nghost();
  c.x1min = coord->x1min(), c.x1max = coord->x1max();
  c.x2min = coord->x2min(), c.x2max = coord->x2max();
  c.x3min = coord->x3min(), c.x3max = coord->x3max();
  c.coord = coord->type() == "spherical-polar"      ? SNAPX_COORD_SPHERICAL_POLAR
            : coord->type() == "gnomonic-equiangle" ? SNAPX_COORD_GNOMONIC_EQUIANGLE
                                                    : SNAPX_COORD_CARTESIAN;
  auto eos = hydro->eos() ? hydro->eos() : EquationOfStateOptionsImpl::create();
  TORCH_CHECK(eos->type() == "ideal-gas", "EquationOfState: type '", eos->type(),
              "' is not supported (ideal-gas only)");
  c.gammad = eos->gammad();
  c.weight = eos->weight();
  c.limiter = eos->limiter();
  c.density_floor = eos->density_floor();
  c.pressure_floor = eos->pressure_floor();
  c.temperature_floor = eos->temperature_floor();
  ReconstructOptions rs[2] = {hydro->recon1() ? hydro->recon1() : ReconstructOptionsImpl::create(),
                              hydro->recon23() ? hydro->recon23()
                                               : ReconstructOptionsImpl::create()};
  for (int r = 0; r < 2; ++r) {
    c.recon_type[r] = interp_code(rs[r]->interp()->type());
    c.recon_scale[r] = rs[r]->interp()->scale();
    c.recon_shock[r] = rs[r]->shock();
  }
  auto rv = hydro->riemann() ? hydro->riemann() : RiemannSolverOptionsImpl::create();
  for (int d = 0; d < 3; ++d) c.riemann[d] = riemann_code(rv->type_of(d + 1));
  auto g = hydro->grav();
  c.grav1 = g ? g->grav1() : 0.;
  c.grav2 = g ? g->grav2() : 0.;
  c.grav3 = g ? g->grav3() : 0.;
  c.non_hydrostatic = g ? g->non_hydrostatic() : 1.;
  c.implicit_scheme = hydro->icorr() ? hydro->icorr()->scheme() : 0;
  TORCH_CHECK(c.implicit_scheme == 0 || g, "[ImplicitHydro] forcing does not have const-gravity");
  c.disable_flux[0] = hydro->disable_flux_x1();
  c.disable_flux[1] = hydro->disable_flux_x2();
  c.disable_flux[2] = hydro->disable_flux_x3();
  for (int i = 0; i < 6; ++i) c.bc[i] = i < (int)bfuncs.size() ? bc_code(bfuncs[i]) : -1;
  return c;
}

// ================================================================ modules

IdealGasImpl::IdealGasImpl(EquationOfStateOptions options_, CoordinateOptions coord)
    : options(std::move(options_)) {
  auto h = HydroOptionsImpl::create();
  h->eos() = options;
  cfg_ = make_config(coord ? coord : CoordinateOptionsImpl::create(), h, {});
}

torch::Tensor IdealGasImpl::forward(torch::Tensor cons, c10::optional<torch::Tensor> out) {
  return compute("U->W", {cons, out.has_value() ? *out : torch::empty_like(cons)});
}

torch::Tensor IdealGasImpl::compute(std::string ab, std::vector<torch::Tensor> const& args) {
  TORCH_CHECK(!args.empty(), "compute: no arguments");
  if (ab == "U->W" || ab == "W->U") {
    auto in = args[0];
    auto out = args.size() > 1 ? args[1] : torch::empty_like(in);
    require_device(in, "input");
    require_device(out, "output");
    TORCH_CHECK(in.size(0) == 5 && in.sizes() == out.sizes(), "compute: shape mismatch");
    // the EOS kernels only need the cell count: present the block as 1 x 1 x n
    snapx_config c = cfg_;
    c.nx1 = in[0].numel(), c.nx2 = 1, c.nx3 = 1, c.nghost = 1;
    c.nx1 -= 2;
    for (int i = 0; i < 6; ++i) c.bc[i] = -1;
    if (c.nx1 < 1) c.nx1 = 1;
    TORCH_CHECK(in[0].numel() >= 3, "compute: at least 3 cells");
    int st = ab == "U->W" ? snapx_ideal_gas_cons2prim(&c, in.data_ptr(), out.data_ptr(),
                                                       dtype_code(in), stream_of(in))
                          : snapx_ideal_gas_prim2cons(&c, in.data_ptr(), out.data_ptr(),
                                                      dtype_code(in), stream_of(in));
    check(st, ab.c_str());
    return out;
  }
  auto w = args[0];
  if (ab == "W->I") return w[IPR] / (options->gammad() - 1);
  if (ab == "W->T") return w[IPR] / (w[IDN] * (8.31446261815324 / options->weight()));
  if (ab == "W->A") return options->gammad() * torch::ones_like(w[IDN]);
  if (ab == "WA->L") return torch::sqrt(args[1] * w[IPR] / w[IDN]);
  TORCH_CHECK(false, "Unknown abbreviation: ", ab);
  return {};
}

ReconstructImpl::ReconstructImpl(ReconstructOptions options_, EquationOfStateOptions eos,
                                 bool has_parent, int nghost)
    : options(std::move(options_)), eos_(std::move(eos)), has_parent_(has_parent),
      nghost_(nghost) {
  if (!options->interp()) options->interp() = InterpOptionsImpl::create();
  interp_code(options->interp()->type());
}

int ReconstructImpl::stencils() const {
  auto t = options->interp()->type();
  return t == "dc" ? 1 : (t == "plm" || t == "cp3" || t == "weno3") ? 3 : 5;
}

torch::Tensor ReconstructImpl::forward(torch::Tensor w, int dim) {
  require_device(w, "w");
  TORCH_CHECK(w.dim() == 4 && w.size(0) == 5, "Reconstruct: w must be [5, nc3, nc2, nc1]");
  TORCH_CHECK(dim >= 1 && dim <= 3, "Reconstruct: dim must be 1, 2 or 3");
  auto vec = w.sizes().vec();
  vec.insert(vec.begin(), 2);
  auto result = torch::empty(vec, w.options());
  int n = w.size(dim);
  int ng = stencils() / 2 + 1;
  TORCH_CHECK(ng <= n - ng, "il > iu");
  bool limiter = eos_ ? eos_->limiter() : options->limiter();
  double df = eos_ ? eos_->density_floor() : options->density_floor();
  double pf = eos_ ? eos_->pressure_floor() : options->pressure_floor();
  check(snapx_reconstruct(w.data_ptr(), result.data_ptr(), 5, w.size(1), w.size(2), w.size(3),
                          dim, interp_code(options->interp()->type()),
                          options->interp()->scale(), options->shock(), has_parent_, limiter,
                          df, pf, 0, nghost_, dtype_code(w), stream_of(w)),
        "reconstruct");
  return result;
}

RiemannSolverImpl::RiemannSolverImpl(RiemannSolverOptions options_, double gammad,
                                     std::string type)
    : options(std::move(options_)), gammad_(gammad), type_(std::move(type)) {
  riemann_code(type_);
}

torch::Tensor RiemannSolverImpl::forward(torch::Tensor wl, torch::Tensor wr, int dim,
                                         torch::Tensor flx, torch::Tensor face_pressure) {
  require_device(wl, "wl");
  require_device(wr, "wr");
  require_device(flx, "flx");
  TORCH_CHECK(wl.sizes() == wr.sizes() && wl.sizes() == flx.sizes(), "RiemannSolver: shapes");
  TORCH_CHECK(wl.size(0) == 5, "RiemannSolver: dry ideal gas (nvar = 5)");
  if (face_pressure.defined()) require_device(face_pressure, "face_pressure");
  check(snapx_riemann(riemann_code(type_), wl.data_ptr(), wr.data_ptr(), flx.data_ptr(),
                      face_pressure.defined() ? face_pressure.data_ptr() : nullptr, 5,
                      wl.size(1), wl.size(2), wl.size(3), dim, gammad_, dtype_code(wl),
                      stream_of(wl)),
        "riemann");
  return flx;
}

RiemannSolver create_riemann_solver(RiemannSolverOptions const& opts, double gammad,
                                    std::string const& type) {
  return std::make_shared<RiemannSolverImpl>(opts, gammad, type.empty() ? opts->type() : type);
}

CoordinateImpl::CoordinateImpl(CoordinateOptions options_) : options(std::move(options_)) {
  auto const& op = options;
  const bool sph = op->type() == "spherical-polar";
  auto axis = [&](int nx, double lo, double hi, int nc, torch::Tensor& xf, torch::Tensor& xv,
                  torch::Tensor& dxf) {
    double dx = (hi - lo) / nx;
    double xmin = nx > 1 ? lo - op->nghost() * dx : lo;
    double xmax = nx > 1 ? hi + op->nghost() * dx : hi;
    xf = torch::linspace(xmin, xmax, nc + 1, torch::kFloat64);
    xv = 0.5 * (xf.slice(0, 0, nc) + xf.slice(0, 1, nc + 1));
    dxf = sph ? xf.slice(0, 1, nc + 1) - xf.slice(0, 0, nc) : torch::ones(nc, torch::kFloat64) * dx;
  };
  axis(op->nx1(), op->x1min(), op->x1max(), op->nc1(), x1f, x1v, dx1f);
  axis(op->nx2(), op->x2min(), op->x2max(), op->nc2(), x2f, x2v, dx2f);
  axis(op->nx3(), op->x3min(), op->x3max(), op->nc3(), x3f, x3v, dx3f);
  if (sph) {
    // volume-weighted centres (src/coord/spherical_polar.cpp:17-28, 60-67)
    int nc1 = op->nc1(), nc2 = op->nc2();
    auto rm = x1f.slice(0, 0, nc1), rp = x1f.slice(0, 1, nc1 + 1);
    x1v = 0.75 * (rp.pow(4) - rm.pow(4)) / (rp.pow(3) - rm.pow(3));
    if (nc2 > 1) {
      auto tm = x2f.slice(0, 0, nc2), tp = x2f.slice(0, 1, nc2 + 1);
      x2v = ((tp.sin() - tp * tp.cos()) - (tm.sin() - tm * tm.cos())) / (tm.cos() - tp.cos());
    }
  }
}

torch::Tensor CoordinateImpl::divergence(torch::Tensor f1, torch::Tensor f2,
                                         torch::Tensor f3) const {
  torch::Tensor ref = f1.defined() ? f1 : (f2.defined() ? f2 : f3);
  TORCH_CHECK(ref.defined(), "At least one flux tensor must be defined");
  require_device(ref, "flux");
  auto h = HydroOptionsImpl::create();
  h->eos() = EquationOfStateOptionsImpl::create();
  snapx_config c = make_config(options, h, {});
  auto div = torch::empty_like(ref);
  check(snapx_divergence(&c, f1.defined() ? f1.data_ptr() : nullptr,
                         f2.defined() ? f2.data_ptr() : nullptr,
                         f3.defined() ? f3.data_ptr() : nullptr, div.data_ptr(), dtype_code(ref),
                         stream_of(ref)),
        "divergence");
  return div;
}

torch::Tensor CoordinateImpl::forward(torch::Tensor prim, torch::Tensor f1, torch::Tensor f2,
                                      torch::Tensor f3, torch::Tensor face_pressure1) {
  torch::Tensor ref = f1.defined() ? f1 : (f2.defined() ? f2 : f3);
  TORCH_CHECK(ref.defined(), "At least one flux tensor must be defined");
  require_device(ref, "flux");
  auto h = HydroOptionsImpl::create();
  h->eos() = EquationOfStateOptionsImpl::create();
  snapx_config c = make_config(options, h, {});
  if (c.coord == SNAPX_COORD_CARTESIAN) return divergence(f1, f2, f3);
  require_device(prim, "prim");
  auto div = torch::empty_like(ref);
  auto ptr = [](torch::Tensor const& t) { return t.defined() ? t.data_ptr() : nullptr; };
  check(snapx_coord_forward(&c, prim.data_ptr(), ptr(f1), ptr(f2), ptr(f3), ptr(face_pressure1),
                            div.data_ptr(), dtype_code(ref), stream_of(ref)),
        "coord_forward");
  return div;
}

ImplicitHydroImpl::ImplicitHydroImpl(ImplicitOptions options_, snapx_config cfg)
    : options(std::move(options_)), cfg_(cfg) {
  cfg_.implicit_scheme = options->scheme();
  options->type();  // validates the scheme
}

torch::Tensor ImplicitHydroImpl::forward(torch::Tensor du, torch::Tensor w, torch::Tensor,
                                         double dt) {
  if (options->scheme() == 0) return torch::zeros_like(du);
  require_device(du, "du");
  require_device(w, "w");
  auto du0 = du.clone();
  size_t n = snapx_implicit_workspace_bytes(&cfg_);
  if (!ws_.defined() || (size_t)ws_.numel() < n || ws_.device() != du.device())
    ws_ = torch::empty({(int64_t)std::max<size_t>(n, 1)}, du.options().dtype(torch::kUInt8));
  check(snapx_implicit_correction(&cfg_, dt, du.data_ptr(), w.data_ptr(), ws_.data_ptr(),
                                  ws_.numel(), dtype_code(du), stream_of(du)),
        "implicit_correction");
  return du - du0;
}

HydroImpl::HydroImpl(HydroOptions options_, CoordinateOptions coord,
                     std::vector<std::string> bfuncs)
    : options(std::move(options_)) {
  if (!options->eos()) options->eos() = EquationOfStateOptionsImpl::create();
  if (!options->recon1()) options->recon1() = ReconstructOptionsImpl::create();
  if (!options->recon23()) options->recon23() = ReconstructOptionsImpl::create();
  if (!options->riemann()) options->riemann() = RiemannSolverOptionsImpl::create();
  cfg_ = make_config(coord, options, bfuncs);
  TORCH_CHECK(snapx_hydro_workspace_bytes(&cfg_, SNAPX_F64) > 0, "invalid configuration: ",
              snapx_last_error());
  peos = std::make_shared<IdealGasImpl>(options->eos(), coord);
  precon1 = std::make_shared<ReconstructImpl>(options->recon1(), options->eos(), true,
                                              coord->nghost());
  precon23 = std::make_shared<ReconstructImpl>(options->recon23(), options->eos(), true,
                                               coord->nghost());
  priemann = create_riemann_solver(options->riemann(), options->eos()->gammad(),
                                   options->riemann()->type_of(1));
  if (options->icorr()) picorr = std::make_shared<ImplicitHydroImpl>(options->icorr(), cfg_);
  pcoord = std::make_shared<CoordinateImpl>(coord);
}

void HydroImpl::ensure_buffers(torch::Tensor const& u) {
  auto shape = u.sizes();
  if (!F1_.defined() || F1_.sizes() != shape || F1_.device() != u.device() ||
      F1_.scalar_type() != u.scalar_type()) {
    F1_ = torch::zeros(shape, u.options());
    F2_ = torch::zeros(shape, u.options());
    F3_ = torch::zeros(shape, u.options());
    size_t n = snapx_hydro_workspace_bytes(&cfg_, dtype_code(u));
    ws_ = torch::empty({(int64_t)std::max<size_t>(n, 1)}, u.options().dtype(torch::kUInt8));
  }
}

double HydroImpl::max_time_step(torch::Tensor w) const {
  require_device(w, "hydro_w");
  auto out = torch::empty({1}, w.options().dtype(torch::kFloat64));
  check(snapx_max_time_step(&cfg_, w.data_ptr(), 1.0, out.data_ptr<double>(), dtype_code(w),
                            stream_of(w)),
        "max_time_step");
  return out.item<double>();
}

torch::Tensor HydroImpl::forward(double dt, torch::Tensor u, Variables const& other) {
  require_device(u, "hydro_u");
  TORCH_CHECK(other.count("hydro_w"), "Hydro::forward: other[\"hydro_w\"] is required");
  TORCH_CHECK(!other.count("solid"), "internal solid boundaries are not supported");
  auto w = other.at("hydro_w");
  require_device(w, "hydro_w");
  TORCH_CHECK(u.sizes() == w.sizes() && u.scalar_type() == w.scalar_type(),
              "hydro_u / hydro_w mismatch");
  TORCH_CHECK(u.size(0) == 5 && u.size(1) == cfg_.nx3 + (cfg_.nx3 > 1 ? 2 * cfg_.nghost : 0) &&
                  u.size(2) == cfg_.nx2 + (cfg_.nx2 > 1 ? 2 * cfg_.nghost : 0) &&
                  u.size(3) == cfg_.nx1 + (cfg_.nx1 > 1 ? 2 * cfg_.nghost : 0),
              "hydro_u shape does not match the geometry");
  ensure_buffers(u);
  auto du = torch::empty_like(u);
  check(snapx_hydro_forward(&cfg_, dt, u.data_ptr(), w.data_ptr(), du.data_ptr(),
                            u.size(3) > 1 ? F1_.data_ptr() : nullptr,
                            u.size(2) > 1 ? F2_.data_ptr() : nullptr,
                            u.size(1) > 1 ? F3_.data_ptr() : nullptr, ws_.data_ptr(),
                            ws_.numel(), dtype_code(u), stream_of(u)),
        "hydro_forward");
  return du;
}

MeshBlockImpl::MeshBlockImpl(MeshBlockOptions options_) : options(std::move(options_)) {
  phydro = std::make_shared<HydroImpl>(options->hydro(), options->coord(), options->bfuncs());
  cfg_ = phydro->config();
  // output types as in src/mesh/meshblock.cpp:128-175
  for (auto const& out_op : options->outputs()) {
    auto const& type = out_op->file_type();
    if (type == "netcdf" || type == "pnetcdf") {
      TORCH_WARN("MeshBlock: output type '", type, "' keeps its schedule but is not written ",
                 "(this library writes restart files only)");
    } else {
      TORCH_CHECK(type == "restart", "Output type '", type, "' is not implemented.");
    }
    auto ot = std::make_shared<OutputType>();
    ot->options = out_op;
    output_types.push_back(ot);
  }
}

void MeshBlockImpl::apply_bc(torch::Tensor var, int var_type) {
  require_device(var, "var");
  check(snapx_apply_bc(&cfg_, var.data_ptr(), var.size(0), var_type, dtype_code(var),
                       stream_of(var)),
        "apply_bc");
}

double MeshBlockImpl::initialize(Variables& vars, char const* restart_file) {
  if (restart_file != nullptr) return init_from_restart(vars, restart_file);
  TORCH_CHECK(vars.count("hydro_w"), "initialize: vars[\"hydro_w\"] is required");
  auto w = vars.at("hydro_w");
  apply_bc(w, kPrimitive);
  auto u = torch::empty_like(w);
  check(snapx_ideal_gas_prim2cons(&cfg_, w.data_ptr(), u.data_ptr(), dtype_code(w),
                                  stream_of(w)),
        "prim2cons");
  apply_bc(u, kConserved);
  vars["hydro_u"] = u;
  return 0.;
}

double MeshBlockImpl::init_from_restart(Variables& vars, std::string const& fname) {
  std::filesystem::path path(fname);
  if (!path.is_absolute() && !std::filesystem::exists(path))
    path = std::filesystem::path(options->output_dir()) / fname;
  auto data = load_restart(path.string(), /*block_rank=*/0);
  for (const char* key : {"hydro_u", "last_time", "last_cycle", "file_number", "next_time"})
    TORCH_CHECK(data.count(key), "Restart file is missing required variable: ", key);
  // resume at the saved cycle, so that nlim-limited runs take no extra step
  cycle = (int)data.at("last_cycle").item<int64_t>();
  const double time = data.at("last_time").item<double>();
  auto const& numbers = data.at("file_number");
  auto const& times = data.at("next_time");
  const int saved = std::min<int>((int)output_types.size(), (int)numbers.size(0));
  for (int n = 0; n < saved; ++n) {
    output_types[n]->file_number = (int)numbers[n].item<int64_t>();
    output_types[n]->next_time = times[n].item<double>();
  }
  // outputs added after the restart file was written start now
  for (size_t n = saved; n < output_types.size(); ++n) output_types[n]->next_time = time;
  for (auto& [name, t] : data) {
    if (name == "last_time" || name == "last_cycle" || name == "file_number" || name == "next_time")
      continue;
    vars[name] = t.to(torch::kCUDA);
  }
  return time;
}

void MeshBlockImpl::write_restart(OutputType const& out, Variables const& vars, double time,
                                  bool final_write) {
  Variables out_vars;
  const bool skip_scalar_r = vars.count("scalar_s") && vars.count("hydro_u");
  for (auto const& [name, t] : vars) {
    if (skip_scalar_r && name == "scalar_r") continue;
    if (t.defined()) out_vars[name] = t.to(torch::kCPU);
  }
  out_vars["last_time"] = torch::tensor({time}, torch::kFloat64);
  out_vars["last_cycle"] = torch::tensor({(int64_t)cycle}, torch::kInt64);
  std::vector<int64_t> numbers;
  std::vector<double> times;
  for (auto const& ot : output_types) {
    numbers.push_back(ot->file_number);
    times.push_back(ot->next_time);
  }
  out_vars["file_number"] = torch::tensor(numbers, torch::kInt64);
  out_vars["next_time"] = torch::tensor(times, torch::kFloat64);

  auto const& dir = options->output_dir();
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  TORCH_CHECK(!ec, "Failed to create output directory '", dir, "': ", ec.message());
  save_tensors(out_vars, restart_part_name(dir, options->basename(), /*rank=*/0,
                                           out.file_number, final_write));
  combine_restart_parts(dir, options->basename(), out.file_number, final_write);
}

void MeshBlockImpl::make_outputs(Variables const& vars, double current_time, bool final_write) {
  for (auto& ot : output_types) {
    const bool due = final_write || current_time >= ot->next_time;
    if (!due) continue;
    if (ot->options->file_type() == "restart") write_restart(*ot, vars, current_time, final_write);
    if (!final_write) {
      ot->next_time += ot->options->dt();
      ot->file_number += 1;
    }
  }
}

void MeshBlockImpl::finalize(Variables const& vars, double time) {
  make_outputs(vars, time, /*final_write=*/true);
}

double MeshBlockImpl::max_time_step(Variables const& vars) {
  auto w = vars.at("hydro_w");
  return std::pow(2., -current_redo) * options->cfl() * phydro->max_time_step(w);
}

int MeshBlockImpl::check_redo(Variables& vars) {
  auto u = vars.at("hydro_u");
  require_device(u, "hydro_u");
  if (!flag_.defined() || flag_.device() != u.device())
    flag_ = torch::empty({1}, u.options().dtype(torch::kInt32));
  check(snapx_check_state(&cfg_, u.data_ptr(), flag_.data_ptr<int>(), dtype_code(u),
                          stream_of(u)),
        "check_state");
  int verdict = flag_.item<int>();
  if (verdict & (SNAPX_REDO_DENSITY | SNAPX_REDO_PRESSURE)) {
    current_redo += 1;
    if (current_redo > options->max_redo()) return -1;
    TORCH_CHECK(u0_.defined() && u0_.sizes() == u.sizes() && u0_.dtype() == u.dtype(),
                "check_redo: no step input to restore (run a step first)");
    u.copy_(u0_);
    cycle -= 1;
    return 1;
  }
  current_redo = 0;
  return 0;
}

void MeshBlockImpl::forward(Variables& vars, double dt, int stage) {
  TORCH_CHECK(stage >= 0 && stage < 3, "Invalid stage: ", stage);
  auto u = vars.at("hydro_u");
  if (!vars.count("hydro_w")) vars["hydro_w"] = torch::empty_like(u);
  if (stage == 0) {
    u0_ = u.clone();
    u1_ = u.clone();
  }
  auto du = phydro->forward(dt, u, vars);
  static const double W[3][3] = {
      {0., 1., 1.}, {3. / 4., 1. / 4., 1. / 4.}, {1. / 3., 2. / 3., 2. / 3.}};
  check(snapx_rk_combine(u.data_ptr(), u0_.data_ptr(), u1_.data_ptr(), du.data_ptr(),
                         W[stage][0], W[stage][1], W[stage][2], u.numel(), dtype_code(u),
                         stream_of(u)),
        "rk_combine");
  if (cfg_.limiter) {
    auto scratch = torch::empty_like(u);
    check(snapx_ideal_gas_cons2prim(&cfg_, u.data_ptr(), scratch.data_ptr(), dtype_code(u),
                                    stream_of(u)),
          "conserved limiter");
  }
  apply_bc(u, kConserved);
  u1_.copy_(u);
}

void MeshBlockImpl::step(Variables& vars, double dt, bool fused) {
  if (!fused) {
    for (int s = 0; s < 3; ++s) forward(vars, dt, s);
    return;
  }
  auto u = vars.at("hydro_u");
  require_device(u, "hydro_u");
  TORCH_CHECK(u.scalar_type() == torch::kFloat32, "the fused path needs float32 hydro_u");
  size_t n = snapx_fused_workspace_bytes(&cfg_);
  TORCH_CHECK(n > 0, "fused path unavailable for this configuration: ", snapx_last_error());
  // step input kept for check_redo, as stage 0 of forward() keeps it
  if (u0_.defined() && u0_.sizes() == u.sizes() && u0_.dtype() == u.dtype())
    u0_.copy_(u);
  else
    u0_ = u.clone();
  if (!fused_ws_.defined() || (size_t)fused_ws_.numel() < n)
    fused_ws_ = torch::empty({(int64_t)n}, u.options().dtype(torch::kUInt8));
  check(snapx_fused_rk3_step(&cfg_, dt, u.data_ptr<float>(), fused_ws_.data_ptr(),
                             fused_ws_.numel(), stream_of(u)),
        "fused_rk3_step");
}

}  // namespace snap
