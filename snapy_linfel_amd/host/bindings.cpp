// Python bindings of the host modules, named like the reference's pybind
// module (python/csrc/snapy.cpp:32-61, python/csrc/py*.cpp).
#include <torch/extension.h>

#include "restart.hpp"
#include "snap_modules.hpp"

namespace py = pybind11;
using namespace snap;

#define OPT(cls, name) .def_property(#name, [](cls& o) { return o.name(); }, \
                                     [](cls& o, decltype(std::declval<cls&>().name()) v) { o.name() = v; })

PYBIND11_MODULE(_snapy, m) {
  m.doc() = "MI355X-native hot path of snapy: host modules over libsnapx";
  m.attr("kIDN") = (int)IDN;
  m.attr("kIV1") = (int)IVX;
  m.attr("kIV2") = (int)IVY;
  m.attr("kIV3") = (int)IVZ;
  m.attr("kIPR") = (int)IPR;
  m.attr("kICY") = (int)ICY;

  py::class_<EquationOfStateOptionsImpl, EquationOfStateOptions>(m, "EquationOfStateOptions")
      .def(py::init(&EquationOfStateOptionsImpl::create))
      .def_static("from_yaml", py::overload_cast<std::string const&, bool>(
                                   &EquationOfStateOptionsImpl::from_yaml),
                  py::arg("filename"), py::arg("verbose") = false)
      OPT(EquationOfStateOptionsImpl, type) OPT(EquationOfStateOptionsImpl, gammad)
      OPT(EquationOfStateOptionsImpl, weight) OPT(EquationOfStateOptionsImpl, density_floor)
      OPT(EquationOfStateOptionsImpl, pressure_floor)
      OPT(EquationOfStateOptionsImpl, temperature_floor) OPT(EquationOfStateOptionsImpl, limiter);

  py::class_<InterpOptionsImpl, InterpOptions>(m, "InterpOptions")
      .def(py::init(&InterpOptionsImpl::create))
      .def(py::init<std::string>())
      OPT(InterpOptionsImpl, type) OPT(InterpOptionsImpl, scale);

  py::class_<ReconstructOptionsImpl, ReconstructOptions>(m, "ReconstructOptions")
      .def(py::init(&ReconstructOptionsImpl::create))
      .def_static("from_yaml", py::overload_cast<std::string const&, std::string const&>(
                                   &ReconstructOptionsImpl::from_yaml))
      OPT(ReconstructOptionsImpl, shock) OPT(ReconstructOptionsImpl, interp)
      OPT(ReconstructOptionsImpl, limiter) OPT(ReconstructOptionsImpl, density_floor)
      OPT(ReconstructOptionsImpl, pressure_floor);

  py::class_<RiemannSolverOptionsImpl, RiemannSolverOptions>(m, "RiemannSolverOptions")
      .def(py::init(&RiemannSolverOptionsImpl::create))
      .def_static("from_yaml", py::overload_cast<std::string const&, std::string const&>(
                                   &RiemannSolverOptionsImpl::from_yaml),
                  py::arg("filename"), py::arg("section") = "dynamics")
      .def("type_of", &RiemannSolverOptionsImpl::type_of)
      OPT(RiemannSolverOptionsImpl, type) OPT(RiemannSolverOptionsImpl, dir)
      OPT(RiemannSolverOptionsImpl, type_x1) OPT(RiemannSolverOptionsImpl, type_x2)
      OPT(RiemannSolverOptionsImpl, type_x3);

  py::class_<ConstGravityOptionsImpl, ConstGravityOptions>(m, "ConstGravityOptions")
      .def(py::init(&ConstGravityOptionsImpl::create))
      OPT(ConstGravityOptionsImpl, grav1) OPT(ConstGravityOptionsImpl, grav2)
      OPT(ConstGravityOptionsImpl, grav3) OPT(ConstGravityOptionsImpl, non_hydrostatic);

  py::class_<ImplicitOptionsImpl, ImplicitOptions>(m, "ImplicitOptions")
      .def(py::init(&ImplicitOptionsImpl::create))
      .def_static("from_yaml", py::overload_cast<std::string const&, bool>(
                                   &ImplicitOptionsImpl::from_yaml),
                  py::arg("filename"), py::arg("verbose") = false)
      .def("size", &ImplicitOptionsImpl::size)
      .def("type", &ImplicitOptionsImpl::type)
      OPT(ImplicitOptionsImpl, scheme);

  py::class_<CoordinateOptionsImpl, CoordinateOptions>(m, "CoordinateOptions")
      .def(py::init(&CoordinateOptionsImpl::create))
      .def_static("from_yaml", py::overload_cast<std::string const&>(
                                   &CoordinateOptionsImpl::from_yaml))
      .def("nc1", &CoordinateOptionsImpl::nc1)
      .def("nc2", &CoordinateOptionsImpl::nc2)
      .def("nc3", &CoordinateOptionsImpl::nc3)
      OPT(CoordinateOptionsImpl, type) OPT(CoordinateOptionsImpl, x1min)
      OPT(CoordinateOptionsImpl, x1max) OPT(CoordinateOptionsImpl, x2min)
      OPT(CoordinateOptionsImpl, x2max) OPT(CoordinateOptionsImpl, x3min)
      OPT(CoordinateOptionsImpl, x3max) OPT(CoordinateOptionsImpl, nx1)
      OPT(CoordinateOptionsImpl, nx2) OPT(CoordinateOptionsImpl, nx3)
      OPT(CoordinateOptionsImpl, nghost);

  py::class_<HydroOptionsImpl, HydroOptions>(m, "HydroOptions")
      .def(py::init(&HydroOptionsImpl::create))
      .def_static("from_yaml", &HydroOptionsImpl::from_yaml, py::arg("filename"),
                  py::arg("verbose") = false)
      OPT(HydroOptionsImpl, verbose) OPT(HydroOptionsImpl, disable_flux_x1)
      OPT(HydroOptionsImpl, disable_flux_x2) OPT(HydroOptionsImpl, disable_flux_x3)
      OPT(HydroOptionsImpl, grav) OPT(HydroOptionsImpl, eos) OPT(HydroOptionsImpl, recon1)
      OPT(HydroOptionsImpl, recon23) OPT(HydroOptionsImpl, riemann) OPT(HydroOptionsImpl, icorr);

  py::class_<OutputOptionsImpl, OutputOptions>(m, "OutputOptions")
      .def(py::init(&OutputOptionsImpl::create))
      OPT(OutputOptionsImpl, fid) OPT(OutputOptionsImpl, dt) OPT(OutputOptionsImpl, file_type)
      OPT(OutputOptionsImpl, variables) OPT(OutputOptionsImpl, combine);

  py::class_<OutputType, std::shared_ptr<OutputType>>(m, "OutputType")
      .def_readonly("options", &OutputType::options)
      .def_readwrite("file_number", &OutputType::file_number)
      .def_readwrite("next_time", &OutputType::next_time);

  // restart files (kintera::save_tensors / load_tensors format,
  // src/input/read_restart_file.cpp, src/output/combine_restart.cpp)
  m.def("save_tensors", [](std::map<std::string, torch::Tensor> const& vars,
                           std::string const& filename) { save_tensors(vars, filename); },
        py::arg("tensor_map"), py::arg("filename"));
  m.def("load_tensors", &load_tensors, py::arg("filename"));
  m.def("load_restart", &load_restart, py::arg("path"), py::arg("block_rank") = 0);
  m.def("make_restart_bundle", &make_restart_bundle, py::arg("bundle_name"),
        py::arg("file_list"));

  py::class_<MeshBlockOptionsImpl, MeshBlockOptions>(m, "MeshBlockOptions")
      .def(py::init(&MeshBlockOptionsImpl::create))
      .def_static("from_yaml", &MeshBlockOptionsImpl::from_yaml, py::arg("filename"),
                  py::arg("verbose") = false)
      OPT(MeshBlockOptionsImpl, coord) OPT(MeshBlockOptionsImpl, hydro)
      OPT(MeshBlockOptionsImpl, integrator) OPT(MeshBlockOptionsImpl, cfl)
      OPT(MeshBlockOptionsImpl, nlim) OPT(MeshBlockOptionsImpl, tlim)
      OPT(MeshBlockOptionsImpl, max_redo)
      OPT(MeshBlockOptionsImpl, bfuncs) OPT(MeshBlockOptionsImpl, basename)
      OPT(MeshBlockOptionsImpl, output_dir) OPT(MeshBlockOptionsImpl, outputs);

  py::class_<IdealGasImpl, IdealGas>(m, "IdealGas")
      .def(py::init<EquationOfStateOptions, CoordinateOptions>(), py::arg("options"),
           py::arg("coord") = nullptr)
      .def_readonly("options", &IdealGasImpl::options)
      .def("nvar", &IdealGasImpl::nvar)
      .def("forward", &IdealGasImpl::forward, py::arg("cons"), py::arg("out") = py::none(),
           py::call_guard<py::gil_scoped_release>())
      .def("compute", &IdealGasImpl::compute);

  py::class_<ReconstructImpl, Reconstruct>(m, "Reconstruct")
      .def(py::init<ReconstructOptions, EquationOfStateOptions, bool, int>(), py::arg("options"),
           py::arg("eos") = nullptr, py::arg("has_parent") = false, py::arg("nghost") = 3)
      .def_readonly("options", &ReconstructImpl::options)
      .def("forward", &ReconstructImpl::forward, py::call_guard<py::gil_scoped_release>());

  py::class_<RiemannSolverImpl, RiemannSolver>(m, "RiemannSolver")
      .def(py::init<RiemannSolverOptions, double, std::string>())
      .def_readonly("options", &RiemannSolverImpl::options)
      .def("type", &RiemannSolverImpl::type)
      .def("forward", &RiemannSolverImpl::forward, py::arg("wl"), py::arg("wr"), py::arg("dim"),
           py::arg("flx"), py::arg("face_pressure") = torch::Tensor(),
           py::call_guard<py::gil_scoped_release>());
  auto solver = [&m](const char* name, const char* type) {
    m.def(name, [type](RiemannSolverOptions opts, double gammad) {
      return create_riemann_solver(opts, gammad, type);
    }, py::arg("options"), py::arg("gammad") = 1.4);
  };
  solver("LmarsSolver", "lmars");
  solver("HLLCSolver", "hllc");
  solver("RoeSolver", "roe");

  py::class_<CoordinateImpl, Coordinate>(m, "Cartesian")
      .def(py::init<CoordinateOptions>())
      .def_readonly("options", &CoordinateImpl::options)
      .def_readonly("x1v", &CoordinateImpl::x1v)
      .def_readonly("x2v", &CoordinateImpl::x2v)
      .def_readonly("x3v", &CoordinateImpl::x3v)
      .def_readonly("x1f", &CoordinateImpl::x1f)
      .def_readonly("x2f", &CoordinateImpl::x2f)
      .def_readonly("x3f", &CoordinateImpl::x3f)
      .def_readonly("dx1f", &CoordinateImpl::dx1f)
      .def_readonly("dx2f", &CoordinateImpl::dx2f)
      .def_readonly("dx3f", &CoordinateImpl::dx3f)
      .def("divergence", &CoordinateImpl::divergence)
      .def("forward", &CoordinateImpl::forward, py::arg("prim"), py::arg("flux1"),
           py::arg("flux2"), py::arg("flux3"), py::arg("face_pressure1") = torch::Tensor());
  // one module serves both geometries (options.type selects)
  m.attr("SphericalPolar") = m.attr("Cartesian");
  m.attr("Coordinate") = m.attr("Cartesian");

  py::class_<ImplicitHydroImpl, ImplicitHydro>(m, "ImplicitHydro")
      .def_readonly("options", &ImplicitHydroImpl::options)
      .def("forward", &ImplicitHydroImpl::forward, py::call_guard<py::gil_scoped_release>());

  py::class_<HydroImpl, Hydro>(m, "Hydro")
      .def(py::init<HydroOptions, CoordinateOptions, std::vector<std::string>>(),
           py::arg("options"), py::arg("coord"), py::arg("bfuncs") = std::vector<std::string>{})
      .def_readonly("options", &HydroImpl::options)
      .def_readonly("peos", &HydroImpl::peos)
      .def_readonly("precon1", &HydroImpl::precon1)
      .def_readonly("precon23", &HydroImpl::precon23)
      .def_readonly("priemann", &HydroImpl::priemann)
      .def_readonly("picorr", &HydroImpl::picorr)
      .def_readonly("pcoord", &HydroImpl::pcoord)
      .def("max_time_step", &HydroImpl::max_time_step)
      .def("flux1", &HydroImpl::flux1)
      .def("flux2", &HydroImpl::flux2)
      .def("flux3", &HydroImpl::flux3)
      .def("forward", &HydroImpl::forward, py::call_guard<py::gil_scoped_release>());

  py::class_<MeshBlockImpl, MeshBlock>(m, "MeshBlock")
      .def(py::init<MeshBlockOptions>())
      .def_readonly("options", &MeshBlockImpl::options)
      .def_readonly("phydro", &MeshBlockImpl::phydro)
      // (vars, start time), as python/csrc/pymesh.cpp:189-203; the dict is
      // also filled in place
      .def("initialize", [](MeshBlockImpl& b, py::dict vars) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             double t = b.initialize(v);
             for (auto& kv : v) vars[py::str(kv.first)] = kv.second;
             return py::make_tuple(vars, t);
           }, py::arg("vars"))
      .def("initialize_from_restart", [](MeshBlockImpl& b, std::string const& restart_file) {
             Variables v;
             double t = b.initialize(v, restart_file.c_str());
             py::dict vars;
             for (auto& kv : v) vars[py::str(kv.first)] = kv.second;
             return py::make_tuple(vars, t);
           }, py::arg("restart_file"))
      .def("make_outputs", [](MeshBlockImpl& b, py::dict vars, double time, bool final_write) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             b.make_outputs(v, time, final_write);
           }, py::arg("vars"), py::arg("time"), py::arg("final_write") = false)
      .def("finalize", [](MeshBlockImpl& b, py::dict vars, double time) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             b.finalize(v, time);
           }, py::arg("vars"), py::arg("time"))
      .def("get_outputs", [](MeshBlockImpl& b) { return b.output_types; })
      .def("max_time_step", [](MeshBlockImpl& b, py::dict vars) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             return b.max_time_step(v);
           })
      .def("forward", [](MeshBlockImpl& b, py::dict vars, double dt, int stage) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             b.forward(v, dt, stage);
             for (auto& kv : v) vars[py::str(kv.first)] = kv.second;
           })
      .def("check_redo", [](MeshBlockImpl& b, py::dict vars) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             return b.check_redo(v);
           })
      .def_readwrite("current_redo", &MeshBlockImpl::current_redo)
      .def_readwrite("cycle", &MeshBlockImpl::cycle)
      .def("step", [](MeshBlockImpl& b, py::dict vars, double dt, bool fused) {
             Variables v;
             for (auto item : vars) v[item.first.cast<std::string>()] = item.second.cast<torch::Tensor>();
             b.step(v, dt, fused);
             for (auto& kv : v) vars[py::str(kv.first)] = kv.second;
           }, py::arg("vars"), py::arg("dt"), py::arg("fused") = false);
}
