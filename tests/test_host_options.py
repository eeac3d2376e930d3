"""C++ host modules: YAML options parse like the reference's from_yaml code
(defaults and quirks), and the modules refuse host tensors (no CPU path)."""
import os
import tempfile

import pytest
import torch

from snapy_linfel_amd import host

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _yaml(text):
    f = tempfile.NamedTemporaryFile("w", suffix=".yaml", delete=False)
    f.write(text)
    f.close()
    return f.name


def test_straka_options():
    op = host.HydroOptions.from_yaml(os.path.join(G, "straka.yaml"))
    assert op.eos.type == "ideal-gas"
    assert abs(op.eos.gammad - 1.4002789) < 1e-12
    assert op.recon1.interp.type == "weno5" and op.recon1.shock is False
    assert op.recon23.interp.type == "weno5" and op.recon23.interp.scale is False
    assert op.riemann.type == "lmars"
    assert op.icorr.scheme == 1 and op.icorr.size() == 3 and op.icorr.type() == "vic-partial"
    assert op.grav.grav1 == -9.81 and op.grav.non_hydrostatic == 1.0
    mb = host.MeshBlockOptions.from_yaml(os.path.join(G, "straka.yaml"))
    assert mb.coord.nx1 == 64 and mb.coord.nx2 == 512 and mb.coord.nc1() == 70
    assert mb.coord.nc3() == 1
    assert mb.bfuncs == ["reflecting", "reflecting", "periodic", "periodic", "", ""]
    assert mb.cfl == 0.9 and mb.nlim == 1000


def test_per_direction_riemann_extension():
    op = host.RiemannSolverOptions.from_yaml(os.path.join(G, "box3d.yaml"))
    assert [op.type_of(a) for a in (1, 2, 3)] == ["lmars", "roe", "roe"]


def test_reference_yaml_quirks():
    # riemann-solver present without type -> "roe"; absent -> struct "hllc"
    # (src/riemann/riemann_solver.cpp:15-28, riemann_solver.hpp:33)
    f = _yaml("dynamics:\n  riemann-solver:\n    dir: omni\n")
    assert host.RiemannSolverOptions.from_yaml(f).type == "roe"
    f2 = _yaml("dynamics:\n  verbose: false\n")
    assert host.RiemannSolverOptions.from_yaml(f2).type == "hllc"
    # reconstruct section present without keys -> type "dc", shock false;
    # absent -> struct default shock true (src/recon/reconstruct.cpp:29-31)
    f3 = _yaml("dynamics:\n  reconstruct:\n    vertical: {scale: true}\n")
    r = host.ReconstructOptions.from_yaml(f3, "vertical")
    assert r.interp.type == "dc" and r.shock is False and r.interp.scale is True
    assert host.ReconstructOptions.from_yaml(f3, "horizontal").shock is True
    # EOS floors: YAML defaults 1e-6 / 1e-3, struct 1e-10
    # (src/eos/equation_of_state.cpp:45-51)
    f4 = _yaml("dynamics:\n  equation-of-state:\n    type: ideal-gas\n")
    e = host.EquationOfStateOptions.from_yaml(f4)
    assert e.density_floor == 1e-6 and e.pressure_floor == 1e-3 and e.gammad == 1.4
    assert host.EquationOfStateOptions().density_floor == 1e-10


def test_implicit_scheme_bits():
    # tests/test_backward_substitution.cpp:155-167
    f = _yaml("integration:\n  implicit-scheme: 9\n")
    op = host.ImplicitOptions.from_yaml(f)
    assert op.scheme == 9 and op.size() == 5 and op.type() == "vic-full"
    assert host.ImplicitOptions.from_yaml(_yaml("integration:\n  cfl: 0.5\n")) is None
    bad = host.ImplicitOptions()
    bad.scheme = 4
    with pytest.raises(RuntimeError, match="Unsupported implicit scheme"):
        bad.type()


def test_disable_flux_zeroes_gravity():
    f = _yaml("dynamics:\n  disable-flux-x1: true\nforcing:\n  const-gravity: {grav1: -9.8}\n")
    op = host.HydroOptions.from_yaml(f)
    assert op.disable_flux_x1 is True and op.grav.grav1 == 0.0


def test_non_hydrostatic_range_checked():
    f = _yaml("forcing:\n  const-gravity: {grav1: -9.8, non-hydrostatic: 1.5}\n")
    with pytest.raises(RuntimeError):
        host.HydroOptions.from_yaml(f)


def test_unsupported_eos_and_geometry_are_errors():
    f = _yaml("dynamics:\n  equation-of-state: {type: moist-mixture}\n"
              "geometry:\n  cells: {nx1: 8, nghost: 3}\n")
    with pytest.raises(RuntimeError, match="ideal-gas only"):
        host.MeshBlock(host.MeshBlockOptions.from_yaml(f))
    g = _yaml("geometry:\n  type: cylindrical\n")
    with pytest.raises(RuntimeError, match="gnomonic-equiangle only"):
        host.CoordinateOptions.from_yaml(g)
    assert host.CoordinateOptions.from_yaml(
        _yaml("geometry:\n  type: gnomonic-equiangle\n")).type == "gnomonic-equiangle"


def test_modules_refuse_host_tensors():
    mb = host.MeshBlock(host.MeshBlockOptions.from_yaml(os.path.join(G, "sod.yaml")))
    u = torch.zeros(5, 1, 1, 262, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="GPU tensor"):
        mb.phydro.forward(0.1, u, {"hydro_w": u.clone()})
    r = host.Reconstruct(host.ReconstructOptions())
    with pytest.raises(RuntimeError, match="GPU tensor"):
        r.forward(u, 3)
