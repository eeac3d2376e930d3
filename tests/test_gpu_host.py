"""GPU: the C++ host modules (YAML-configured MeshBlock / Hydro / Reconstruct
/ RiemannSolver) produce the oracle's results through libsnapx."""
import os

import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import host
from test_gpu_plain import host as to_host, var_errors

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEV = torch.device("cuda:0")


def _straka_oracle():
    oc = O.straka_config(nx2=512, nx1=64, implicit_scheme=1)
    oc.gammad = 1.4002789
    oc.weight = 28.970601e-3
    return oc


def test_meshblock_straka_from_yaml_matches_oracle():
    mb = host.MeshBlock(host.MeshBlockOptions.from_yaml(os.path.join(G, "straka.yaml")))
    oc = _straka_oracle()
    w0 = O.straka_init(oc)
    _, u_init = O.initialize(oc, w0)
    vars_ = {"hydro_w": torch.from_numpy(w0).to(DEV)}
    mb.initialize(vars_)
    assert np.allclose(to_host(vars_["hydro_u"]), u_init, rtol=1e-15, atol=0)
    for _ in range(2):
        mb.step(vars_, 0.1)
    u_ref = u_init
    for _ in range(2):
        u_ref, _ = O.rk3_step(oc, 0.1, u_ref)
    e = var_errors(to_host(vars_["hydro_u"]), u_ref, O.interior(oc))
    assert (e < 1e-12).all(), e


def test_meshblock_fused_step():
    mb = host.MeshBlock(host.MeshBlockOptions.from_yaml(os.path.join(G, "straka.yaml")))
    oc = _straka_oracle()
    _, u_init = O.initialize(oc, O.straka_init(oc))
    u_ref, _ = O.rk3_step(oc, 0.1, u_init)
    vars_ = {"hydro_u": torch.from_numpy(u_init).to(DEV, torch.float32)}
    mb.step(vars_, 0.1, fused=True)
    e = var_errors(to_host(vars_["hydro_u"]), u_ref, O.interior(oc))
    assert (e < 1e-5).all(), e


def test_hydro_forward_and_fluxes():
    mb = host.MeshBlock(host.MeshBlockOptions.from_yaml(os.path.join(G, "sod.yaml")))
    oc = O.sod_config(256)
    w, u = O.initialize(oc, O.sod_init(oc))
    du_ref, w_ref, F = O.hydro_forward(oc, 1e-3, u)
    ud = torch.from_numpy(u).to(DEV)
    wd = torch.empty_like(ud)
    du = mb.phydro.forward(1e-3, ud, {"hydro_w": wd})
    assert np.allclose(to_host(du), du_ref, rtol=1e-12, atol=1e-15)
    assert np.allclose(to_host(mb.phydro.flux1()), F[0], rtol=1e-13, atol=1e-15)
    assert np.allclose(to_host(wd), w_ref, rtol=1e-15, atol=0)
    dt = mb.max_time_step({"hydro_w": wd})
    assert abs(dt - O.max_time_step(oc, w_ref, 0.9)) < 1e-15


def test_reconstruct_and_riemann_modules():
    rng = np.random.default_rng(0)
    w = np.empty((5, 6, 7, 20))
    w[0] = rng.uniform(0.5, 2, w.shape[1:])
    w[1:4] = rng.uniform(-1, 1, (3,) + w.shape[1:])
    w[4] = rng.uniform(0.5, 2, w.shape[1:])
    ro = host.ReconstructOptions()
    ro.interp = host.InterpOptions("weno5")
    ro.shock = False
    rec = host.Reconstruct(ro, None, True)
    wlr = rec.forward(torch.from_numpy(w).to(DEV), 3)
    ref = O.reconstruct(w, 3, "weno5", shock=False)
    assert np.allclose(to_host(wlr), ref, rtol=1e-13, atol=1e-14)
    rs = host.HLLCSolver(host.RiemannSolverOptions(), 1.4)
    flx = torch.empty_like(wlr[0])
    fp = torch.empty_like(wlr[0][0])
    rs.forward(wlr[0].contiguous(), wlr[1].contiguous(), 3, flx, fp)
    fr, fpr = O.riemann("hllc", ref[0], ref[1], 3)
    assert np.allclose(to_host(flx), fr, rtol=1e-12, atol=1e-13)
    assert np.allclose(to_host(fp), fpr, rtol=1e-12, atol=1e-13)


def test_meshblock_check_redo_restores_and_halves_dt():
    """MeshBlockImpl::check_redo (src/mesh/meshblock.cpp:974-1028): a bad
    state after a step is replaced by the step's input, each redo halves the
    next CFL step, a good step resets the count, and more than max_redo
    redos in a row give -1."""
    opts = host.MeshBlockOptions.from_yaml(os.path.join(G, "straka.yaml"))
    assert opts.max_redo == 5
    opts.max_redo = 2
    mb = host.MeshBlock(opts)
    oc = _straka_oracle()
    vars_ = {"hydro_w": torch.from_numpy(O.straka_init(oc)).to(DEV)}
    mb.initialize(vars_)

    def dt_ratio():
        r, mb.current_redo = mb.current_redo, 0
        base = mb.max_time_step(vars_)
        mb.current_redo = r
        return mb.max_time_step(vars_) / base

    for fused in (False, True):
        if fused:
            vars_ = {k: v.float() for k, v in vars_.items()}
        mb.step(vars_, 0.1, fused=fused)
        assert mb.check_redo(vars_) == 0 and mb.current_redo == 0
        u_in = vars_["hydro_u"].clone()
        for r in (1, 2):
            mb.step(vars_, 0.1, fused=fused)
            vars_["hydro_u"][0, 0, 100, 20] = -1.0
            assert mb.check_redo(vars_) == 1 and mb.current_redo == r
            assert torch.equal(vars_["hydro_u"], u_in)
            assert dt_ratio() == pytest.approx(0.5 ** r, rel=1e-14)
        mb.step(vars_, 0.1, fused=fused)
        vars_["hydro_u"][0, 0, 100, 20] = -1.0
        assert mb.check_redo(vars_) == -1
        vars_["hydro_u"].copy_(u_in)
        mb.step(vars_, 0.1, fused=fused)
        assert mb.check_redo(vars_) == 0 and mb.current_redo == 0


def _run_to(mb, vars_, time, nlim, dt_fixed=None):
    """The time loop of examples/straka.cpp:138-162 (restart outputs written by
    make_outputs); returns the final time and (cycle, time, dt) per cycle."""
    log = []
    cycle = mb.cycle
    while cycle < nlim:
        cycle += 1
        mb.cycle = cycle
        dt = dt_fixed if dt_fixed is not None else mb.max_time_step(vars_)
        log.append((cycle, time, dt))
        mb.step(vars_, dt, fused=dt_fixed is not None)
        redo = mb.check_redo(vars_)
        if redo > 0:
            cycle = mb.cycle
            continue
        assert redo == 0
        time += dt
        mb.make_outputs(vars_, time)
    mb.finalize(vars_, time)
    return time, log


@pytest.mark.parametrize("fused", [False, True], ids=["fp64_cfl_dt", "fp32_fused_fixed_dt"])
def test_restart_resumes_the_uninterrupted_run(tmp_path, fused):
    """tests/run_restart_cycle_limit.py: a run restarted from a restart file
    written mid-way resumes at the cycle and time of the uninterrupted run,
    takes the same steps (cycle, time and dt equal) and ends at the same
    cycle, with a bit-identical state."""
    def block(sub):
        op = host.MeshBlockOptions.from_yaml(os.path.join(G, "straka.yaml"))
        out = host.OutputOptions()
        out.file_type, out.dt = "restart", 0.35
        op.outputs = [out]
        op.output_dir = str(tmp_path / sub)
        return host.MeshBlock(op)

    oc = _straka_oracle()
    w0 = torch.from_numpy(O.straka_init(oc)).to(DEV)
    dt_fixed = 0.1 if fused else None
    nlim = 9

    base = block("base")
    vars_ = {"hydro_w": w0.clone()}
    _, t0 = base.initialize(vars_)
    if fused:
        vars_ = {k: v.float() for k, v in vars_.items()}
    base.make_outputs(vars_, t0)
    t_end, base_log = _run_to(base, vars_, t0, nlim, dt_fixed)
    assert base.cycle == nlim
    files = sorted(os.listdir(tmp_path / "base"))
    assert files[0] == "straka.00000.restart" and files[-1] == "straka.final.restart"
    assert "straka.00002.restart" in files

    again = block("restart")
    rvars, t_r = again.initialize_from_restart(str(tmp_path / "base" / "straka.00002.restart"))
    assert rvars["hydro_u"].device.type == "cuda"
    assert rvars["hydro_u"].dtype == (torch.float32 if fused else torch.float64)
    start = again.cycle
    again.make_outputs(rvars, t_r)
    t_end2, log2 = _run_to(again, rvars, t_r, nlim, dt_fixed)
    matched = [e for e in base_log if e[1] == t_r]
    assert len(matched) == 1 and matched[0][0] == start + 1
    assert log2 == base_log[start:]
    assert t_end2 == t_end and again.cycle == base.cycle == nlim
    assert torch.equal(rvars["hydro_u"], vars_["hydro_u"])
    final = host.load_restart(str(tmp_path / "base" / "straka.final.restart"))
    assert torch.equal(final["hydro_u"], vars_["hydro_u"].cpu())
