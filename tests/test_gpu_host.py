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
