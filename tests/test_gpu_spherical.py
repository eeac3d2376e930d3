"""GPU parity of the spherical-polar plain path against the oracle.

Same bars as the Cartesian plain path: the coordinate forward to 1e-13 of
each variable's largest value, the hydro tendency to the tolerance of the
implicit scheme, one RK3 step to 1e-12 per conserved variable, the CFL step
to 1e-14."""
import os

import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi, host as H
from test_gpu_plain import _du_tol, cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BC = ("reflecting", "reflecting", "reflecting", "reflecting", "periodic", "periodic")


def _sector(riemann="hllc", scheme=1, n=(14, 10, 8), seed=3, grav=-0.5):
    """r in [1, 2], theta in [pi/4, 3pi/4], phi in [-pi/2, pi/2] (periodic):
    an isothermal atmosphere (R T = 1) under gravity grav, perturbed."""
    oc = O.make_config(nx1=n[0], nx2=n[1], nx3=n[2], nghost=3, x1min=1.0, x1max=2.0,
                       x2min=0.25 * np.pi, x2max=0.75 * np.pi, x3min=-0.5 * np.pi,
                       x3max=0.5 * np.pi, coord="spherical-polar", gammad=1.4,
                       recon_vertical=("weno5", False, False), riemann=riemann, grav1=grav,
                       implicit_scheme=scheme, bc=BC)
    n3, n2, n1 = O.shape(oc)
    r = O.geometry(oc, "x1v")
    rng = np.random.default_rng(seed)
    p = np.exp(grav * (r - 1.0))
    w = np.zeros((5, n3, n2, n1))
    w[4] = p * (1.0 + 0.02 * rng.uniform(-1, 1, (n3, n2, n1)))
    w[0] = p * (1.0 + 0.05 * rng.uniform(-1, 1, (n3, n2, n1)))
    w[1:4] = 0.05 * rng.uniform(-1, 1, (3, n3, n2, n1))
    return oc, *O.initialize(oc, w)


def test_coord_forward_parity():
    oc, w, u = _sector()
    n3, n2, n1 = O.shape(oc)
    rng = np.random.default_rng(21)
    f1, f2, f3 = (rng.standard_normal((5, n3, n2, n1)) for _ in range(3))
    fp1 = rng.standard_normal((n3, n2, n1))
    cfg = cfg_of(oc)
    for fp in (fp1, None):
        ref = O.coord_forward(oc, w, f1, f2, f3, fp)
        div = torch.empty((5, n3, n2, n1), dtype=torch.float64, device="cuda")
        capi.coord_forward(cfg, to_dev(w), to_dev(f1), to_dev(f2), to_dev(f3),
                           None if fp is None else to_dev(fp), div)
        got = host(div)
        for v in range(5):
            scale = np.abs(ref[v]).max()
            assert np.abs(got[v] - ref[v]).max() < 1e-13 * scale, (v, fp is None)


@pytest.mark.parametrize("riemann", ["lmars", "hllc", "roe"])
@pytest.mark.parametrize("scheme", [0, 1, 9])
def test_hydro_forward_parity(riemann, scheme):
    oc, w, u = _sector(riemann, scheme)
    du_ref, w_ref, _ = O.hydro_forward(oc, 0.01, u)
    cfg = cfg_of(oc)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    dud = torch.empty_like(ud)
    capi.hydro_forward(cfg, 0.01, ud, wd, dud)
    I = O.interior(oc)
    scale = np.abs(du_ref[I]).max(axis=(1, 2, 3)) + 1e-300
    rel = np.abs(host(dud)[I] - du_ref[I]).max(axis=(1, 2, 3)) / scale
    assert (rel < _du_tol(scheme)).all(), rel
    # the gravity ghost refill uses the spherical centre distances
    assert np.allclose(host(wd), w_ref, rtol=1e-14, atol=0)


@pytest.mark.parametrize("case", [("lmars", 0), ("hllc", 1), ("roe", 9)])
def test_rk3_step_parity(case):
    oc, w, u = _sector(*case)
    u_ref = u
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    cfg = cfg_of(oc)
    for _ in range(3):
        u_ref, _ = O.rk3_step(oc, 0.01, u_ref)
        capi.rk3_step(cfg, 0.01, ud, wd)
    I = O.interior(oc)
    e = var_errors(host(ud), u_ref, I, vel_scale=1.0)
    assert (e < 1e-12).all(), e
    vol = O.geometry(oc, "vol")
    m0 = (vol[I[1:]] * u[0][I[1:]]).sum()
    m1 = (vol[I[1:]] * host(ud)[0][I[1:]]).sum()
    assert abs(m1 - m0) / m0 < 1e-13


@pytest.mark.parametrize("scheme", [0, 1])
def test_max_time_step_parity(scheme):
    oc, w, u = _sector(scheme=scheme)
    ref = O.max_time_step(oc, w, 0.9)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    capi.max_time_step(cfg_of(oc), to_dev(w), 0.9, out)
    assert abs(host(out)[0] - ref) < 1e-14 * ref


def test_rest_state_has_no_tendency():
    oc = O.make_config(nx1=14, nx2=10, nx3=8, nghost=3, x1min=1.0, x1max=2.0,
                       x2min=0.25 * np.pi, x2max=0.75 * np.pi, x3min=-0.5 * np.pi,
                       x3max=0.5 * np.pi, coord="spherical-polar", riemann="lmars", bc=BC)
    n3, n2, n1 = O.shape(oc)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = 1.3, 2.5
    u = O.prim2cons(oc, w)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    dud = torch.empty_like(ud)
    capi.hydro_forward(cfg_of(oc), 1e-3, ud, wd, dud)
    assert np.abs(host(dud)[O.interior(oc)]).max() < 1e-14


def test_meshblock_from_yaml_matches_oracle():
    """The host modules with geometry.type spherical-polar (bounds in units
    of pi) step like the oracle."""
    mb = H.MeshBlock(H.MeshBlockOptions.from_yaml(os.path.join(G, "spherical.yaml")))
    coord = mb.phydro.pcoord
    oc = O.make_config(nx1=16, nx2=8, nx3=8, nghost=3, x1min=1.0, x1max=2.0,
                       x2min=0.25 * np.pi, x2max=0.75 * np.pi, x3min=-0.5 * np.pi,
                       x3max=0.5 * np.pi, coord="spherical-polar", gammad=1.4,
                       recon_vertical=("weno5", False, False), riemann="hllc", grav1=-0.5,
                       implicit_scheme=1, bc=BC)
    assert np.allclose(coord.x1v.numpy(), O.geometry(oc, "x1v"), rtol=1e-15)
    assert np.allclose(coord.x2v.numpy(), O.geometry(oc, "x2v"), rtol=1e-15)
    _, w, u = _sector(n=(16, 8, 8))
    vars_ = {"hydro_w": to_dev(w)}
    mb.initialize(vars_)
    dt = mb.max_time_step(vars_)
    assert dt == pytest.approx(O.max_time_step(oc, w, 0.9), rel=1e-14)
    mb.step(vars_, 0.01)
    u_ref, _ = O.rk3_step(oc, 0.01, u)
    e = var_errors(host(vars_["hydro_u"]), u_ref, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-12).all(), e
    # the FP32 fused path of the same block
    vf = {"hydro_u": to_dev(u, torch.float32)}
    mb.step(vf, 0.01, fused=True)
    e = var_errors(host(vf["hydro_u"]), u_ref, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-5).all(), e


def test_fp32_plain_path_close_to_fp64():
    oc, w, u = _sector()
    cfg = cfg_of(oc)
    u64, u32 = to_dev(u), to_dev(u, torch.float32)
    w64, w32 = torch.empty_like(u64), torch.empty_like(u32)
    for _ in range(3):
        capi.rk3_step(cfg, 0.01, u64, w64)
        capi.rk3_step(cfg, 0.01, u32, w32)
    e = var_errors(host(u32), host(u64), O.interior(oc), vel_scale=1.0)
    assert (e < 1e-5).all(), e


@pytest.mark.parametrize("case", [("hllc", 1, (14, 10, 8)), ("lmars", 0, (14, 10, 8)),
                                  ("roe", 9, (14, 10, 8)), ("lmars", 1, (64, 24, 1)),
                                  ("hllc", 1, (100, 12, 1))],
                         ids=["3d_hllc_vic3", "3d_lmars_explicit", "3d_roe_vic5", "2d_lmars_vic3",
                              "2d_tall_hllc_vic3"])
def test_fused_steps_vs_fp64_oracle(case):
    """The FP32 fused path on spherical-polar blocks (per-face areas and
    geometric sources in K1, per-layer areas in the FP64 column solve) against
    the FP64 oracle: the fused bar, 1e-5 per conserved variable after three
    steps; interior mass sum V rho conserved to FP32 roundoff; stored ghosts as
    the reference boundary pass leaves them."""
    riemann, scheme, n = case
    oc, w, u = _sector(riemann, scheme, n=n)
    cfg = cfg_of(oc)
    uf = to_dev(u, torch.float32)
    u_ref = u
    for _ in range(3):
        u_ref, _ = O.rk3_step(oc, 0.01, u_ref)
        capi.fused_rk3_step(cfg, 0.01, uf)
    got = host(uf)
    I = O.interior(oc)
    e = var_errors(got, u_ref, I, vel_scale=1.0)
    assert (e < 1e-5).all(), e
    vol = O.geometry(oc, "vol")
    m0 = (vol[I[1:]] * u[0][I[1:]]).sum()
    assert abs((vol[I[1:]] * got[0][I[1:]]).sum() - m0) / m0 < 1e-6
    assert np.array_equal(O.apply_bc(oc, got, O.CONSERVED), got)


@pytest.mark.parametrize("scheme", [1, 9])
def test_spherical_members_as_batch(scheme):
    """Spherical-polar blocks of different shapes and seeds as members of one
    batch (persistent K1 with geometric sources, cell-parallel spherical K2):
    after the autotuner and two steps each member matches the FP64 oracle at
    the fused bar; a batch that mixes coordinate systems is rejected."""
    shapes = [(14, 10, 8), (32, 12, 1), (20, 6, 4)] if scheme == 1 else [(14, 10, 8), (24, 8, 1)]
    cases = [_sector("hllc" if m % 2 else "lmars", scheme, n=n, seed=40 + m)
             for m, n in enumerate(shapes)]
    dt = 0.01
    b = capi.Batch([cfg_of(oc) for oc, _, _ in cases], dt)
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, _, u0 in cases])
    before = u.clone()
    b.autotune(u, dt, trials=2)
    assert torch.equal(u, before)
    b.steps(u, 2)
    for m, (oc, _, u0) in enumerate(cases):
        u_ref = u0
        for _ in range(2):
            u_ref, _ = O.rk3_step(oc, dt, u_ref)
        e = var_errors(host(b.member(u, m)), u_ref, O.interior(oc), vel_scale=1.0)
        assert (e < 1e-5).all(), (m, shapes[m], e)
    cart = O.make_config(nx1=16, nx2=8, nx3=1, nghost=3, x1min=0.0, x1max=1.0, x2min=0.0,
                         x2max=1.0, implicit_scheme=scheme)
    with pytest.raises(capi.SnapxError, match="same coordinate"):
        capi.Batch([cfg_of(cases[0][0]), cfg_of(cart)], dt)
