"""GPU parity of the gnomonic-equiangle panel against the oracle.

Plain path, same bars as its other geometries: EOS and the coordinate
forward to 1e-13 of each variable's largest value, the hydro tendency to the
tolerance of the implicit scheme, RK3 steps to 1e-12 per conserved variable,
the CFL step to 1e-14; a uniform pressure at rest keeps no tendency with the
solvers that work in the face frames (LMARS, HLLC).  Fused FP32 path: 1e-5
per conserved variable against the FP64 oracle, single blocks and batch
members."""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from test_gpu_plain import _du_tol, cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu
PI4 = np.pi / 4
BC = ("reflecting", "reflecting", "outflow", "outflow", "outflow", "outflow")


def _panel(riemann="lmars", scheme=1, n=(12, 10, 10), seed=3, grav=-0.5, lo=-PI4, hi=PI4):
    """r in [1, 1.5], a whole panel by default: an isothermal atmosphere
    (R T = 1) under gravity grav, perturbed, with a random flow."""
    oc = O.make_config(nx1=n[0], nx2=n[1], nx3=n[2], nghost=3, x1min=1.0, x1max=1.5,
                       x2min=lo, x2max=hi, x3min=lo, x3max=hi, coord="gnomonic-equiangle",
                       gammad=1.4, recon_vertical=("weno5", False, False), riemann=riemann,
                       grav1=grav, implicit_scheme=scheme, bc=BC)
    n3, n2, n1 = O.shape(oc)
    r = O.geometry(oc, "x1v")
    rng = np.random.default_rng(seed)
    p = np.exp(grav * (r - 1.0))
    w = np.zeros((5, n3, n2, n1))
    w[4] = p * (1.0 + 0.02 * rng.uniform(-1, 1, (n3, n2, n1)))
    w[0] = p * (1.0 + 0.05 * rng.uniform(-1, 1, (n3, n2, n1)))
    w[1:4] = 0.05 * rng.uniform(-1, 1, (3, n3, n2, n1))
    return oc, *O.initialize(oc, w)


def test_eos_parity():
    oc, w, u = _panel()
    cfg = cfg_of(oc)
    wd = torch.empty_like(to_dev(u))
    capi.cons2prim(cfg, to_dev(u), wd)
    ref = O.cons2prim(oc, u)
    assert np.allclose(host(wd), ref, rtol=1e-13, atol=1e-15)
    # (not u itself: its ghosts are copies of interior conserved values, made
    # with the interior cells' metric)
    ud = torch.empty_like(wd)
    capi.prim2cons(cfg, to_dev(w), ud)
    assert np.allclose(host(ud), O.prim2cons(oc, w), rtol=1e-13, atol=1e-15)


def test_coord_forward_parity():
    oc, w, u = _panel()
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
    oc, w, u = _panel(riemann, scheme)
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
    assert np.allclose(host(wd), w_ref, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("case", [("lmars", 0), ("hllc", 1), ("roe", 9),
                                  (("lmars", "hllc", "roe"), 1)],
                         ids=["lmars_explicit", "hllc_vic3", "roe_vic5", "mixed_vic3"])
def test_rk3_step_parity(case):
    oc, w, u = _panel(*case)
    u_ref = u
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    cfg = cfg_of(oc)
    for _ in range(3):
        u_ref, _ = O.rk3_step(oc, 0.01, u_ref)
        capi.rk3_step(cfg, 0.01, ud, wd)
    e = var_errors(host(ud), u_ref, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-12).all(), e


def test_part_of_a_panel():
    """A block covering part of a panel (x2, x3 in [-0.3, 0.5]) with more
    cells: the same parity."""
    oc, w, u = _panel("hllc", 1, n=(16, 20, 20), lo=-0.3, hi=0.5)
    u_ref, _ = O.rk3_step(oc, 0.005, u)
    ud = to_dev(u)
    capi.rk3_step(cfg_of(oc), 0.005, ud, torch.empty_like(ud))
    e = var_errors(host(ud), u_ref, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-12).all(), e


@pytest.mark.parametrize("scheme", [0, 1])
def test_max_time_step_parity(scheme):
    oc, w, u = _panel(scheme=scheme)
    ref = O.max_time_step(oc, w, 0.9)
    out = torch.empty(1, dtype=torch.float64, device="cuda")
    capi.max_time_step(cfg_of(oc), to_dev(w), 0.9, out)
    assert abs(host(out)[0] - ref) < 1e-14 * ref


@pytest.mark.parametrize("riemann", ["lmars", "hllc"])
def test_rest_state_has_no_tendency(riemann):
    oc = O.make_config(nx1=12, nx2=10, nx3=10, nghost=3, x1min=1.0, x1max=1.5, x2min=-PI4,
                       x2max=PI4, x3min=-PI4, x3max=PI4, coord="gnomonic-equiangle",
                       riemann=riemann, bc=("outflow",) * 6)
    n3, n2, n1 = O.shape(oc)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = 1.3, 2.5
    u = O.prim2cons(oc, w)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    dud = torch.empty_like(ud)
    capi.hydro_forward(cfg_of(oc), 1e-3, ud, wd, dud)
    assert np.abs(host(dud)[O.interior(oc)]).max() < 1e-14


def test_check_state_on_a_panel():
    """The positivity check with the panel's cons2prim: a clean state passes;
    a cell with non-positive energy (whose pressure is negative with any
    metric, the kinetic energy being positive) or density asks for a redo."""
    oc, w, u = _panel()
    cfg = cfg_of(oc)
    flag = torch.zeros(1, dtype=torch.int32, device="cuda")
    capi.check_state(cfg, to_dev(u), flag)
    assert int(flag.item()) == 0
    x = u.copy()
    x[4][5, 6, 7] = -1.0
    capi.check_state(cfg, to_dev(x), flag)
    assert int(flag.item()) == capi.REDO_PRESSURE
    x = u.copy()
    x[0][6, 5, 8] = 0.0
    capi.check_state(cfg, to_dev(x), flag)
    assert int(flag.item()) & capi.REDO_DENSITY


def test_fp32_plain_path_close_to_fp64():
    oc, w, u = _panel()
    cfg = cfg_of(oc)
    u64, u32 = to_dev(u), to_dev(u, torch.float32)
    w64, w32 = torch.empty_like(u64), torch.empty_like(u32)
    for _ in range(3):
        capi.rk3_step(cfg, 0.01, u64, w64)
        capi.rk3_step(cfg, 0.01, u32, w32)
    e = var_errors(host(u32), host(u64), O.interior(oc), vel_scale=1.0)
    assert (e < 1e-5).all(), e


@pytest.mark.parametrize("case", [("lmars", 1, (12, 10, 10)), ("hllc", 9, (12, 10, 10)),
                                  ("roe", 1, (12, 10, 10)), ("lmars", 0, (12, 10, 10)),
                                  ("hllc", 1, (100, 8, 8))],
                         ids=["lmars_vic3", "hllc_vic5", "roe_vic3", "lmars_explicit",
                              "tall_hllc_vic3"])
def test_fused_steps_vs_fp64_oracle(case):
    """The FP32 fused path on a whole panel (face frames and metric sources
    in K1, per-layer areas and the metric EOS in the FP64 column solve)
    against the FP64 oracle: the fused bar, 1e-5 per conserved variable after
    three steps; stored ghosts as the reference boundary pass leaves them."""
    riemann, scheme, n = case
    oc, w, u = _panel(riemann, scheme, n=n)
    cfg = cfg_of(oc)
    uf = to_dev(u, torch.float32)
    u_ref = u
    for _ in range(3):
        u_ref, _ = O.rk3_step(oc, 0.01, u_ref)
        capi.fused_rk3_step(cfg, 0.01, uf)
    got = host(uf)
    e = var_errors(got, u_ref, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-5).all(), e
    assert np.array_equal(O.apply_bc(oc, got, O.CONSERVED), got)


def test_fused_part_of_a_panel():
    """Part of a panel (x2, x3 in [-0.3, 0.5], 16 x 20 x 20): the fused bar."""
    oc, w, u = _panel("hllc", 1, n=(16, 20, 20), lo=-0.3, hi=0.5)
    uf = to_dev(u, torch.float32)
    u_ref = u
    for _ in range(2):
        u_ref, _ = O.rk3_step(oc, 0.005, u_ref)
        capi.fused_rk3_step(cfg_of(oc), 0.005, uf)
    e = var_errors(host(uf), u_ref, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-5).all(), e


@pytest.mark.parametrize("riemann", ["lmars", "hllc"])
def test_fused_rest_state_keeps_still(riemann):
    """A uniform pressure at rest stays at rest on the fused path: the face
    frames and the face-pressure form of the sources balance in FP32 to
    roundoff (velocities below 1e-5 of the sound speed after three steps)."""
    oc = O.make_config(nx1=12, nx2=10, nx3=10, nghost=3, x1min=1.0, x1max=1.5, x2min=-PI4,
                       x2max=PI4, x3min=-PI4, x3max=PI4, coord="gnomonic-equiangle",
                       riemann=riemann, implicit_scheme=0, bc=("outflow",) * 6)
    n3, n2, n1 = O.shape(oc)
    w = np.zeros((5, n3, n2, n1))
    w[0], w[4] = 1.3, 2.5
    u = O.prim2cons(oc, w)
    uf = to_dev(u, torch.float32)
    for _ in range(3):
        capi.fused_rk3_step(cfg_of(oc), 1e-3, uf)
    got = host(uf)[O.interior(oc)]
    cs = np.sqrt(1.4 * 2.5 / 1.3)
    assert np.abs(got[1:4] / 1.3).max() < 1e-5 * cs
    assert np.abs(got[0] - 1.3).max() < 1e-6


@pytest.mark.parametrize("scheme", [1, 9])
def test_panels_as_batch(scheme):
    """Gnomonic panels of different shapes and seeds as members of one batch
    (persistent K1 with face frames and metric sources, cell-parallel K2):
    after the autotuner and two steps each member matches the FP64 oracle at
    the fused bar; a batch mixing panels and spherical-polar blocks is
    rejected."""
    shapes = [(12, 10, 10), (16, 12, 12), (20, 8, 8)] if scheme == 1 else [(12, 10, 10), (10, 8, 8)]
    cases = [_panel("hllc" if m % 2 else "lmars", scheme, n=n, seed=50 + m)
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
    sph = O.make_config(nx1=12, nx2=10, nx3=10, nghost=3, x1min=1.0, x1max=1.5, x2min=0.5,
                        x2max=2.5, x3min=0.0, x3max=1.0, coord="spherical-polar",
                        implicit_scheme=scheme, grav1=-0.5)
    with pytest.raises(capi.SnapxError, match="same coordinate"):
        capi.Batch([cfg_of(cases[0][0]), cfg_of(sph)], dt)


def test_fused_columns_too_tall_are_refused():
    """Curvilinear columns run on the cell-parallel solve only: a panel with
    more than 256 cells along x1 (3x3 blocks) is refused with a message."""
    oc, w, u = _panel(n=(12, 10, 10))
    cfg = cfg_of(oc)
    cfg.nx1 = 300
    uf = torch.zeros((5, 16, 16, 306), dtype=torch.float32, device="cuda")
    with pytest.raises(capi.SnapxError, match="curvilinear columns"):
        capi.fused_rk3_step(cfg, 0.01, uf)
