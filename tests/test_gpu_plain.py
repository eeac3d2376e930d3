"""GPU parity of the plain path (C ABI of libsnapx.so) against the oracle.

FP64 bar: per conserved variable, max |gpu - oracle| / scale <= 1e-12, where
scale is max |oracle value| of that variable (momentum: max(rho) * 300 m/s
when the field is near zero).
Module-level checks (EOS, reconstruction, Riemann) use 1e-13 relative.
"""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def to_dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)


def host(t):
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def cfg_of(oc):
    return capi.Config.from_buffer_copy(oc)


def var_errors(got, ref, I=None, vel_scale=300.0):
    if I is not None:
        got, ref = got[I], ref[I]
    rho_max = np.abs(ref[0]).max()
    errs = []
    for v in range(5):
        scale = np.abs(ref[v]).max()
        if v in (1, 2, 3):
            scale = max(scale, rho_max * vel_scale)
        errs.append(np.abs(got[v] - ref[v]).max() / scale)
    return np.array(errs)


def _rand_prim(shape, seed):
    rng = np.random.default_rng(seed)
    w = np.empty((5,) + shape)
    w[0] = rng.uniform(0.5, 2.0, shape)
    w[1:4] = rng.uniform(-30.0, 30.0, (3,) + shape)
    w[4] = rng.uniform(0.5e5, 1.5e5, shape)
    return w


def test_library_reports_device():
    assert capi.lib().snapx_device_count() >= 1


def test_cons2prim_prim2cons():
    oc = O.make_config(nx1=20, nx2=12, nx3=6, nghost=3)
    w = _rand_prim(O.shape(oc), 1)
    u = O.prim2cons(oc, w)
    cfg = cfg_of(oc)
    ud = to_dev(w)
    uo = torch.empty_like(ud)
    capi.prim2cons(cfg, ud, uo)
    assert np.allclose(host(uo), u, rtol=1e-15, atol=0)
    wo = torch.empty_like(ud)
    capi.cons2prim(cfg, to_dev(u), wo)
    assert np.allclose(host(wo), O.cons2prim(oc, u), rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("kind", ["weno5", "cp5", "weno3", "cp3", "plm", "dc"])
@pytest.mark.parametrize("dim", [3, 2, 1])
@pytest.mark.parametrize("shock", [True, False])
def test_reconstruct_parity(kind, dim, shock):
    shape = (14, 11, 16)
    w = _rand_prim(shape, 7 + dim)
    ref = O.reconstruct(w, dim, kind, shock=shock)
    wd = to_dev(w)
    out = torch.empty((2,) + w.shape, dtype=torch.float64, device=DEV)
    capi.reconstruct(wd, out, dim, kind, shock=shock)
    got = host(out)
    scale = np.abs(ref).max(axis=(0, 2, 3, 4), keepdims=True)
    assert np.abs(got - ref).max() <= 0 or (np.abs(got - ref) / scale).max() < 1e-13


def test_reconstruct_scale_and_limiter_options():
    w = _rand_prim((10, 9, 12), 3)
    w[1] -= 40.0
    for kw in ({"scale": True, "shock": True}, {"shock": False, "limiter": True,
                                                 "dfloor": 1.2, "pfloor": 1.0e5}):
        ref = O.reconstruct(w, 3, "weno5", **kw)
        out = torch.empty((2,) + w.shape, dtype=torch.float64, device=DEV)
        capi.reconstruct(to_dev(w), out, 3, "weno5", **kw)
        got = host(out)
        scale = np.abs(ref).max(axis=(0, 2, 3, 4), keepdims=True)
        assert (np.abs(got - ref) / scale).max() < 1e-13


@pytest.mark.parametrize("solver", ["lmars", "hllc", "roe"])
@pytest.mark.parametrize("dim", [3, 2, 1])
def test_riemann_parity(solver, dim):
    shape = (12, 10, 9)
    wl = _rand_prim(shape, 21)
    wr = _rand_prim(shape, 22)
    wr[1:4] *= -1.0
    ref, fpr = O.riemann(solver, wl, wr, dim)
    flx = torch.empty(wl.shape, dtype=torch.float64, device=DEV)
    fp = torch.empty(shape, dtype=torch.float64, device=DEV)
    capi.riemann(solver, to_dev(wl), to_dev(wr), flx, fp, dim)
    got = host(flx)
    scale = np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    assert (np.abs(got - ref) / scale).max() < 1e-13
    assert (np.abs(host(fp) - fpr) / np.abs(fpr).max()).max() < 1e-13


def _branch_faces(dim):
    """Faces of four regimes interleaved along x1: supersonic to +n, to -n,
    strong rarefaction (LLF fallback of Roe) and subsonic close states."""
    from test_oracle_vic_blocks import faces
    parts = [faces(kind, dim, n=48, seed=5) for kind in ("right", "left", "llf", "sub")]
    wl = np.concatenate([p[0] for p in parts], axis=-1)
    wr = np.concatenate([p[1] for p in parts], axis=-1)
    return wl, wr


@pytest.mark.parametrize("solver", ["roe", "lmars", "hllc"])
@pytest.mark.parametrize("dim", [3, 2, 1])
def test_riemann_parity_every_branch(solver, dim):
    """Supersonic (lambda_- > 0, lambda_+ < 0), LLF-fallback and subsonic
    faces (src/riemann/roe_impl.h:178-211) against the oracle, whose branches
    are pinned by test_oracle_vic_blocks.py: FP64 to 1e-13, FP32 to 1e-5
    (per variable, scaled by its largest flux)."""
    from test_oracle_vic_blocks import expected_flux
    wl, wr = _branch_faces(dim)
    ref, fpr = O.riemann(solver, wl, wr, dim)
    if solver == "roe":
        # the oracle itself takes the pinned branch on each regime
        for q, kind in enumerate(("right", "left", "llf")):
            s = slice(48 * q, 48 * (q + 1))
            e = expected_flux(kind, wl[..., s], wr[..., s], dim)
            assert np.abs(ref[..., s] - e).max() < 1e-12 * np.abs(e).max()
    scale = np.abs(ref).max(axis=(1, 2, 3), keepdims=True)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 1e-5)):
        flx = torch.empty(wl.shape, dtype=dt, device=DEV)
        fp = torch.empty(wl.shape[1:], dtype=dt, device=DEV)
        capi.riemann(solver, to_dev(wl, dt), to_dev(wr, dt), flx, fp, dim)
        got = host(flx)
        assert (np.abs(got - ref) / scale).max() < tol, (dt, (np.abs(got - ref) / scale).max())
        assert (np.abs(host(fp) - fpr) / np.abs(fpr).max()).max() < tol


def test_riemann_reference_face_pressure_seeds():
    """tests/test_riemann.cpp:412-570 inputs (torch seeds 77/78/79) through the
    GPU solvers against the oracle (itself pinned to the closed forms)."""
    from test_oracle_riemann import _states
    for solver, seed in (("hllc", 77), ("lmars", 78), ("roe", 79)):
        wl, wr = _states(seed)
        ref, fpr = O.riemann(solver, wl, wr, 3)
        flx = torch.empty(wl.shape, dtype=torch.float64, device=DEV)
        fp = torch.empty(wl.shape[1:], dtype=torch.float64, device=DEV)
        capi.riemann(solver, to_dev(wl), to_dev(wr), flx, fp, 3)
        assert np.allclose(host(fp), fpr, rtol=1e-13, atol=1e-13)
        assert np.allclose(host(flx), ref, rtol=1e-12, atol=1e-12)


def test_boundary_functions_exact():
    oc = O.make_config(nx1=7, nx2=6, nx3=5, nghost=3,
                       bc=("reflecting", "outflow", "periodic", "periodic", "outflow",
                           "reflecting"))
    v = np.random.default_rng(4).normal(size=(5,) + O.shape(oc))
    ref = O.apply_bc(oc, v, O.CONSERVED)
    d = to_dev(v)
    capi.apply_bc(cfg_of(oc), d, capi.CONSERVED)
    assert np.array_equal(host(d), ref)


def _straka_small(scheme, nx2=48, nx1=24, riemann="lmars", wind=0.0):
    """Reduced Straka block (dx1 = 6.4 km / nx1, dx2 = 100 m).  ``wind`` adds
    a smooth velocity field of that amplitude (m/s), so that every tendency
    is O(dt u / dx) instead of the roundoff residual of a state at rest."""
    oc = O.straka_config(nx2=nx2, nx1=nx1, implicit_scheme=scheme, riemann=riemann)
    oc.x2max = 25.6e3 * nx2 / 512 * 2
    wp = O.straka_init(oc)
    if wind:
        n3, n2, n1 = O.shape(oc)
        z = (np.arange(n1) - oc.nghost + 0.5) / oc.nx1
        x = (np.arange(n2) - oc.nghost + 0.5) / oc.nx2
        X, Z = np.meshgrid(x, z, indexing="ij")
        wp[1, 0] = wind * np.sin(np.pi * Z) * np.sin(2 * np.pi * X)
        wp[2, 0] = wind * np.cos(np.pi * Z) * np.cos(2 * np.pi * X)
    w, u = O.initialize(oc, wp)
    return oc, w, u


def _du_tol(scheme):
    """Bar for du vs the oracle, relative to the largest tendency.  Host and
    device pow/sqrt differ by an ulp; the 5x5 block solve (scheme 9, partial
    pivoting) amplifies that by the condition number of its blocks (acoustic
    CFL ~2 in the vertical), so it gets 1e-9 instead of 1e-11."""
    return 1e-9 if scheme == 9 else 1e-11


@pytest.mark.parametrize("scheme", [0, 1, 9])
def test_hydro_forward_straka(scheme):
    oc, w, u = _straka_small(scheme, wind=10.0)
    du_ref, w_ref, F = O.hydro_forward(oc, 0.5, u)
    cfg = cfg_of(oc)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    dud = torch.empty_like(ud)
    capi.hydro_forward(cfg, 0.5, ud, wd, dud)
    I = O.interior(oc)
    e = var_errors(host(dud), du_ref, I, vel_scale=1.0)
    # du itself: scale by the largest tendency of the variable
    scale = np.abs(du_ref[I]).max(axis=(1, 2, 3)) + 1e-300
    rel = np.abs(host(dud)[I] - du_ref[I]).max(axis=(1, 2, 3)) / scale
    assert (rel < _du_tol(scheme)).all(), rel
    assert np.allclose(host(wd), w_ref, rtol=1e-14, atol=0)


def _box3d(riemann=("lmars", "roe", "roe"), scheme=1, n=(16, 12, 10)):
    oc = O.make_config(nx1=n[0], nx2=n[1], nx3=n[2], nghost=3, x1min=0.0, x1max=1600.0,
                       x2min=0.0, x2max=1200.0, x3min=0.0, x3max=1000.0, gammad=1.4,
                       weight=8.31446261815324 / 287.0, recon_vertical=("weno5", False, False),
                       riemann=riemann, grav1=-9.81, implicit_scheme=scheme,
                       bc=("reflecting", "reflecting", "periodic", "periodic", "periodic",
                           "periodic"))
    n3, n2, n1 = O.shape(oc)
    z = (np.arange(n1) - 3 + 0.5) * 100.0
    T = 250.0 - 9.81 * z / 1004.5
    p = 1e5 * (T / 250.0) ** (1004.5 / 287.0)
    rng = np.random.default_rng(1234)
    w = np.zeros((5, n3, n2, n1))
    w[4] = p
    th = rng.uniform(-0.5, 0.5, size=(n3, n2, n1)) * (np.arange(n1) < 3 + 5)
    w[0] = p / (287.0 * (T + th))
    w[1:4] = rng.uniform(-1, 1, size=(3, n3, n2, n1))
    return oc, *O.initialize(oc, w)


@pytest.mark.parametrize("scheme", [0, 1, 9])
def test_hydro_forward_3d_per_direction_solvers(scheme):
    oc, w, u = _box3d(scheme=scheme)
    du_ref, w_ref, _ = O.hydro_forward(oc, 0.5, u)
    cfg = cfg_of(oc)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    dud = torch.empty_like(ud)
    capi.hydro_forward(cfg, 0.5, ud, wd, dud)
    I = O.interior(oc)
    scale = np.abs(du_ref[I]).max(axis=(1, 2, 3)) + 1e-300
    rel = np.abs(host(dud)[I] - du_ref[I]).max(axis=(1, 2, 3)) / scale
    assert (rel < _du_tol(scheme)).all(), rel


@pytest.mark.parametrize("case", ["straka1", "straka9", "box_lmars_roe", "box_hllc"])
def test_rk3_step_fp64_parity(case):
    """One full SSP-RK3 step from identical seeded state: <= 1e-12 per
    conserved variable."""
    if case.startswith("straka"):
        oc, w, u = _straka_small(int(case[-1]))
    elif case == "box_hllc":
        oc, w, u = _box3d(riemann="hllc", scheme=1)
    else:
        oc, w, u = _box3d()
    u_ref, _ = O.rk3_step(oc, 0.5, u)
    cfg = cfg_of(oc)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    capi.rk3_step(cfg, 0.5, ud, wd)
    I = O.interior(oc)
    e = var_errors(host(ud), u_ref, I)
    assert (e < 1e-12).all(), e
    # mass: the interior total changes only by roundoff (closed boundaries)
    m0 = u[0][I[1:]].sum()
    m1 = host(ud)[0][I[1:]].sum()
    assert abs(m1 - m0) / m0 < 1e-13


def test_sod_200_steps_parity():
    """Configuration 1 (Sod, 256 cells, WENO5 + HLLC, CFL 0.9, 200 steps) on
    the GPU plain path with device-side CFL time steps vs the oracle."""
    oc = O.sod_config(256)
    w, u = O.initialize(oc, O.sod_init(oc))
    cfg = cfg_of(oc)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    dt_d = torch.empty(1, dtype=torch.float64, device=DEV)
    ws = capi.rk3_workspace(cfg, capi.F64)
    u_ref = u.copy()
    for _ in range(200):
        w_ref = O.cons2prim(oc, u_ref)
        dt = O.max_time_step(oc, w_ref, 0.9)
        capi.cons2prim(cfg, ud, wd)
        capi.max_time_step(cfg, wd, 0.9, dt_d)
        dtg = float(dt_d.item())
        assert abs(dtg - dt) <= 1e-12 * dt
        u_ref, _ = O.rk3_step(oc, dt, u_ref)
        capi.rk3_step(cfg, dt, ud, wd, ws)
    I = O.interior(oc)
    e = var_errors(host(ud), u_ref, I, vel_scale=1.0)
    assert (e < 1e-10).all(), e


def test_plain_fp32_close_to_fp64():
    oc, w, u = _straka_small(1)
    u_ref, _ = O.rk3_step(oc, 0.5, u)
    cfg = cfg_of(oc)
    ud = to_dev(u, torch.float32)
    wd = torch.empty_like(ud)
    capi.rk3_step(cfg, 0.5, ud, wd)
    I = O.interior(oc)
    e = var_errors(host(ud), u_ref, I)
    assert (e < 1e-5).all(), e


def test_invalid_arguments_raise():
    oc = O.make_config(nx1=16, nghost=3)
    cfg = cfg_of(oc)
    cfg.implicit_scheme = 4
    u = torch.zeros(cfg.shape, dtype=torch.float64, device=DEV)
    with pytest.raises(capi.SnapxError, match="implicit scheme"):
        capi.cons2prim(cfg, u, u.clone())


def _redo_cases(u):
    """(name, state) pairs covering every branch of the positivity check."""
    nan, inf = float("nan"), float("inf")
    edits = {
        "clean": [],
        "ghost_negative_rho": [((0, 0, 5, 1), -1.0)],
        "zero_rho": [((0, 0, 5, 6), 0.0)],
        "nan_hides_negative_rho": [((0, 0, 5, 6), -2.0), ((0, 0, 7, 7), nan)],
        "negative_energy": [((4, 0, 4, 4), -1.0)],
        "ke_above_energy_screen_passes": [((1, 0, 4, 4), 1e6)],
        "screen_and_pressure": [((1, 0, 4, 4), 1e6), ((4, 0, 6, 6), 0.0)],
        "infinite_momentum": [((2, 0, 3, 3), inf)],
        "last_interior_cell": [((0, 0, -4, -4), -1e-30)],
    }
    for name, ed in edits.items():
        x = u.clone()
        for idx, v in ed:
            x[idx] = v
        yield name, x


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_check_state_matches_reference_rules(dtype):
    import torch_oracle as T

    oc, _, u = _straka_small(0, nx2=16, nx1=12)
    cfg = cfg_of(oc)
    flag = torch.full((1,), -7, dtype=torch.int32, device=DEV)
    for name, x in _redo_cases(torch.from_numpy(u)):
        expect = T.check_redo(oc, x)
        capi.check_state(cfg, x.to(DEV, dtype).contiguous(), flag)
        assert int(flag.item()) == expect, (name, int(flag.item()), expect)


def test_check_state_with_limiter():
    """With the limiter on, cons2prim floors the pressure (> 0), so only the
    density test can ask for a redo; the input is left untouched."""
    oc, _, u = _straka_small(0, nx2=16, nx1=12)
    cfg = cfg_of(oc)
    cfg.limiter = 1
    cfg.pressure_floor = 1e-3
    flag = torch.zeros(1, dtype=torch.int32, device=DEV)
    x = torch.from_numpy(u).to(DEV)
    x[4, 0, 6, 6] = -5.0
    before = x.clone()
    capi.check_state(cfg, x, flag)
    assert int(flag.item()) == 0
    assert torch.equal(x, before)
    x[0, 0, 6, 7] = -1.0
    capi.check_state(cfg, x, flag)
    assert int(flag.item()) == capi.REDO_DENSITY


def test_check_state_in_a_captured_graph():
    """The positivity check clears its flag with a kernel (no memset), so a
    captured graph replays it correctly: three replays on changed states, with
    eager work between them, each against the eager verdict."""
    oc, _, u = _straka_small(0, nx2=16, nx1=12)
    cfg = cfg_of(oc)
    x = torch.from_numpy(u).to(DEV)
    flag = torch.zeros(1, dtype=torch.int32, device=DEV)
    eager = torch.zeros(1, dtype=torch.int32, device=DEV)
    capi.check_state(cfg, x, flag)  # warm up outside the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            capi.check_state(cfg, x, flag)
    torch.cuda.synchronize()
    good = x.clone()
    for edit, expect in (((0, 0, 5, 6), capi.REDO_DENSITY), (None, 0),
                         ((4, 0, 4, 4), capi.REDO_PRESSURE)):
        x.copy_(good)
        if edit is not None:
            x[edit] = -1.0
        y = x * 2.0  # eager work between replays
        del y
        g.replay()
        torch.cuda.synchronize()
        capi.check_state(cfg, x, eager)
        assert int(flag.item()) == int(eager.item()) == expect, (edit, int(flag.item()))


def test_custom_boundaries_plain_path_parity():
    """custom x2 sides (src/bc/bc_func.cpp:4-5): the ghost cells are the
    caller's, and the stage combine and gravity still advance them as in the
    reference; three FP64 steps match the oracle to 1e-12.  The fused path
    refuses such blocks."""
    oc = O.straka_config(nx2=48, nx1=24, implicit_scheme=1)
    oc.x2max = 25.6e3 * 48 / 512 * 2
    oc.bc[2] = oc.bc[3] = O.BC["custom"]
    wp = O.straka_init(oc, xc=2400.0)
    w, u = O.initialize(oc, wp)  # x2 ghosts: the initial profile itself
    u_ref = u
    cfg = cfg_of(oc)
    ud = to_dev(u)
    wd = torch.empty_like(ud)
    for _ in range(3):
        u_ref, _ = O.rk3_step(oc, 0.5, u_ref)
        capi.rk3_step(cfg, 0.5, ud, wd)
    e = var_errors(host(ud), u_ref)  # ghosts included
    assert (e < 1e-12).all(), e
    with pytest.raises(RuntimeError, match="custom boundaries"):
        capi.fused_rk3_step(cfg, 0.5, to_dev(u, torch.float32))
