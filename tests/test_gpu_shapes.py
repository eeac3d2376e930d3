"""GPU: seeded random block shapes, boundary functions, solvers and implicit
schemes, one RK3 step of the FP64 plain path (1e-12) and of the FP32 fused
path (1e-5) against the oracle.

The shapes reach the edges of the tilings: ragged tiles along every axis, last
tiles of one or two cells, single-tile axes, x1 columns from 3 to 70 cells,
and 1-, 2- and 3-D blocks."""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from test_gpu_plain import cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu

SOLVERS = ("lmars", "hllc", "roe")
BCS = ("reflecting", "outflow", "periodic")


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    nx1 = int(rng.integers(3, 71))
    dim = int(rng.choice([1, 2, 2, 3, 3, 3]))
    nx2 = int(rng.integers(3, 41)) if dim >= 2 else 1
    nx3 = int(rng.integers(3, 13)) if dim == 3 else 1
    scheme = int(rng.choice([0, 1, 9]))
    bc = []
    for ax in range(3):
        kind = str(rng.choice(BCS))
        if ax == 0 and kind == "periodic":
            kind = "reflecting"  # a hydrostatic column has walls
        bc += [kind, kind]
    dz = 100.0
    oc = O.make_config(nx1=nx1, nx2=nx2, nx3=nx3, nghost=3, x1min=0.0, x1max=dz * nx1,
                       x2min=0.0, x2max=dz * nx2, x3min=0.0, x3max=dz * nx3, gammad=1.4,
                       weight=8.31446261815324 / 287.0,
                       recon_vertical=("weno5", False, False),
                       recon_horizontal=("weno5", False, bool(rng.integers(0, 2))),
                       riemann=tuple(str(rng.choice(SOLVERS)) for _ in range(3)),
                       grav1=-9.81, implicit_scheme=scheme, bc=tuple(bc))
    n3, n2, n1 = O.shape(oc)
    z = (np.arange(n1) - 3 + 0.5) * dz
    T = 250.0 - 9.81 * z / 1004.5
    p = 1e5 * (T / 250.0) ** (1004.5 / 287.0)
    w = np.zeros((5, n3, n2, n1))
    w[4] = p
    w[0] = p / (287.0 * (T + rng.uniform(-0.5, 0.5, size=(n3, n2, n1))))
    w[1:4] = rng.uniform(-3, 3, size=(3, n3, n2, n1))
    w, u = O.initialize(oc, w)
    dt = 0.5 * O.max_time_step(oc, w, 1.0)
    return oc, u, dt


@pytest.mark.parametrize("seed", range(16))
def test_random_block_one_step(seed):
    oc, u, dt = _case(seed)
    desc = (oc.nx1, oc.nx2, oc.nx3, oc.implicit_scheme, list(oc.bc), list(oc.riemann))
    u_ref, _ = O.rk3_step(oc, dt, u)
    cfg = cfg_of(oc)
    I = O.interior(oc)
    ud = to_dev(u)
    capi.rk3_step(cfg, dt, ud, torch.empty_like(ud))
    e64 = var_errors(host(ud), u_ref, I)
    assert (e64 < 1e-12).all(), (desc, e64)
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_step(cfg, dt, uf)
    e32 = var_errors(host(uf), u_ref, I)
    assert (e32 < 1e-5).all(), (desc, e32)


@pytest.mark.parametrize("scheme", [0, 1, 9])
def test_random_blocks_as_batches(scheme):
    """The same random blocks, grouped by implicit scheme, as members of one
    batch (persistent kernels, default tiles): each member after one step
    against the oracle at the fused bar."""
    cases = [c for c in (_case(s) for s in range(16)) if c[0].implicit_scheme == scheme]
    assert cases
    dt = min(c[2] for c in cases)
    b = capi.Batch([cfg_of(oc) for oc, _, _ in cases], dt)
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, u0, _ in cases])
    b.steps(u, 1)
    for m, (oc, u0, _) in enumerate(cases):
        u_ref, _ = O.rk3_step(oc, dt, u0)
        e = var_errors(host(b.member(u, m)), u_ref, O.interior(oc))
        assert (e < 1e-5).all(), (m, oc.nx1, oc.nx2, oc.nx3, e)


def _sph_case(seed):
    rng = np.random.default_rng(2000 + seed)
    nx1 = int(rng.integers(3, 71))
    dim = int(rng.choice([2, 3]))
    nx2 = int(rng.integers(3, 33))
    nx3 = int(rng.integers(3, 11)) if dim == 3 else 1
    scheme = int(rng.choice([0, 1, 9]))
    t0 = float(rng.uniform(0.1, 0.4)) * np.pi
    t1 = float(rng.uniform(0.6, 0.9)) * np.pi
    bc = ["reflecting", str(rng.choice(["reflecting", "outflow"]))]
    bc += [str(rng.choice(["reflecting", "outflow"]))] * 2
    bc += [str(rng.choice(BCS))] * 2
    oc = O.make_config(nx1=nx1, nx2=nx2, nx3=nx3, nghost=3, x1min=1.0, x1max=2.0, x2min=t0,
                       x2max=t1, x3min=0.0, x3max=float(rng.uniform(0.3, 1.5)),
                       coord="spherical-polar", gammad=1.4,
                       recon_vertical=("weno5", False, False),
                       riemann=tuple(str(rng.choice(SOLVERS)) for _ in range(3)),
                       grav1=-0.5, implicit_scheme=scheme, bc=tuple(bc))
    n3, n2, n1 = O.shape(oc)
    r = O.geometry(oc, "x1v")
    p = np.exp(-0.5 * (r - 1.0))
    w = np.zeros((5, n3, n2, n1))
    w[4] = p * (1.0 + 0.02 * rng.uniform(-1, 1, (n3, n2, n1)))
    w[0] = p * (1.0 + 0.05 * rng.uniform(-1, 1, (n3, n2, n1)))
    w[1:4] = 0.05 * rng.uniform(-1, 1, (3, n3, n2, n1))
    w, u = O.initialize(oc, w)
    return oc, u, 0.5 * O.max_time_step(oc, w, 1.0)


@pytest.mark.parametrize("seed", range(8))
def test_random_spherical_block_one_step(seed):
    oc, u, dt = _sph_case(seed)
    desc = (oc.nx1, oc.nx2, oc.nx3, oc.implicit_scheme, list(oc.bc), list(oc.riemann))
    u_ref, _ = O.rk3_step(oc, dt, u)
    cfg = cfg_of(oc)
    I = O.interior(oc)
    ud = to_dev(u)
    capi.rk3_step(cfg, dt, ud, torch.empty_like(ud))
    e64 = var_errors(host(ud), u_ref, I, vel_scale=1.0)
    assert (e64 < 1e-12).all(), (desc, e64)
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_step(cfg, dt, uf)
    e32 = var_errors(host(uf), u_ref, I, vel_scale=1.0)
    assert (e32 < 1e-5).all(), (desc, e32)


@pytest.mark.parametrize("seed", range(6))
def test_random_block_adjoint_matches_torch_autograd(seed):
    """The hand-written adjoint on random (smaller) blocks against libtorch
    autograd through the differentiable restatement of the step (CPU FP64):
    full gradients, <= 1e-9 of the largest entry per variable.

    LMARS is kept off a horizontal axis with reflecting walls and WENO5
    velocities.  Mirrored states there make the LMARS interface velocity zero up
    to roundoff (exactly 0 on about half the wall faces, below 4e-14 on the
    rest; at least 5e-4 inside), which is exactly where its upwind switch sits.
    The flux is not differentiable there: its one-sided derivatives differ.
    Each implementation takes a side from the roundoff sign of its own WENO5
    sum. Flipping only that tie in the restatement moves the gradient by up to
    1.2e-4 of the largest entry. HLLC keeps the same walls and WENO5 paths in
    the test."""
    import torch_oracle as T
    oc, u_np, dt = _case(100 + seed)
    for ax in (1, 2):
        if (oc.bc[2 * ax] == O.BC["reflecting"] and oc.riemann[ax] == O.RIEMANN["lmars"]
                and oc.recon_shock[1]):
            oc.riemann[ax] = O.RIEMANN["hllc"]
    oc.nx1 = min(oc.nx1, 40)
    oc.nx2 = min(oc.nx2, 20) if oc.nx2 > 1 else 1
    oc.nx3 = min(oc.nx3, 6) if oc.nx3 > 1 else 1
    oc.x1max, oc.x2max, oc.x3max = 100.0 * oc.nx1, 100.0 * oc.nx2, 100.0 * oc.nx3
    n3, n2, n1 = O.shape(oc)
    u_np = np.ascontiguousarray(u_np[:, :n3, :n2, :n1])
    u_np = O.apply_bc(oc, u_np, O.CONSERVED)
    rng = np.random.default_rng(31 + seed)
    scale = np.maximum(np.abs(u_np).reshape(5, -1).max(axis=1), 1e-3).reshape(5, 1, 1, 1)
    lam = rng.standard_normal(u_np.shape) / scale
    x = torch.from_numpy(u_np.copy()).requires_grad_(True)
    y = T.rk3_step(oc, dt, x)
    (g_ref,) = torch.autograd.grad((y * torch.from_numpy(lam)).sum(), x)
    g_ref = g_ref.numpy()
    gl = to_dev(lam)
    capi.adjoint_rk3_step(cfg_of(oc), dt, to_dev(u_np), gl)
    got = host(gl)
    gmax = np.abs(g_ref).reshape(5, -1).max(axis=1)
    err = np.abs(got - g_ref).reshape(5, -1).max(axis=1) / np.maximum(gmax, gmax.max() * 1e-12)
    assert (err < 1e-9).all(), ((oc.nx1, oc.nx2, oc.nx3, oc.implicit_scheme, list(oc.bc)), err)
