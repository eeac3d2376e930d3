"""GPU checks of the hand-written adjoint (snapx_adjoint_rk3_step).

The adjoint has no counterpart in the reference; what pins it is the forward
map it differentiates (the FP64 plain path, itself pinned to the oracle):

* dot-product test  <lambda, (S(u + e v) - S(u - e v)) / 2e>  ==  <J^T lambda, v>
  for random directions v and weights lambda, per configuration (central
  differences of the FP64 step; bar 1e-6 relative);
* torch.autograd.gradcheck of the RK3Step node on tiny grids (every column of
  the Jacobian by finite differences against every row by the adjoint);
* the same gradient through a 3-step ``integrate`` chain as the product of
  single-step adjoints (checkpointing restores the stage states exactly).
"""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from snapy_linfel_amd import adjoint as A
from test_gpu_plain import _box3d, _straka_small, cfg_of, to_dev

pytestmark = pytest.mark.gpu


def _step(cfg, dt, u):
    out = u.clone()
    capi.rk3_step(cfg, dt, out, torch.empty_like(out))
    return out


def _central(cfg, dt, u, v, lam, eps):
    return ((_step(cfg, dt, u + eps * v) - _step(cfg, dt, u - eps * v)) * lam).sum() / (2 * eps)


def _scales(u):
    s = u.abs().amax(dim=(1, 2, 3))
    return torch.maximum(s, torch.full_like(s, 1e-3)).view(5, 1, 1, 1)


def _case(name):
    if name.startswith("straka_tall"):
        # columns of two wavefronts (scheme 1: cross-wavefront PCR) and taller
        # than the 5x5 column workgroup (scheme 9: serial block Thomas).  24
        # columns: with 8 the smooth wind puts a face exactly on an upwind
        # switch, where one-sided differences differ (a kink, not an error;
        # libtorch autograd through the torch oracle agrees with the adjoint)
        nx1 = 100 if name.endswith("1") else 150
        oc, w, u = _straka_small(int(name[-1]), nx2=24, nx1=nx1, wind=10.0)
        return oc, u, 0.2
    if name.startswith("straka_mid"):
        # columns of 44 cells: two rows per lane in 32-lane groups (two
        # columns per wavefront).  (At 45 and 48 cells a face of the flow
        # sits on an upwind switch, which central differences straddle; the
        # odd height is in the autograd comparison.)
        oc, w, u = _straka_small(int(name[-1]), nx2=24, nx1=44, wind=10.0)
        return oc, u, 0.2
    if name.startswith("straka"):
        oc, w, u = _straka_small(int(name[6]), nx2=24, nx1=16, wind=10.0)
        return oc, u, 0.2
    if name == "box_lmars_roe":
        oc, w, u = _box3d(n=(10, 8, 6))
        return oc, u, 0.2
    if name == "box_hllc":
        oc, w, u = _box3d(riemann="hllc", n=(10, 8, 6))
        return oc, u, 0.2
    if name == "box_nograv":
        oc, w, u = _box3d(riemann=("hllc", "lmars", "roe"), scheme=0, n=(10, 8, 6))
        oc.grav1 = 0.0
        w, u = O.initialize(oc, w)
        return oc, u, 0.2
    if name.startswith("sph"):
        # spherical-polar sector (r in [1, 2], theta in [pi/4, 3pi/4], phi
        # periodic) of tests/test_gpu_spherical.py: "sph<dims>_<solver><scheme>"
        from test_gpu_spherical import _sector
        dims, rest = name[3:].split("_")
        # 2dtall: 40-cell radial columns (two rows per lane)
        n = {"2d": (14, 10, 1), "2dtall": (40, 10, 1), "3d": (10, 8, 6)}[dims]
        oc, w, u = _sector(rest[:-1], int(rest[-1]), n=n, seed=9)
        return oc, u, 0.01
    if name == "sod":
        oc = O.sod_config(64)
        w, u = O.initialize(oc, _sod_moving(oc))
        return oc, u, O.max_time_step(oc, w, 0.5)
    raise KeyError(name)


def _sod_moving(oc):
    """Sod density / pressure jump with the velocity 0.1 + 0.3 sin(2 pi x).
    HLLC's branches are continuous but not differentiable where the contact
    speed is zero or where the PVRS star pressure equals a side pressure
    (every face of a uniform region at rest, or in uniform motion, sits
    there); the velocity varies from face to face by far more than the
    finite-difference step.  The WENO smoothness indicators of the velocity
    stay well above the WENO epsilon (1e-6), where the weights would be
    strongly nonlinear on the scale of that step."""
    n1 = O.shape(oc)[2]
    x = (np.arange(n1) - oc.nghost + 0.5) / oc.nx1
    w = O.sod_init(oc)
    w[1] = 0.1 + 0.3 * np.sin(2 * np.pi * x)
    return w


def _bc_consistent(cfg, v):
    """Directions are taken in the space of states that satisfy the boundary
    conditions (every state the step sees does): at a reflecting wall the
    face velocity is then exactly zero for all of them, and the upwind branch
    of the solver never switches."""
    capi.apply_bc(cfg, v, capi.CONSERVED)
    return v


CASES = ["straka0", "straka1", "straka9", "straka_mid1", "straka_tall1", "straka_tall9",
         "box_lmars_roe", "box_hllc", "box_nograv", "sod"]
# spherical-polar blocks: per-cell areas and volumes, the geometric sources
# (with the x1 face pressure), per-layer column geometry and ghost distances
SPH_CASES = ["sph3d_hllc1", "sph3d_lmars9", "sph3d_roe0", "sph2d_hllc1", "sph2d_roe9",
             "sph2dtall_hllc1"]


@pytest.mark.parametrize("case", CASES + SPH_CASES)
def test_adjoint_dot_product_vs_central_differences(case):
    oc, u_np, dt = _case(case)
    cfg = cfg_of(oc)
    u = to_dev(u_np)
    g = torch.Generator(device="cpu").manual_seed(7)
    sc = _scales(u)
    ws = capi.adjoint_workspace(cfg)
    for trial in range(2):
        v = _bc_consistent(cfg, torch.randn(u.shape, generator=g, dtype=torch.float64)
                           .to(u.device) * sc)
        lam = torch.randn(u.shape, generator=g, dtype=torch.float64).to(u.device) / sc
        # central differences at eps and 2 eps, Richardson-extrapolated: the
        # O(eps^2) term (1e-5 relative for WENO weights at eps = 1e-6) cancels
        fd = (4 * _central(cfg, dt, u, v, lam, 1e-6) - _central(cfg, dt, u, v, lam, 2e-6)) / 3
        gl = lam.clone()
        capi.adjoint_rk3_step(cfg, dt, u, gl, ws)
        ad = (gl * v).sum()
        rel = abs(float(ad - fd)) / max(abs(float(fd)), 1e-300)
        print(f"{case} trial {trial}: adjoint {float(ad):.12e} fd {float(fd):.12e} rel {rel:.2e}")
        assert rel < 1e-6, (case, trial, float(ad), float(fd), rel)


def _bc_box_tiny(xi, ng=3):
    """Boundary pass of the tiny box as torch ops (x1 reflecting with the
    normal momentum flipped, then x2 periodic over whole rows, the order of
    the library's boundary pass)."""
    sgn = torch.tensor([1.0, -1.0, 1.0, 1.0, 1.0], dtype=xi.dtype, device=xi.device)
    sgn = sgn.view(5, 1, 1, 1)
    lo = xi[..., :ng].flip(-1) * sgn
    hi = xi[..., -ng:].flip(-1) * sgn
    x = torch.cat([lo, xi, hi], dim=-1)
    return torch.cat([x[:, :, -ng:], x, x[:, :, :ng]], dim=2)


def test_adjoint_gradcheck_sod_tiny():
    oc = O.sod_config(12)
    w, u_np = O.initialize(oc, _sod_moving(oc))
    dt = O.max_time_step(oc, w, 0.5)
    cfg = cfg_of(oc)
    u = to_dev(u_np)
    sc = _scales(u)
    # every entry (ghosts included) in scaled variables, so that all are O(1)
    x = (u / sc).detach().requires_grad_(True)

    def f(x):
        return A.rk3_step(x * sc, cfg, dt) / sc

    # eps 1e-7: the O(eps^2) error of central differences through the WENO
    # weights is 1e-5 relative at eps = 1e-6 (it falls as eps^2)
    assert torch.autograd.gradcheck(f, (x,), eps=1e-7, atol=1e-7, rtol=1e-5)


def test_adjoint_gradcheck_box_tiny():
    """x1 reflecting + gravity + 3x3 column solve, x2 periodic, LMARS / Roe;
    the inputs are the interior cells, the boundary pass is part of f."""
    oc, w, u_np = _box3d(n=(6, 4, 1), scheme=1)
    dt = 0.2
    cfg = cfg_of(oc)
    u = to_dev(u_np)
    assert u.shape == (5, 1, 10, 12)
    sc = _scales(u)
    xi = (u[:, :, 3:-3, 3:-3] / sc).detach().requires_grad_(True)

    def f(xi):
        out = A.rk3_step(_bc_box_tiny(xi * sc), cfg, dt)
        return out[:, :, 3:-3, 3:-3] / sc

    # the torch boundary pass reproduces the library's
    full = _bc_box_tiny(u[:, :, 3:-3, 3:-3])
    ref = full.clone()
    capi.apply_bc(cfg, ref, capi.CONSERVED)
    assert torch.equal(full, ref)
    assert torch.autograd.gradcheck(f, (xi,), eps=1e-6, atol=1e-7, rtol=1e-5)


def test_integrate_chain_matches_stepwise_adjoints():
    oc, u_np, dt = _case("straka1")
    cfg = cfg_of(oc)
    u0 = to_dev(u_np)
    g = torch.Generator(device="cpu").manual_seed(3)
    lam = torch.randn(u0.shape, generator=g, dtype=torch.float64).to(u0.device) / _scales(u0)
    x = u0.clone().requires_grad_(True)
    y = A.integrate(x, cfg, dt, 3)
    (grad,) = torch.autograd.grad((y * lam).sum(), x)
    # by hand: forward states, then adjoints in reverse
    us = [u0]
    for _ in range(2):
        us.append(_step(cfg, dt, us[-1]))
    l = lam.clone()
    for k in (2, 1, 0):
        capi.adjoint_rk3_step(cfg, dt, us[k], l)
    assert torch.equal(grad, l)
    # forward values are the plain path's
    assert torch.equal(y.detach(), _step(cfg, dt, us[2]))


def test_adjoint_rejects_limiter():
    oc, u_np, dt = _case("sod")
    oc.limiter = 1
    cfg = cfg_of(oc)
    u = to_dev(u_np)
    with pytest.raises(capi.SnapxError, match="limiter"):
        capi.adjoint_rk3_step(cfg, dt, u, u.clone())


def _box_autograd(n=(24, 32, 32)):
    """The configuration-3 physics on a 32 x 32 x 24 box (x1 vertical)."""
    return _box3d(n=n)


@pytest.mark.parametrize("case", ["box_32x32x24", "straka9", "straka1_45"])
def test_adjoint_matches_torch_autograd(case):
    """The hand-written adjoint (GPU) against libtorch autograd through the
    differentiable restatement of the step (oracle/torch_oracle.py, CPU FP64,
    itself pinned to the C oracle): full gradients, <= 1e-9 of the largest
    entry per variable.  straka1_45: 45-cell 3x3 columns, two rows per lane
    in 32-lane groups with an odd row count."""
    import torch_oracle as T
    if case == "straka9":
        oc, w, u_np = _straka_small(9, nx2=24, nx1=16, wind=10.0)
        dt = 0.2
    elif case == "straka1_45":
        oc, w, u_np = _straka_small(1, nx2=12, nx1=45, wind=10.0)
        dt = 0.2
    else:
        oc, w, u_np = _box_autograd()
        dt = 0.2
    rng = np.random.default_rng(21)
    # a variable that is zero everywhere (v3 in 2-D) gets weights of order 1e3
    scale = np.maximum(np.abs(u_np).reshape(5, -1).max(axis=1), 1e-3).reshape(5, 1, 1, 1)
    lam = rng.standard_normal(u_np.shape) / scale
    x = torch.from_numpy(u_np.copy()).requires_grad_(True)
    y = T.rk3_step(oc, dt, x)
    (g_ref,) = torch.autograd.grad((y * torch.from_numpy(lam)).sum(), x)
    g_ref = g_ref.numpy()
    cfg = cfg_of(oc)
    gl = to_dev(lam)
    capi.adjoint_rk3_step(cfg, dt, to_dev(u_np), gl)
    got = gl.cpu().numpy()
    gmax = np.abs(g_ref).reshape(5, -1).max(axis=1)
    err = np.abs(got - g_ref).reshape(5, -1).max(axis=1) / np.maximum(gmax, gmax.max() * 1e-12)
    print(case, "relative error per variable", err)
    assert (err < 1e-9).all(), err


def test_segment_checkpointing_gives_the_same_gradient():
    """integrate(..., checkpoint_every=3) over 7 steps: the same gradient, bit
    for bit, as keeping every step input."""
    oc, u_np, dt = _case("straka1")
    cfg = cfg_of(oc)
    u0 = to_dev(u_np)
    g = torch.Generator(device="cpu").manual_seed(4)
    lam = torch.randn(u0.shape, generator=g, dtype=torch.float64).to(u0.device) / _scales(u0)
    grads = []
    for k in (None, 3):
        x = u0.clone().requires_grad_(True)
        y = A.integrate(x, cfg, dt, 7, checkpoint_every=k, tape=False)
        (gr,) = torch.autograd.grad((y * lam).sum(), x)
        grads.append(gr)
    assert torch.equal(grads[0], grads[1])


@pytest.mark.parametrize("case", ["straka1", "straka9", "straka_tall9", "box_lmars_roe"])
def test_taped_segments_match_stepwise_adjoints(case):
    """integrate(..., checkpoint_every=3, tape=True) over 7 steps: the
    segments' replay records every step once and sweeps the tapes; the
    gradient equals the step-by-step adjoint to 1e-11 (the replay's states
    come from the recompute's cyclic-reduction column solve and agree with
    the forward's to rounding).  Record + sweep of one step gives the bits
    of snapx_adjoint_rk3_step, and the recorded step output agrees with the
    plain step to 1e-13."""
    oc, u_np, dt = _case(case)
    cfg = cfg_of(oc)
    u0 = to_dev(u_np)
    g = torch.Generator(device="cpu").manual_seed(5)
    lam = torch.randn(u0.shape, generator=g, dtype=torch.float64).to(u0.device) / _scales(u0)
    grads = []
    for tape in (False, True):
        x = u0.clone().requires_grad_(True)
        y = A.integrate(x, cfg, dt, 7, checkpoint_every=3, tape=tape)
        (gr,) = torch.autograd.grad((y * lam).sum(), x)
        grads.append(gr)
    sc = _scales(grads[0])
    err = ((grads[1] - grads[0]).abs() / sc).amax(dim=(1, 2, 3))
    assert (err < 1e-11).all(), err
    # one step, both ways
    tape = torch.empty(capi.adjoint_tape_bytes(cfg), dtype=torch.uint8, device=u0.device)
    work = capi.adjoint_work(cfg, u0.device)
    out = torch.empty_like(u0)
    l1, l2 = lam.clone(), lam.clone()
    capi.adjoint_record(cfg, dt, u0, out, tape, work)
    capi.adjoint_sweep(cfg, dt, tape, l1, work)
    capi.adjoint_rk3_step(cfg, dt, u0, l2)
    torch.cuda.synchronize()
    assert torch.equal(l1, l2)
    ref = _step(cfg, dt, u0)
    rel = ((out - ref).abs() / _scales(ref)).amax()
    assert rel < 1e-13, float(rel)
