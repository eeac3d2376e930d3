"""GPU: batched members of uneven size in one persistent launch per kernel
reproduce the single-block fused path exactly, and the oracle to the FP32
bar."""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from test_gpu_plain import cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu


def _members(seed=7, n=5, riemann="hllc"):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx2 = int(rng.choice([16, 24, 40]))
        nx1 = int(rng.choice([8, 12, 20]))
        oc = O.straka_config(nx2=nx2, nx1=nx1, implicit_scheme=1, riemann=riemann)
        oc.x2max = 50.0 * nx2
        oc.x1max = 50.0 * nx1
        w = O.straka_init(oc, xc=float(rng.uniform(0, 50.0 * nx2)), xr=float(rng.uniform(250, 600)),
                          zc=25.0 * nx1, zr=float(rng.uniform(150, 300)), dT=0.5)
        w, u = O.initialize(oc, w)
        out.append((oc, u))
    return out


def test_batch_matches_single_block_fused_exactly():
    mem = _members()
    cfgs = [cfg_of(oc) for oc, _ in mem]
    b = capi.Batch(cfgs, 0.25)
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, u0 in mem])
    b.steps(u, 3)
    for m, (oc, u0) in enumerate(mem):
        single = to_dev(u0, torch.float32)
        capi.fused_rk3_steps(cfgs[m], 0.25, 3, single)
        assert torch.equal(b.member(u, m), single), m


def test_batch_vs_oracle_one_step():
    mem = _members(seed=11, n=4)
    cfgs = [cfg_of(oc) for oc, _ in mem]
    b = capi.Batch(cfgs, 0.25, tile=(8, 4, 1))
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, u0 in mem])
    b.steps(u, 1)
    for m, (oc, u0) in enumerate(mem):
        ref, _ = O.rk3_step(oc, 0.25, u0)
        e = var_errors(host(b.member(u, m)), ref, O.interior(oc))
        assert (e < 1e-5).all(), (m, e)


def test_batch_rejects_mixed_schemes():
    a, _ = _members(n=1)[0]
    b = O.straka_config(nx2=16, nx1=8, implicit_scheme=9)
    with pytest.raises(capi.SnapxError, match="same implicit scheme"):
        capi.Batch([cfg_of(a), cfg_of(b)], 0.25)


def test_batch_autotune_keeps_state_and_results():
    """The autotuner times RHS launches only (u unchanged) and the batch it
    plans reproduces the single-block fused path exactly."""
    mem = _members(seed=5, n=4)
    cfgs = [cfg_of(oc) for oc, _ in mem]
    b = capi.Batch(cfgs, 0.25)
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, u0 in mem])
    u_before = u.clone()
    tile, ms = b.autotune(u, 0.25, trials=2)
    assert torch.equal(u, u_before)
    assert ms > 0 and tile[0] >= 16 and tile[2] == 1
    b.steps(u, 2)
    for m, (oc, u0) in enumerate(mem):
        single = to_dev(u0, torch.float32)
        capi.fused_rk3_steps(cfgs[m], 0.25, 2, single)
        e = var_errors(host(b.member(u, m)), host(single), O.interior(oc))
        # a different tile changes no arithmetic of a cell
        assert (e == 0).all(), (m, tile, e)


def test_batch_graph_replays_match_eager():
    """Replays of a captured 2-step batch graph, with eager work and
    synchronisations in between, each equal the eager steps."""
    mem = _members(seed=9, n=3)
    cfgs = [cfg_of(oc) for oc, _ in mem]
    b = capi.Batch(cfgs, 0.25)
    u0 = torch.cat([to_dev(u, torch.float32).reshape(-1) for _, u in mem])
    ug = u0.clone()
    g = b.graph(ug, 2)
    ue = u0.clone()
    for rep in range(3):
        g.launch()
        b.steps(ue, 2)
        torch.cuda.synchronize()
        assert torch.equal(ug, ue), rep
        _ = (ue * 2.0).sum().item()  # unrelated eager work between replays


def test_fused_graph_replays_match_eager():
    oc, u_np = _members(seed=3, n=1)[0]
    cfg = cfg_of(oc)
    ug = to_dev(u_np, torch.float32)
    ue = ug.clone()
    g = capi.fused_graph(cfg, 0.25, 3, ug)
    for rep in range(3):
        g.launch()
        capi.fused_rk3_steps(cfg, 0.25, 3, ue)
        torch.cuda.synchronize()
        assert torch.equal(ug, ue), rep


def test_batch_mixed_column_heights_exact():
    """Members of every column form in one launch reproduce the single-block
    path exactly: 20-cell columns (one row per lane, 64 lanes), 40- and
    64-cell columns (two rows per lane, 32-lane groups, two columns per
    wavefront), 90 and 150 cells (two rows per lane, 64 and 128 lanes)."""
    mem = []
    for nx2, nx1 in ((12, 20), (8, 90), (6, 150), (10, 64), (9, 40)):
        oc = O.straka_config(nx2=nx2, nx1=nx1, implicit_scheme=1, riemann="hllc")
        oc.x2max = 50.0 * nx2
        oc.x1max = 50.0 * nx1
        w = O.straka_init(oc, xc=25.0 * nx2, xr=300.0, zc=25.0 * nx1, zr=200.0, dT=0.5)
        w, u = O.initialize(oc, w)
        mem.append((oc, u))
    cfgs = [cfg_of(oc) for oc, _ in mem]
    b = capi.Batch(cfgs, 0.25)
    u = torch.cat([to_dev(u0, torch.float32).reshape(-1) for _, u0 in mem])
    b.steps(u, 2)
    for m, (oc, u0) in enumerate(mem):
        single = to_dev(u0, torch.float32)
        capi.fused_rk3_steps(cfgs[m], 0.25, 2, single)
        assert torch.equal(b.member(u, m), single), m


@pytest.mark.parametrize("n", [(37, 21, 45), (33, 65, 460), (33, 65, 130)],
                         ids=["tiled", "marching", "marching4"])
def test_fused_3d_marching_rhs_matches_tiled_rhs(n):
    """3-D blocks: the single-block path's RHS kernel (merged 16x4x2 tiles for
    the small block with 72 (tile, 4-slice slab) workgroups, the x3-marching
    kernel with 8-slice slabs for 33 x 65 x 460 (1044 workgroups) and with
    4-slice slabs for 33 x 65 x 130 (594)) and the batch path's tiled RHS
    kernel perform the same operations in
    the same order; the compiler contracts a few multiply-adds differently in
    the kernels, so they agree to FP32 rounding (not bit for bit), and both
    meet the oracle's FP32 bar.  Ragged tiles in every direction, reflecting
    and outflow x3 ends."""
    oc = O.make_config(nx1=n[0], nx2=n[1], nx3=n[2], nghost=3, x1min=0.0, x1max=100.0 * n[0],
                       x2min=0.0, x2max=100.0 * n[1], x3min=0.0, x3max=100.0 * n[2], gammad=1.4,
                       weight=8.31446261815324 / 287.0, recon_vertical=("weno5", False, False),
                       riemann=("lmars", "roe", "hllc"), grav1=-9.81, implicit_scheme=1,
                       bc=("reflecting", "reflecting", "periodic", "periodic", "reflecting",
                           "outflow"))
    n3, n2, n1 = O.shape(oc)
    rng = np.random.default_rng(5)
    z = (np.arange(n1) - 3 + 0.5) * 100.0
    T = 250.0 - 9.81 * z / 1004.5
    p = 1e5 * (T / 250.0) ** (1004.5 / 287.0)
    w = np.zeros((5, n3, n2, n1))
    w[4] = p
    w[0] = p / (287.0 * (T + rng.uniform(-0.5, 0.5, size=(n3, n2, n1))))
    w[1:4] = rng.uniform(-2, 2, size=(3, n3, n2, n1))
    w, u0 = O.initialize(oc, w)
    cfg = cfg_of(oc)
    single = to_dev(u0, torch.float32)
    capi.fused_rk3_steps(cfg, 0.3, 2, single)
    b = capi.Batch([cfg], 0.3)
    ub = to_dev(u0, torch.float32).reshape(-1).clone()
    b.steps(ub, 2)
    I = O.interior(oc)
    e_mb = var_errors(host(b.member(ub, 0)), host(single), I)
    assert (e_mb < 1e-6).all(), e_mb
    u_ref = u0
    for _ in range(2):
        u_ref, _ = O.rk3_step(oc, 0.3, u_ref)
    e = var_errors(host(single), u_ref, O.interior(oc))
    assert (e < 1e-5).all(), e
