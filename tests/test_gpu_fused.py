"""GPU checks of the fused FP32 path (snapx_fused_rk3_step) against the FP64
oracle.

Bar: after one RK3 step from identical
state, per conserved variable max |fp32 - fp64| / scale <= 1e-5 (scale as in
test_gpu_plain); after N steps the per-field L2 / Linf relative errors are
reported and bounded; interior mass is conserved to FP32 roundoff.
"""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from test_gpu_plain import _box3d, _straka_small, cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["straka0", "straka1", "straka9", "box_lmars_roe",
                                  "box_hllc", "sod"])
def test_fused_one_step_vs_fp64_oracle(case):
    if case.startswith("straka"):
        oc, w, u = _straka_small(int(case[-1]))
        dt = 0.5
    elif case == "box_hllc":
        oc, w, u = _box3d(riemann="hllc", scheme=1)
        dt = 0.5
    elif case == "sod":
        oc = O.sod_config(256)
        w, u = O.initialize(oc, O.sod_init(oc))
        dt = O.max_time_step(oc, w, 0.9)
    else:
        oc, w, u = _box3d()
        dt = 0.5
    u_ref, _ = O.rk3_step(oc, dt, u)
    cfg = cfg_of(oc)
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_step(cfg, dt, uf)
    got = host(uf)
    I = O.interior(oc)
    e = var_errors(got, u_ref, I)
    assert (e < 1e-5).all(), e
    # the stored ghosts match the reference boundary pass
    ghosts = O.apply_bc(oc, got, O.CONSERVED)
    assert np.array_equal(ghosts, got)


def test_fused_matches_plain_fp32_closely():
    oc, w, u = _straka_small(1)
    cfg = cfg_of(oc)
    a = to_dev(u, torch.float32)
    b = a.clone()
    capi.fused_rk3_step(cfg, 0.5, a)
    capi.rk3_step(cfg, 0.5, b, torch.empty_like(b))
    I = O.interior(oc)
    e = var_errors(host(a), host(b), I)
    assert (e < 1e-5).all(), e


def test_fused_straka_multistep_error_and_mass():
    """50 steps of a reduced Straka block (dt = 0.2 s; the horizontal
    acoustic CFL limit of this 100 m grid is 0.26 s): per-field relative L2
    and Linf error vs FP64 stay below 1e-4 and mass drifts at FP32
    roundoff."""
    oc, w, u = _straka_small(1, nx2=64, nx1=32)
    cfg = cfg_of(oc)
    uf = to_dev(u, torch.float32)
    ws = capi.fused_workspace(cfg)
    u_ref = u.copy()
    for _ in range(50):
        capi.fused_rk3_step(cfg, 0.2, uf, ws)
        u_ref, _ = O.rk3_step(oc, 0.2, u_ref)
    got = host(uf)
    I = O.interior(oc)
    e = var_errors(got, u_ref, I)
    assert (e < 1e-4).all(), e
    for v in (0, 4):
        l2 = np.linalg.norm(got[v][I[1:]] - u_ref[v][I[1:]]) / np.linalg.norm(u_ref[v][I[1:]])
        assert l2 < 1e-5, (v, l2)
    m0 = u[0][I[1:]].sum()
    m1 = got[0][I[1:]].astype(np.float64).sum()
    assert abs(m1 - m0) / m0 < 1e-6


@pytest.mark.parametrize("scheme", [1, 9])
@pytest.mark.parametrize("nx1", [64, 65, 99, 100, 128, 200, 300])
def test_fused_column_heights(nx1, scheme):
    """Columns of one wavefront, of several (cross-wavefront PCR barriers and
    column sums), and taller than the column workgroup (the lane-per-column
    block Thomas kernel), with a ragged x2 tile (6 columns, tiles of 4).
    Scheme 1 at 65-128 cells: one wavefront of two rows per lane (65 and 99
    cells leave an odd row count after the in-lane reduction)."""
    oc, w, u = _straka_small(scheme, nx2=6, nx1=nx1, wind=10.0)  # dx1 = 6.4 km / nx1
    dt = 0.2
    u_ref, _ = O.rk3_step(oc, dt, u)
    cfg = cfg_of(oc)
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_step(cfg, dt, uf)
    e = var_errors(host(uf), u_ref, O.interior(oc))
    assert (e < 1e-5).all(), (nx1, scheme, e)


def test_fused_hydrostatic_rest_state_drift():
    """The Straka background at rest (no bubble): the FP32 fused path stays at
    rest to the FP32 storage rounding of the total state (measured 1.2e-4 m/s
    after 100 steps; bar 1e-3), the FP64 plain path to roundoff (bar 1e-9)."""
    oc = O.straka_config(nx2=64, nx1=64, implicit_scheme=1)
    oc.x2max = 25.6e3 * 64 / 512
    w0, u0 = O.initialize(oc, O.straka_init(oc, dT=0.0))
    cfg = cfg_of(oc)
    I = O.interior(oc)[1:]
    u32 = to_dev(u0, torch.float32)
    capi.fused_rk3_steps(cfg, 0.1, 100, u32)
    u64 = to_dev(u0)
    w64 = torch.empty_like(u64)
    for _ in range(100):
        capi.rk3_step(cfg, 0.1, u64, w64)
    for u, bar in ((host(u32), 1e-3), (host(u64), 1e-9)):
        vmax = max(np.abs(u[v][I] / u[0][I]).max() for v in (1, 2, 3))
        assert vmax < bar, (vmax, bar)


@pytest.mark.parametrize("nx1", [65, 66])
def test_fused_ragged_last_x1_tile_refill(nx1):
    """Regression: 2-D blocks on 32x8 tiles (one face pass per direction),
    with a last x1 tile of one or two cells.  The tile below it has the
    physical x1 ghosts in its 3-cell halo, and those must get the isentropic
    refill too.  A single step against the FP64 oracle, at the fused bar."""
    oc = O.straka_config(nx2=2048, nx1=nx1, implicit_scheme=1)
    oc.x1max = 100.0 * nx1
    oc.x2max = 50.0 * 2048
    w, u = O.initialize(oc, O.straka_init(oc, xc=51200.0))
    u_ref, _ = O.rk3_step(oc, 0.2, u)
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_step(cfg_of(oc), 0.2, uf)
    e = var_errors(host(uf), u_ref, O.interior(oc))
    assert (e < 1e-5).all(), e


@pytest.mark.parametrize("nx1", [66, 99, 128])
def test_fused_packed_columns_full_workgroups(nx1):
    """2048 columns, so that K2 runs workgroups of four one-wavefront
    columns of two rows per lane (one step against the FP64 oracle, the 1e-5
    bar)."""
    oc, w, u = _straka_small(1, nx2=2048, nx1=nx1, wind=10.0)
    dt = 0.2
    u_ref, _ = O.rk3_step(oc, dt, u)
    cfg = cfg_of(oc)
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_step(cfg, dt, uf)
    e = var_errors(host(uf), u_ref, O.interior(oc))
    assert (e < 1e-5).all(), (nx1, e)
