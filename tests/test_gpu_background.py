"""FP32 perturbation path about an FP64 hydrostatic background
(snapx_fused_background / split / rk3_steps_bg / join) against the FP64
oracle stepping the total state.

Bars:
* every variable within 1e-5 of the oracle after the steps (scale as in
  test_gpu_plain: momentum by max(rho) * 300 m/s), like the fused path;
* on top of that, each momentum within 1e-5 of its own largest value, where
  the fused path on the total state is off by more (its FP32 storage of the
  hydrostatic pressure, 0.01 Pa on 1e5 Pa, drives spurious motion);
* a background at rest stays at rest: max |v| after 100 steps below 1e-7 m/s
  (total-state FP32: about 1e-4 m/s).
"""
import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from test_gpu_plain import _box3d, cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu


def _own_norm(got, ref, I):
    """Per variable max |got - ref| / max |ref| (a field that is zero in the
    oracle, such as the third momentum of a 2-D block, must stay zero)."""
    out = []
    for v in range(5):
        d = np.abs(got[v][I] - ref[v][I]).max()
        m = np.abs(ref[v][I]).max()
        out.append(d / m if m > 0 else d)
    return np.array(out)


def _straka(nx2=48, nx1=24, scheme=1):
    oc = O.straka_config(nx2=nx2, nx1=nx1, implicit_scheme=scheme)
    oc.x2max = 25.6e3 * nx2 / 512 * 2
    w, u = O.initialize(oc, O.straka_init(oc))
    wb = O.straka_init(oc, dT=0.0)
    return oc, u, wb[0, 0, 0].copy(), wb[4, 0, 0].copy()


def _box_background(oc):
    n1 = O.shape(oc)[2]
    z = (np.arange(n1) - 3 + 0.5) * (oc.x1max - oc.x1min) / oc.nx1
    T = 250.0 - 9.81 * z / 1004.5
    p = 1e5 * (T / 250.0) ** (1004.5 / 287.0)
    return p / (287.0 * T), p


def _run_bg(oc, u, rb, pb, dt, n):
    cfg = cfg_of(oc)
    bg = capi.Background(cfg, rb, pb)
    u64 = to_dev(u)
    up = bg.split(u64)
    bg.steps(dt, n, up)
    out = torch.empty_like(u64)
    bg.join(up, out)
    return host(out)


def _run_fused(oc, u, dt, n):
    uf = to_dev(u, torch.float32)
    capi.fused_rk3_steps(cfg_of(oc), dt, n, uf)
    return host(uf)


@pytest.mark.parametrize("scheme", [0, 1, 9])
def test_straka_background_path_vs_oracle(scheme):
    oc, u, rb, pb = _straka(scheme=scheme)
    dt, n = 0.2, 3
    ref = u
    for _ in range(n):
        ref, _ = O.rk3_step(oc, dt, ref)
    got = _run_bg(oc, u, rb, pb, dt, n)
    I = O.interior(oc)
    e = var_errors(got, ref, I)
    assert (e < 1e-5).all(), e
    own = _own_norm(got, ref, I[1:])
    tot = _own_norm(_run_fused(oc, u, dt, n), ref, I[1:])
    print("own-norm errors: background", own, "total", tot)
    assert (own < 1e-5).all(), own
    # the join applies the reference's boundary pass
    assert np.array_equal(O.apply_bc(oc, got, O.CONSERVED), got)


@pytest.mark.parametrize("riemann", [("lmars", "roe", "roe"), "hllc", "lmars", "roe"])
def test_box3d_background_path_vs_oracle(riemann):
    """Merged 16x4x2 tiles on a small 3-D box, every solver."""
    oc, w, u = _box3d(riemann=riemann, n=(20, 14, 12))
    rb, pb = _box_background(oc)
    dt, n = 0.2, 2
    ref = u
    for _ in range(n):
        ref, _ = O.rk3_step(oc, dt, ref)
    got = _run_bg(oc, u, rb, pb, dt, n)
    I = O.interior(oc)
    e = var_errors(got, ref, I)
    assert (e < 1e-5).all(), e
    own = _own_norm(got, ref, I[1:])
    print("own-norm errors", own)
    assert (own < 1e-5).all(), own


@pytest.mark.parametrize("n", [(70, 65, 460), (33, 65, 130)], ids=["slab16", "slab4"])
def test_marching_background_path_vs_oracle(n):
    """x3-marching K1 on the perturbation path with two rows per lane in the
    column solve, one step: 70 x 65 x 460 cells (16-slice slabs, 783
    workgroups, 70-cell columns) and 33 x 65 x 130 cells (ragged 4-slice
    slabs, 594 workgroups, 33-cell columns of 17 lanes)."""
    oc, w, u = _box3d(n=n)
    oc.x1max, oc.x2max, oc.x3max = 100.0 * n[0], 100.0 * n[1], 100.0 * n[2]
    w, u = O.initialize(oc, w)
    rb, pb = _box_background(oc)
    dt = 0.2
    ref, _ = O.rk3_step(oc, dt, u)
    got = _run_bg(oc, u, rb, pb, dt, 1)
    I = O.interior(oc)
    e = var_errors(got, ref, I)
    assert (e < 1e-5).all(), e
    own = _own_norm(got, ref, I[1:])
    print("own-norm errors", own)
    assert (own < 1e-5).all(), own


def test_rest_state_stays_at_rest():
    """The Straka background without the bubble, 100 steps of 0.1 s."""
    oc = O.straka_config(nx2=32, nx1=64, implicit_scheme=1)
    oc.x2max = 3200.0
    wb = O.straka_init(oc, dT=0.0)
    w, u = O.initialize(oc, wb)
    rb, pb = wb[0, 0, 0].copy(), wb[4, 0, 0].copy()
    got = _run_bg(oc, u, rb, pb, 0.1, 100)
    I = O.interior(oc)
    vmax = max(np.abs(got[v][I[1:]] / got[0][I[1:]]).max() for v in (1, 2))
    tot = _run_fused(oc, u, 0.1, 100)
    vtot = max(np.abs(tot[v][I[1:]] / tot[0][I[1:]]).max() for v in (1, 2))
    print("max |v| after 100 steps: background", vmax, "total", vtot)
    assert vmax < 1e-7, vmax


def test_background_rejects_unsupported_options():
    oc, u, rb, pb = _straka()
    cfg = cfg_of(oc)
    cfg.recon_shock[0] = 1
    with pytest.raises(capi.SnapxError):
        capi.Background(cfg, rb, pb)
    cfg = cfg_of(oc)
    with pytest.raises(capi.SnapxError):
        capi.Background(cfg, rb[:-1], pb[:-1])
    bg = capi.Background(cfg, rb, pb)
    with pytest.raises(capi.SnapxError):
        bg.steps(0.1, 1, torch.zeros(10, dtype=torch.float32, device="cuda"))
