"""GPU: time steps read from device memory, and the CFL loop with redo run
on the device (snapx_rk3_attempts / snapx_fused_attempts), captured in one
HIP graph, against the host loop (tests/cfl_loop.py) around the FP64 oracle.

Bars: every dt of the loop within 1e-12 (relative) of the oracle's, the
state within 1e-10 (scales as in test_gpu_plain, velocity scale 1 for the
Sod tube), the loop counters exactly; the device-dt entry points give the
same bits as their host-dt forms."""
import math

import numpy as np
import pytest
import torch

import oracle as O
import torch_oracle as T
from cfl_loop import DENSITY, DT_FROM_U, DT_FROM_W, NONFINITE, PRESSURE, Loop
from snapy_linfel_amd import capi
from test_gpu_plain import DEV, cfg_of, host, to_dev, var_errors

pytestmark = pytest.mark.gpu


def _sod():
    oc = O.sod_config(256)
    w, u = O.initialize(oc, O.sod_init(oc))
    return oc, w, u


def _oracle_loop(oc, loop, n, u, w):
    verdict = lambda un: T.check_redo(oc, torch.from_numpy(un))
    return loop.run(n, u, w, lambda dt, x: O.rk3_step(oc, dt, x),
                    lambda x: O.max_time_step(oc, x, 1.0), verdict,
                    lambda x: O.cons2prim(oc, x))


def _check_state_record(st, loop):
    assert (st.cycle, st.redo, st.status, st.steps, st.redos) == (
        loop.cycle, loop.redo, loop.status, loop.steps, loop.redos)
    assert abs(st.time - loop.time) <= 1e-12 * max(loop.time, 1e-300)


def _check_dt(got, ref):
    ref = np.asarray(ref)
    assert got.shape == ref.shape
    assert np.array_equal(got == 0, ref == 0)
    nz = ref != 0
    assert (np.abs(got[nz] - ref[nz]) <= 1e-12 * ref[nz]).all(), (got, ref)


def _capture(fn):
    """fn() captured in a torch CUDA graph on a side stream."""
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fn()
    torch.cuda.synchronize()
    return g


def test_step_dt_entry_points_match_host_dt():
    """rk3_step_dt / fused_rk3_step_dt / batch set_dt give the bits of the
    host-dt calls (plain FP64 and FP32 with both implicit schemes)."""
    oc = O.straka_config(nx2=24, nx1=16, implicit_scheme=1)
    oc.x2max = 25.6e3 * 24 / 512 * 2
    w0, u0 = O.initialize(oc, O.straka_init(oc))
    dt = 0.37
    dtd = torch.tensor([dt], dtype=torch.float64, device=DEV)
    for scheme in (1, 9):
        cfg = cfg_of(oc)
        cfg.implicit_scheme = scheme
        for dtype in (torch.float64, torch.float32):
            a, b = to_dev(u0, dtype), to_dev(u0, dtype)
            wa, wb = torch.empty_like(a), torch.empty_like(b)
            capi.rk3_step(cfg, dt, a, wa)
            capi.rk3_step_dt(cfg, dtd, b, wb)
            torch.cuda.synchronize()
            assert torch.equal(a, b) and torch.equal(wa, wb), (scheme, dtype)
        a, b = to_dev(u0, torch.float32), to_dev(u0, torch.float32)
        capi.fused_rk3_step(cfg, dt, a)
        capi.fused_rk3_step_dt(cfg, dtd, b)
        torch.cuda.synchronize()
        assert torch.equal(a, b), scheme
    # batch: two members planned with dt = 1, then both set to dt on the device
    cfg = cfg_of(oc)
    bt = capi.Batch([cfg, cfg], 1.0)
    ref = capi.Batch([cfg, cfg], dt)
    ub = torch.cat([to_dev(u0, torch.float32).flatten()] * 2)
    ur = ub.clone()
    bt.set_dt(torch.tensor([dt, dt], dtype=torch.float64, device=DEV))
    bt.steps(ub, 2)
    ref.steps(ur, 2)
    torch.cuda.synchronize()
    assert torch.equal(ub, ur)


def test_cfl_loop_graph_matches_oracle_sod():
    """Sod tube (256 cells, WENO5 + HLLC, CFL 0.9), the reference's loop
    (dt from the last stage's primitives): 40 attempts captured in one graph,
    replayed twice with eager work between the replays."""
    oc, w, u = _sod()
    cfg = cfg_of(oc)
    n = 40
    ud, wd = to_dev(u), to_dev(w)
    loop = capi.StepLoop(cfg, 0.9)
    log = torch.zeros(n, dtype=torch.float64, device=DEV)
    loop.prepare(ud)
    g = _capture(lambda: loop.attempts(ud, wd, n, DT_FROM_W, log))
    ref = Loop(0.9)
    ur, wr = u, w
    scratch = torch.empty_like(ud)
    for rep in range(2):
        g.replay()
        capi.cons2prim(cfg, ud, scratch)  # eager work between replays
        torch.cuda.synchronize()
        ref.dt_log = []
        ur, wr = _oracle_loop(oc, ref, n, ur, wr)
        _check_dt(host(log), ref.dt_log)
        _check_state_record(loop.state(), ref)
        I = O.interior(oc)
        e = var_errors(host(ud), ur, I, vel_scale=1.0)
        assert (e < 1e-10).all(), (rep, e)
        ew = var_errors(host(wd), wr, I, vel_scale=1.0)
        assert (ew < 1e-10).all(), (rep, ew)
    assert ref.redos == 0 and ref.steps == 2 * n


def test_forced_redo_in_graph_matches_oracle_sod():
    """Sod tube at CFL 3 with non-finite states also asking for a redo and
    dt from the attempt's input: steps at CFL 3 blow up (NaN in some cells),
    the redo restores the input and halves dt; 12 attempts in one graph."""
    oc, w, u = _sod()
    cfg = cfg_of(oc)
    n = 12
    mask = DENSITY | PRESSURE | NONFINITE
    ud, wd = to_dev(u), to_dev(w)
    loop = capi.StepLoop(cfg, 3.0, redo_mask=mask)
    log = torch.zeros(n, dtype=torch.float64, device=DEV)
    loop.prepare(ud)
    g = _capture(lambda: loop.attempts(ud, wd, n, DT_FROM_U, log))
    g.replay()
    torch.cuda.synchronize()
    ref = Loop(3.0, redo_mask=mask, dt_source=DT_FROM_U)
    ur, wr = _oracle_loop(oc, ref, n, u, w)
    print("dt", ref.dt_log, "redos", ref.redos, "steps", ref.steps)
    assert ref.redos >= 3 and max(ref.dt_log) / min(ref.dt_log) >= 4  # a double redo
    _check_dt(host(log), ref.dt_log)
    _check_state_record(loop.state(), ref)
    e = var_errors(host(ud), ur, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-10).all(), e


def test_redo_limit_and_stop_limits():
    """More than max_redo redos in a row end the loop with the failed step
    kept; attempts after that, and after nlim / tlim, change nothing."""
    oc, w, u = _sod()
    cfg = cfg_of(oc)
    mask = DENSITY | PRESSURE | NONFINITE
    cases = [dict(cfl=50.0, max_redo=2, redo_mask=mask, dt_source=DT_FROM_U),
             dict(cfl=0.9, nlim=5), dict(cfl=0.9, tlim=0.006)]
    for kw in cases:
        src = kw.pop("dt_source", DT_FROM_W)
        n = 9
        ud, wd = to_dev(u), to_dev(w)
        loop = capi.StepLoop(cfg, **kw)
        log = torch.zeros(n, dtype=torch.float64, device=DEV)
        loop.attempts(ud, wd, n, src, log)
        ref = Loop(dt_source=src, **kw)
        ur, wr = _oracle_loop(oc, ref, n, u, w)
        assert ref.status != 0, kw
        _check_dt(host(log), ref.dt_log)
        _check_state_record(loop.state(), ref)
        got = host(ud)
        I = O.interior(oc)
        if ref.status < 0:
            # the failed step is kept: same NaN cells, the rest close
            nan = np.isnan(ur)
            assert np.array_equal(np.isnan(got), nan) and nan.any()
            assert np.abs(got[~nan] - ur[~nan]).max() <= 1e-8 * np.abs(ur[~nan]).max()
        else:
            e = var_errors(got, ur, I, vel_scale=1.0)
            assert (e < 1e-10).all(), (kw, e)
            ew = var_errors(host(wd), wr, I, vel_scale=1.0)
            assert (ew < 1e-10).all(), (kw, ew)


@pytest.mark.parametrize("dtype", [torch.float32])
def test_plain_fp32_loop_close_to_oracle(dtype):
    """The plain loop in FP32 storage: dt within 1e-5 of the oracle's."""
    oc, w, u = _sod()
    cfg = cfg_of(oc)
    n = 20
    ud, wd = to_dev(u, dtype), to_dev(w, dtype)
    loop = capi.StepLoop(cfg, 0.9)
    log = torch.zeros(n, dtype=torch.float64, device=DEV)
    loop.attempts(ud, wd, n, DT_FROM_W, log)
    ref = Loop(0.9)
    ur, _ = _oracle_loop(oc, ref, n, u, w)
    d = host(log)
    assert (np.abs(d - ref.dt_log) <= 1e-5 * np.asarray(ref.dt_log)).all()
    e = var_errors(host(ud), ur, O.interior(oc), vel_scale=1.0)
    assert (e < 1e-4).all(), e


def test_fused_loop_graph_straka():
    """The fused FP32 loop (dt from the attempt's input) on a small Straka
    block: 8 attempts in a graph, replayed twice; against eager fused
    attempts bit for bit, and against the FP64 oracle loop to 1e-5."""
    oc = O.straka_config(nx2=48, nx1=24, implicit_scheme=1)
    oc.x2max = 25.6e3 * 48 / 512 * 2
    w, u = O.initialize(oc, O.straka_init(oc))
    cfg = cfg_of(oc)
    n = 8
    ug = to_dev(u, torch.float32)
    ue = ug.clone()
    lg = capi.StepLoop(cfg, 0.9)
    le = capi.StepLoop(cfg, 0.9)
    logg = torch.zeros(n, dtype=torch.float64, device=DEV)
    loge = torch.zeros(n, dtype=torch.float64, device=DEV)
    lg.prepare(ug, fused=True)
    g = _capture(lambda: lg.fused_attempts(ug, n, logg))
    ref = Loop(0.9, dt_source=DT_FROM_U)
    ur, wr = u, w
    for rep in range(2):
        g.replay()
        le.fused_attempts(ue, n, loge)
        torch.cuda.synchronize()
        assert torch.equal(ug, ue) and torch.equal(logg, loge), rep
        ref.dt_log = []
        ur, wr = _oracle_loop(oc, ref, n, ur, wr)
        d = host(logg)
        assert (np.abs(d - ref.dt_log) <= 1e-5 * np.asarray(ref.dt_log)).all(), (d, ref.dt_log)
        st = lg.state()
        assert (st.cycle, st.steps, st.redos, st.status) == (ref.cycle, ref.steps, 0, 0)
        e = var_errors(host(ug), ur, O.interior(oc))
        assert (e < 1e-5).all(), (rep, e)
