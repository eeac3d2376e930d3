"""snapx_fused_rk3_steps_static (the entry point that allocates nothing)
against snapx_fused_rk3_steps with a caller workspace.

Both run the same kernels on the same state, so the results must agree bit
for bit.  Consecutive calls share the library's static per-GPU workspace:
each call here is followed by a smaller one, which must not see the larger
call's leftovers.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from snapy_linfel_amd import capi
from test_gpu_plain import _box3d, _straka_small, cfg_of, to_dev

pytestmark = pytest.mark.gpu


def _pair(oc, u, dt, nsteps):
    cfg = cfg_of(oc)
    ref = to_dev(u, torch.float32)
    got = ref.clone()
    capi.fused_rk3_steps(cfg, dt, nsteps, ref)
    capi.fused_rk3_steps_static(cfg, dt, nsteps, got)
    torch.cuda.synchronize()
    return ref, got


def _tall_column():
    """Columns of 300 cells with 5x5 blocks: the lane-per-column solve, whose
    column scratch lives in the static workspace too."""
    oc = O.straka_config(nx2=8, nx1=300, implicit_scheme=9)
    oc.x2max = 800.0
    oc.x1max = 300 * 20.0
    w, u = O.initialize(oc, O.straka_init(oc))
    return oc, u


def test_static_workspace_matches_caller_workspace():
    cases = []
    oc, _, u = _box3d(n=(40, 36, 28))
    cases.append(("box3d 40x36x28", oc, u, 0.05, 3))  # 33 m cells: acoustic CFL 0.5
    oc, u = _tall_column()
    cases.append(("tall 300x8 scheme 9", oc, u, 0.1, 2))
    oc, _, u = _straka_small(1)
    cases.append(("straka 24x48", oc, u, 0.1, 2))
    oc, _, u = _straka_small(9, nx2=16, nx1=12)
    cases.append(("straka 12x16 scheme 9", oc, u, 0.1, 1))
    sizes = [capi.lib().snapx_fused_workspace_bytes(C.byref(cfg_of(c[1]))) for c in cases]
    assert all(s2 < s1 for s1, s2 in zip(sizes, sizes[1:])), sizes
    for name, oc, u, dt, n in cases:
        print(f"no-alloc case {name}", flush=True)
        ref, got = _pair(oc, u, dt, n)
        assert torch.isfinite(ref).all(), name
        assert torch.equal(ref, got), (name, (ref - got).abs().max().item())


def test_static_workspace_rejects_large_blocks():
    oc = O.make_config(nx1=64, nx2=256, nx3=256, nghost=3, x1max=6400.0, x2max=25600.0,
                       x3max=25600.0, implicit_scheme=1, grav1=-9.81,
                       bc=("reflecting", "reflecting", "periodic", "periodic", "periodic",
                           "periodic"))
    cfg = cfg_of(oc)
    assert capi.lib().snapx_fused_workspace_bytes(C.byref(cfg)) > capi.STATIC_SCRATCH_BYTES
    u = torch.zeros(1, dtype=torch.float32, device="cuda")
    status = capi.lib().snapx_fused_rk3_steps_static(C.byref(cfg), 0.5, 1, capi.ptr(u),
                                                     capi.stream_ptr())
    assert status == capi.ESPACE
    assert b"static" in capi.lib().snapx_last_error()
    with pytest.raises(capi.SnapxError):
        capi.fused_rk3_steps_static(cfg, 0.5, 1, u)
