"""torch.autograd binding of the hand-written discrete adjoint.

``RK3Step`` is one FP64 plain-path SSP-RK3 step (``snapx_rk3_step``, the
stage loop of ``MeshBlockImpl::advance_local``, src/mesh/meshblock.cpp:577-786)
as an autograd node.  Its backward calls ``snapx_adjoint_rk3_step``, which
recomputes the three stage inputs from the saved step input and sweeps the
stages in reverse.  Checkpointing is therefore per step: a trajectory of K
steps keeps K step inputs alive (5 x cells doubles each) and never the stage
states, primitives, interface states or fluxes.

``integrate(u, cfg, dt, nsteps)`` chains steps so that
``torch.autograd.grad(loss(integrate(u0, ...)), u0)`` is the exact gradient of
the discrete map (limiter off, non-hydrostatic = 1, no disable-flux: the
library rejects the rest).

With ``checkpoint_every = K`` the trajectory keeps every K-th state.  By
default (``tape=True``) the backward of a segment replays its K steps once
through ``snapx_adjoint_record``, which keeps each step's stage inputs,
primitives, tendencies and column blocks in a tape (K tapes, reused by every
segment: about 26 fields of 5 x cells doubles per step), and then sweeps the
tapes in reverse (``snapx_adjoint_sweep``): one forward recompute per step
instead of two.  The replay follows the recompute's own column solve (cyclic
reduction), so its states agree with the forward's to rounding.

float32 states select the FP32-storage adjoint (``snapx_adjoint_f32_*``):
FP32 tapes and adjoint fields, FP64 arithmetic in registers and FP64 column
solves.  Their forward steps are plain FP32 steps (``RK3Step``) or, in the
segments of ``integrate`` whose tapes are recorded later, fused FP32 steps
(``snapx_fused_rk3_steps``) when the configuration runs on the fused path.

``integrate_batch`` differentiates steps of a batch of members
(``capi.Batch``): fused FP32 batch steps forward, the FP32-storage adjoint of
each member backward.
"""
from __future__ import annotations

import torch

from . import capi


_DTYPES = (torch.float64, torch.float32)


def _code(dtype):
    return capi.F64 if dtype == torch.float64 else capi.F32


class RK3Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, cfg, dt, ws):
        if u.dtype not in _DTYPES or not u.is_cuda:
            raise TypeError("RK3Step: u must be a float64 or float32 GPU tensor")
        u_in = u.detach().contiguous()
        out = u_in.clone()
        w = torch.empty_like(out)
        capi.rk3_step(cfg, dt, out, w)
        ctx.save_for_backward(u_in)
        ctx.cfg, ctx.dt, ctx.ws = cfg, dt, ws
        return out

    @staticmethod
    def backward(ctx, grad):
        (u_in,) = ctx.saved_tensors
        lam = grad.detach().contiguous().clone()
        ws = (ctx.ws if ctx.ws is not None else
              capi.adjoint_workspace(ctx.cfg, u_in.device, u_in.dtype))
        capi.adjoint_rk3_step(ctx.cfg, ctx.dt, u_in, lam, ws)
        return lam, None, None, None


class _Tapes:
    """The tapes of one segment and the scratch of the record and sweep
    calls, allocated on first use and shared by all segments of a
    trajectory (the segments' backwards run one after another)."""

    def __init__(self, cfg, device, dtype=torch.float64):
        self.cfg = cfg
        self.nbytes = capi.adjoint_tape_bytes(cfg, dtype)
        if self.nbytes == 0:
            raise capi.SnapxError("adjoint: " + capi.lib().snapx_last_error().decode())
        self.work = capi.adjoint_work(cfg, device, dtype)
        self.tapes = []
        self.states = None

    def get(self, n, like):
        while len(self.tapes) < n:
            self.tapes.append(torch.empty(self.nbytes, dtype=torch.uint8, device=like.device))
        if self.states is None:
            self.states = (torch.empty_like(like), torch.empty_like(like))
        return self.tapes[:n], self.states


def _record(cfg, dt, u_in, n, st, out=None):
    """Records n steps from u_in into the store's tapes; out (or None) gets
    the last step's output."""
    tapes, (a, b) = st.get(n, u_in)
    x = u_in
    for i in range(n):
        y = (a if i % 2 == 0 else b) if i < n - 1 else out
        capi.adjoint_record(cfg, dt, x, y, tapes[i], st.work)
        x = y
    return tapes


def _fused_ok(cfg):
    return capi.lib().snapx_fused_workspace_bytes(capi.C.byref(cfg)) > 0


class TapedSegment(torch.autograd.Function):
    """n steps as one autograd node: the forward keeps only the segment's
    input (plain-path steps; fused steps for float32 states when the
    configuration runs on the fused path); the backward records the n steps
    once and sweeps their tapes in reverse.  With ``last`` (the segment
    whose backward runs first) the forward records the tapes itself, through
    the recompute's steps, and the backward only sweeps."""

    @staticmethod
    def forward(ctx, u, cfg, dt, n, store, last=False):
        if u.dtype not in _DTYPES or not u.is_cuda:
            raise TypeError("TapedSegment: u must be a float64 or float32 GPU tensor")
        u_in = u.detach().contiguous()
        out = torch.empty_like(u_in)
        if last:
            _record(cfg, dt, u_in, n, store, out)
        elif u_in.dtype == torch.float32 and _fused_ok(cfg):
            out.copy_(u_in)
            capi.fused_rk3_steps(cfg, dt, n, out)
        else:
            out.copy_(u_in)
            w = torch.empty_like(out)
            ws = capi.rk3_workspace(cfg, _code(out.dtype), out.device)
            for _ in range(n):
                capi.rk3_step(cfg, dt, out, w, ws)
        ctx.save_for_backward(u_in)
        ctx.cfg, ctx.dt, ctx.n, ctx.store, ctx.recorded = cfg, dt, n, store, last
        return out

    @staticmethod
    def backward(ctx, grad):
        (u_in,) = ctx.saved_tensors
        cfg, dt, n, st = ctx.cfg, ctx.dt, ctx.n, ctx.store
        tapes = st.get(n, u_in)[0] if ctx.recorded else _record(cfg, dt, u_in, n, st)
        lam = grad.detach().contiguous().clone()
        for i in reversed(range(n)):
            capi.adjoint_sweep(cfg, dt, tapes[i], lam, st.work)
        return lam, None, None, None, None, None


class _MemberStores:
    """One tape store per batch member (created on first use)."""

    def __init__(self, batch, device):
        self.batch, self.device, self.stores = batch, device, {}

    def get(self, m):
        if m not in self.stores:
            self.stores[m] = _Tapes(self.batch.cfgs[m], self.device, torch.float32)
        return self.stores[m]


class BatchSegment(torch.autograd.Function):
    """n steps of a batch (``capi.Batch``: fused FP32 steps of every member in
    one launch per kernel and stage) as one autograd node.  The members are
    independent, so the batch's Jacobian is block-diagonal: the backward
    replays each member's n steps through the FP32-storage adjoint's record
    half and sweeps that member's tapes in reverse on its slice of lambda."""

    @staticmethod
    def forward(ctx, u, batch, dt, n, stores):
        batch._check(u)
        u_in = u.detach().contiguous()
        out = u_in.clone()
        batch.steps(out, n)
        ctx.save_for_backward(u_in)
        ctx.batch, ctx.dt, ctx.n, ctx.stores = batch, dt, n, stores
        return out

    @staticmethod
    def backward(ctx, grad):
        (u_in,) = ctx.saved_tensors
        b, dt, n = ctx.batch, ctx.dt, ctx.n
        lam = grad.detach().contiguous().clone()
        for m in range(b.n):
            cfg, st = b.cfgs[m], ctx.stores.get(m)
            tapes = _record(cfg, dt, b.member(u_in, m), n, st)
            lm = b.member(lam, m)
            for i in reversed(range(n)):
                capi.adjoint_sweep(cfg, dt, tapes[i], lm, st.work)
        return lam, None, None, None, None


def integrate_batch(u, batch, dt, nsteps, checkpoint_every=None):
    """``nsteps`` differentiable steps of a batch: ``u`` is the float32 array
    of the members' states in member order (``batch.total`` elements), ``dt``
    the batch's planned step.  Every member must be a block the adjoint
    takes (Cartesian, limiter off, ...).  With ``checkpoint_every = K`` only
    every K-th batch state is kept; each member keeps K tapes of its own."""
    for m in range(batch.n):
        if capi.adjoint_tape_bytes(batch.cfgs[m], torch.float32) == 0:
            raise capi.SnapxError(f"integrate_batch: member {m}: " +
                                  capi.lib().snapx_last_error().decode())
    every = checkpoint_every if checkpoint_every else nsteps
    stores = _MemberStores(batch, u.device)
    done = 0
    while done < nsteps:
        n = min(every, nsteps - done)
        u = BatchSegment.apply(u, batch, dt, n, stores)
        done += n
    return u


def rk3_step(u, cfg, dt, ws=None):
    """One differentiable SSP-RK3 step (FP64, or FP32 storage for float32
    states); ``ws`` is an optional adjoint workspace
    (``capi.adjoint_workspace`` of u's dtype) reused by the backward."""
    return RK3Step.apply(u, cfg, dt, ws)


def integrate(u, cfg, dt, nsteps, ws=None, checkpoint_every=None, tape=True):
    """``nsteps`` differentiable steps; one adjoint workspace is shared by
    every backward step (they run one after another on the stream).

    With ``checkpoint_every = K`` only every K-th state is kept.  With
    ``tape`` (the default) each segment is one TapedSegment node; without,
    the backward of a segment replays its forward steps from the segment's
    first state (torch.utils.checkpoint) and each step's adjoint recomputes
    its stages again, so memory holds nsteps / K + K states and no tapes."""
    if (checkpoint_every and checkpoint_every < nsteps and tape and torch.is_grad_enabled()
            and u.requires_grad):
        store = _Tapes(cfg, u.device, u.dtype)
        done = 0
        while done < nsteps:
            n = min(checkpoint_every, nsteps - done)
            u = TapedSegment.apply(u, cfg, dt, n, store, done + n == nsteps)
            done += n
        return u
    if ws is None and torch.is_grad_enabled() and u.requires_grad:
        ws = capi.adjoint_workspace(cfg, u.device, u.dtype)
    if not checkpoint_every or checkpoint_every >= nsteps or not torch.is_grad_enabled():
        for _ in range(nsteps):
            u = RK3Step.apply(u, cfg, dt, ws)
        return u
    from torch.utils.checkpoint import checkpoint

    def segment(x, n):
        for _ in range(n):
            x = RK3Step.apply(x, cfg, dt, ws)
        return x

    done = 0
    while done < nsteps:
        n = min(checkpoint_every, nsteps - done)
        u = checkpoint(segment, u, n, use_reentrant=False)
        done += n
    return u
