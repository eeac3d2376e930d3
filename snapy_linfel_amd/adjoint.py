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
"""
from __future__ import annotations

import torch

from . import capi


class RK3Step(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, cfg, dt, ws):
        if u.dtype != torch.float64 or not u.is_cuda:
            raise TypeError("RK3Step: u must be a float64 GPU tensor")
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
        ws = ctx.ws if ctx.ws is not None else capi.adjoint_workspace(ctx.cfg, u_in.device)
        capi.adjoint_rk3_step(ctx.cfg, ctx.dt, u_in, lam, ws)
        return lam, None, None, None


def rk3_step(u, cfg, dt, ws=None):
    """One differentiable FP64 SSP-RK3 step; ``ws`` is an optional adjoint
    workspace (``capi.adjoint_workspace``) reused by the backward."""
    return RK3Step.apply(u, cfg, dt, ws)


def integrate(u, cfg, dt, nsteps, ws=None, checkpoint_every=None):
    """``nsteps`` differentiable steps; one adjoint workspace is shared by
    every backward step (they run one after another on the stream).

    With ``checkpoint_every = K`` only every K-th state is kept: the backward
    of a segment of K steps first replays its forward steps from the
    segment's first state (torch.utils.checkpoint), so memory holds
    nsteps / K + K states instead of nsteps."""
    if ws is None and torch.is_grad_enabled() and u.requires_grad:
        ws = capi.adjoint_workspace(cfg, u.device)
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
