"""Gradients through level steps: a torch.autograd.Function over the C-ABI
adjoint (camr_level_adjoint_forward / _backward).

``LevelSteps.apply(U0, level, dts, every)`` maps the initial state of every
box (a tensor of shape (nbox, NVAR, nz, ny, nx), equal box shapes, or a
single box (NVAR, nz, ny, nx)) to the state after len(dts) steps with that
frozen dt sequence; its backward runs the hand-written adjoint kernels,
recomputing the steps from device checkpoints taken every `every` steps.
The dt sequence is frozen: derivatives do not flow through the CFL
estimate.  Use ``primal_dts`` to get the reference dt sequence of a run.

``EbForce.apply(U0, level, dts, every, w)`` maps the initial state of an
embedded-boundary level to w . F, F the force on the body after the steps
(``Level.eb_force``); its backward seeds the adjoint with the force's own
(``Level.seed_force_adjoint``).
"""
from __future__ import annotations

import numpy as np
import torch

from .hydro import Level


def primal_dts(level: Level, U0, nsteps: int):
    """The dt sequence camr_level_run takes from U0 (the level is left at
    the final state)."""
    _set_state(level, U0)
    return level.run(nsteps)


def _boxes_of(U):
    a = U.detach().cpu().numpy() if isinstance(U, torch.Tensor) else np.asarray(U)
    return [a] if a.ndim == 4 else [a[b] for b in range(a.shape[0])]


def _set_state(level: Level, U):
    for b, ub in enumerate(_boxes_of(U)):
        level.set_box(b, ub)


def _get_state(level: Level, like):
    boxes = [level.get_box(b) for b in range(len(level.prob.boxes))]
    out = boxes[0] if like.dim() == 4 else np.stack(boxes)
    return torch.as_tensor(out, dtype=like.dtype, device=like.device)


class LevelSteps(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, level: Level, dts, every: int = 5):
        _set_state(level, U0)
        level.adjoint_forward(np.asarray(dts, dtype=np.float64), every)
        ctx.level = level
        ctx.like = U0
        return _get_state(level, U0)

    @staticmethod
    def backward(ctx, gU):
        level = ctx.level
        for b, gb in enumerate(_boxes_of(gU)):
            level.set_adjoint(b, gb)
        level.adjoint_backward()
        boxes = [level.get_adjoint(b) for b in range(len(level.prob.boxes))]
        g = boxes[0] if ctx.like.dim() == 4 else np.stack(boxes)
        return torch.as_tensor(g, dtype=ctx.like.dtype, device=ctx.like.device), None, None, None


class EbForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, level: Level, dts, every: int = 5, w=(1.0, 0.0, 0.0)):
        _set_state(level, U0)
        level.adjoint_forward(np.asarray(dts, dtype=np.float64), every)
        f = level.eb_force()
        ctx.level = level
        ctx.like = U0
        ctx.w = tuple(float(x) for x in w)
        return torch.tensor(float(np.dot(ctx.w, f)), dtype=U0.dtype, device=U0.device)

    @staticmethod
    def backward(ctx, gF):
        level = ctx.level
        level.seed_force_adjoint(ctx.w)
        level.adjoint_backward()
        boxes = [level.get_adjoint(b) for b in range(len(level.prob.boxes))]
        g = boxes[0] if ctx.like.dim() == 4 else np.stack(boxes)
        g = torch.as_tensor(g, dtype=ctx.like.dtype, device=ctx.like.device) * gF
        return g, None, None, None, None
