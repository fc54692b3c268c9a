"""torch.autograd wrappers around the forward kernels and their hand-written
adjoints, and a checkpointed multi-step driver.

Each Function runs the library's forward kernel (the masked variant for the
limiters) and, in backward, the matching nicam_*_adjoint entry point; the
coefficients the adjoint holds fixed (ck and d, Mc/Ml/Mu and the metrics)
get no gradient.  Tensors are float64 on the context's device in device
pointer mode.  ``checkpointed`` runs n steps in segments whose forward is
recomputed during the backward pass, so only the segment boundaries are
kept."""
from __future__ import annotations

from typing import Callable

import torch
from torch.utils.checkpoint import checkpoint

from .runtime import Context


class VerticalLimiter(torch.autograd.Function):
    """q_h' = vertical_limiter_thuburn(q_h, q, ck; d) (VLIM:41-243); the inflow
    flags (signs of ck) and d are held fixed."""

    @staticmethod
    def forward(fctx, nctx: Context, q_h, q_h_pl, q, q_pl, d, d_pl, ck, ck_pl):
        out, out_pl = q_h.clone(), q_h_pl.clone()
        M, MP = nctx.vlim_mask_words()
        mask = torch.empty(M, dtype=torch.int32, device=q.device)
        mask_pl = torch.empty(max(MP, 1), dtype=torch.int32, device=q.device)[:MP]
        nctx.vertical_limiter_thuburn_mask(out, out_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl)
        fctx.nctx = nctx
        fctx.save_for_backward(out, out_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl)
        return out, out_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        out, out_pl, q, q_pl, d, d_pl, ck, ck_pl, mask, mask_pl = fctx.saved_tensors
        qhb, qhb_pl = g.contiguous().clone(), g_pl.contiguous().clone()
        qb, qb_pl = torch.empty_like(q), torch.empty_like(q_pl)
        ckb, ckb_pl = torch.empty_like(ck), torch.empty_like(ck_pl)
        fctx.nctx.vertical_limiter_thuburn_adjoint(qhb, qhb_pl, qb, qb_pl, out, out_pl, q, q_pl, d, d_pl, ck, ck_pl,
                                                   mask, mask_pl, ckb, ckb_pl)
        return None, qhb, qhb_pl, qb, qb_pl, None, None, ckb, ckb_pl


class RhowSolve(torch.autograd.Function):
    """rhogw' = vi_rhow_solver(rhogw, rhogw0, preg0, rhog0, Srho, Sw, Spre; dt)
    (VI:93-306) at the context's Mc/Ml/Mu and metrics.  Inputs and outputs
    come as (regions, poles) pairs."""

    @staticmethod
    def forward(fctx, nctx: Context, dt: float, *fields):
        w, w_pl = fields[0].clone(), fields[1].clone()
        nctx.vi_rhow_solver(w, w_pl, *fields[2:], dt)
        fctx.nctx, fctx.dt = nctx, dt
        fctx.shapes = [(f.shape, f.device) for f in fields[2:]]
        return w, w_pl

    @staticmethod
    def backward(fctx, g, g_pl):
        wb, wb_pl = g.contiguous().clone(), g_pl.contiguous().clone()
        rest = [torch.empty(s, dtype=torch.float64, device=dv) for s, dv in fctx.shapes]
        fctx.nctx.vi_rhow_solver_adjoint(wb, wb_pl, *rest, fctx.dt)
        return (None, None, wb, wb_pl, *rest)


def checkpointed(step: Callable, state, n: int, every: int):
    """Applies ``state = step(*state)`` n times; the steps are grouped in
    segments of ``every`` whose intermediate states are recomputed in the
    backward pass instead of kept."""
    state = tuple(state)

    def segment(m):
        def run(*s):
            for _ in range(m):
                s = tuple(step(*s))
            return s
        return run

    done = 0
    while done < n:
        m = min(every, n - done)
        state = checkpoint(segment(m), *state, use_reentrant=False)
        done += m
    return state
