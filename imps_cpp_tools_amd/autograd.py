"""Differentiable transfer-matrix, H_AC and H_C applications (torch.autograd.Function wrappers).

The forward passes are the library's grouped GEMM plans (TMPlan, HACPlan, HCPlan; FP64, FP32 or
BF16X3).  The backward passes are the library's hand-written adjoints through the C ABI
(imps_tm_vjp, imps_hac_vjp, imps_hc_vjp, FP64): torch only records the graph and hands the
cotangents over.

  left TM   y = sum_s B_s^T x A_s (TMBlockFunctions.hpp:352-390; B = A for the regular TM)
  right TM  y = sum_s A_s x B_s^T (:494-533)
  H_AC      y[i] = EHL x[i] + x[i] EHR + sum_j HL[i][j] x[j] + x[j] HR[i][j] (VUMPS_helpers.hpp:38-53)
  H_C       y = EHL C + C EHR + sum_ij HL[i][j] C AR[j] AR[i]^T, diagonal blocks (:91-111, reduced form)

Each gradient is restricted onto the sector structure of its parameter (the blocks the reference
stores), so a gradient step never creates sectors.  The forward's intermediates are recomputed in
the backward (nothing but the inputs is kept between the passes).
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import torch

from . import FP64, BlockLayout, Context, HACPlan, HCPlan, TMPlan


def _c(t):
    return t.detach().contiguous()


def _empty(n, like):
    return torch.empty(n, dtype=torch.float64, device=like.device)


class TransferMatrix:
    """Differentiable (mixed) transfer-matrix apply on fixed sector structures."""

    def __init__(self, ctx: Context, direction: int, A: BlockLayout, B: Optional[BlockLayout], x: BlockLayout, y: BlockLayout,
                 prec: int = FP64):
        self.ctx, self.direction = ctx, direction
        self.Al, self.Bl, self.xl, self.yl = A, (B if B is not None else A), x, y
        self.mixed = B is not None
        self.plan = TMPlan(ctx, direction, A, B, x, y, prec)

    def __call__(self, x, A, B=None):
        if self.mixed:
            return _TMFn.apply(self, x, A, B)
        return _TMRegFn.apply(self, x, A)

    def forward_raw(self, x, A, B):
        y = torch.empty(self.yl.nelem, dtype=torch.float64, device=x.device)
        self.plan.set_mps(_c(A), _c(B) if self.mixed else None)
        self.plan.apply(_c(x), y)
        return y

    def vjp(self, x, A, B, ybar):
        xbar, Abar = _empty(self.xl.nelem, x), _empty(self.Al.nelem, x)
        Bbar = _empty(self.Bl.nelem, x) if self.mixed else None
        self.plan.vjp(_c(A), _c(B) if self.mixed else None, _c(x), _c(ybar), xbar, Abar, Bbar)
        return xbar, Abar, Bbar


class _TMFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tm: TransferMatrix, x, A, B):
        ctx.tm = tm
        ctx.save_for_backward(x, A, B)
        return tm.forward_raw(x, A, B)

    @staticmethod
    def backward(ctx, ybar):
        x, A, B = ctx.saved_tensors
        xbar, Abar, Bbar = ctx.tm.vjp(x, A, B, ybar)
        return None, xbar, Abar, Bbar


class _TMRegFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, tm: TransferMatrix, x, A):
        ctx.tm = tm
        ctx.save_for_backward(x, A)
        return tm.forward_raw(x, A, None)

    @staticmethod
    def backward(ctx, ybar):
        x, A = ctx.saved_tensors
        xbar, Abar, _ = ctx.tm.vjp(x, A, None, ybar)
        return None, xbar, Abar


class TransferChain:
    """y = T_{L-1}(... T_1(T_0(x))) for a sequence of transfer-matrix applies (TransferMatrix objects,
    one per step, each with its own MPS tensors): a unit cell's chain (TMBlockFunctions.hpp:542-788)
    or several cells of it, as an environment is built.  The forward keeps only the input and every
    `every`-th environment (default ceil(sqrt(L))); the backward recomputes each segment's
    environments from its checkpoint and runs the steps' adjoints in reverse, so the memory is
    O(L / every + every) environments instead of L, for one extra forward."""

    def __init__(self, tms: Sequence[TransferMatrix], every: Optional[int] = None):
        self.tms = list(tms)
        if not self.tms:
            raise ValueError("TransferChain: no steps")
        for a, b in zip(self.tms, self.tms[1:]):
            if not a.yl == b.xl:
                raise ValueError("TransferChain: a step's output structure is not the next step's input")
        if any(t.mixed != self.tms[0].mixed for t in self.tms):
            raise ValueError("TransferChain: all steps regular or all mixed")
        self.every = int(every) if every else max(1, math.ceil(math.sqrt(len(self.tms))))
        self.last_checkpoints = 0  # environments the last forward kept (input included)

    def __call__(self, x, As: Sequence, Bs: Optional[Sequence] = None):
        L = len(self.tms)
        if len(As) != L or (Bs is not None and len(Bs) != L) or (Bs is not None) != self.tms[0].mixed:
            raise ValueError("TransferChain: one tensor (A, and B when mixed) per step")
        return _ChainFn.apply(self, x, *As, *(Bs if Bs is not None else ()))


class _ChainFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, chain: TransferChain, x, *ts):
        L = len(chain.tms)
        As, Bs = ts[:L], (ts[L:] if chain.tms[0].mixed else (None,) * L)
        keep, idx = [x], [0]
        y = x
        for i, tm in enumerate(chain.tms):
            y = tm.forward_raw(y, As[i], Bs[i])
            if (i + 1) % chain.every == 0 and i + 1 < L:
                keep.append(y)
                idx.append(i + 1)
        ctx.chain, ctx.idx, ctx.nkeep = chain, idx, len(keep)
        chain.last_checkpoints = len(keep)
        ctx.save_for_backward(*keep, *ts)
        return y

    @staticmethod
    def backward(ctx, ybar):
        chain = ctx.chain
        L = len(chain.tms)
        saved = ctx.saved_tensors
        keep, ts = saved[:ctx.nkeep], saved[ctx.nkeep:]
        As, Bs = ts[:L], (ts[L:] if chain.tms[0].mixed else (None,) * L)
        gA, gB = [None] * L, [None] * L
        g = ybar.contiguous()
        for c in range(len(ctx.idx) - 1, -1, -1):
            s = ctx.idx[c]
            e = ctx.idx[c + 1] if c + 1 < len(ctx.idx) else L
            xs = [keep[c]]
            for i in range(s, e - 1):  # the segment's environments, recomputed
                xs.append(chain.tms[i].forward_raw(xs[-1], As[i], Bs[i]))
            for i in range(e - 1, s - 1, -1):
                g, gA[i], gB[i] = chain.tms[i].vjp(xs[i - s], As[i], Bs[i], g)
        return (None, g, *gA, *(gB if chain.tms[0].mixed else ()))


class TwoSiteHAC:
    """Differentiable H_AC apply: y = H_AC(x; EHL, EHR, HL, HR)."""

    def __init__(self, ctx: Context, AC: BlockLayout, EHL: BlockLayout, EHR: BlockLayout, HL: BlockLayout, HR: BlockLayout,
                 prec: int = FP64):
        self.ctx = ctx
        self.lays = (AC, EHL, EHR, HL, HR)
        self.plan = HACPlan(ctx, AC, EHL, EHR, HL, HR, prec)

    def __call__(self, x, EHL, EHR, HL, HR):
        return _OpFn.apply(self, x, EHL, EHR, HL, HR)

    def forward_raw(self, x, *ops):
        y = torch.empty(self.lays[0].nelem, dtype=torch.float64, device=x.device)
        self.plan.set_operator(*[_c(o) for o in ops])
        self.plan.apply(_c(x), y)
        return y

    def vjp(self, x, ops, ybar):
        xbar = _empty(self.lays[0].nelem, x)
        bars = [_empty(l.nelem, x) for l in self.lays[1:]]
        self.plan.vjp([_c(o) for o in ops], _c(x), _c(ybar), xbar, bars)
        return (xbar, *bars)


class BondHC(TwoSiteHAC):
    """Differentiable H_C apply (reduced form): y = H_C(C; EHL, EHR, HL, AR)."""

    def __init__(self, ctx: Context, C: BlockLayout, EHL: BlockLayout, EHR: BlockLayout, HL: BlockLayout, AR: BlockLayout,
                 prec: int = FP64):
        self.ctx = ctx
        self.lays = (C, EHL, EHR, HL, AR)
        self.plan = HCPlan(ctx, C, EHL, EHR, HL, AR, prec)


class _OpFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h, x, *ops):
        ctx.h = h
        ctx.save_for_backward(x, *ops)
        return h.forward_raw(x, *ops)

    @staticmethod
    def backward(ctx, ybar):
        x, *ops = ctx.saved_tensors
        return (None,) + tuple(ctx.h.vjp(x, ops, ybar))
