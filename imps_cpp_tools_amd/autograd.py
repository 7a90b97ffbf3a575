"""Differentiable transfer-matrix and H_AC applications (torch.autograd.Function wrappers).

The forward passes are the library's grouped GEMM plans (TMPlan, HACPlan; FP64 or FP32).  The
backward passes are hand-written adjoints built from the library's FP64 block products:

  left TM   y = sum_s B_s^T x A_s (TMBlockFunctions.hpp:352-390; B = A for the regular TM)
            xbar   = sum_s B_s ybar A_s^T          (the right mixed TM with A and B exchanged)
            Abar_s = x^T B_s ybar,   Bbar_s = x A_s ybar^T
  right TM  y = sum_s A_s x B_s^T (:494-533)
            xbar   = sum_s A_s^T ybar B_s          (the left mixed TM with A and B exchanged)
            Abar_s = ybar B_s x^T,   Bbar_s = ybar^T A_s x
  H_AC      y[i] = EHL x[i] + x[i] EHR + sum_j HL[i][j] x[j] + x[j] HR[i][j] (VUMPS_helpers.hpp:38-53)
            xbar[j] = EHL^T ybar[j] + ybar[j] EHR^T + sum_i HL[i][j]^T ybar[i] + ybar[i] HR[i][j]^T
            EHLbar = sum_i ybar[i] x[i]^T,  EHRbar = sum_i x[i]^T ybar[i]
            HLbar[i][j] = ybar[i] x[j]^T,   HRbar[i][j] = x[j]^T ybar[i]
Each gradient is restricted onto the sector structure of its parameter (the blocks the reference
stores), so a gradient step never creates sectors.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import (FP64, LEFT, RIGHT, MAT_MPS, MATT_MPS, MPS_MAT, MPS_MATT, TM_LEFT_UNITY, TM_RIGHT_UNITY, BlockLayout, Context,
               HACPlan, TMPlan)
from .excitations import BlockOps, DObj, _index


class TransferMatrix:
    """Differentiable (mixed) transfer-matrix apply on fixed sector structures."""

    def __init__(self, ctx: Context, direction: int, A: BlockLayout, B: Optional[BlockLayout], x: BlockLayout, y: BlockLayout,
                 prec: int = FP64):
        self.ctx, self.direction = ctx, direction
        self.Al, self.Bl, self.xl, self.yl = A, (B if B is not None else A), x, y
        self.mixed = B is not None
        self.plan = TMPlan(ctx, direction, A, B, x, y, prec)
        # adjoint in x: the opposite-direction mixed TM with A and B exchanged, from y to x
        self.plan_t = TMPlan(ctx, RIGHT if direction == LEFT else LEFT, self.Bl, self.Al, y, x, FP64)
        self.ops = BlockOps(ctx, A.nsym)

    def __call__(self, x, A, B=None):
        return _TMFn.apply(self, x, A, A if B is None else B)

    def forward_raw(self, x, A, B):
        y = torch.empty(self.yl.nelem, dtype=torch.float64, device=x.device)
        self.plan.set_mps(A, B if self.mixed else None)
        self.plan.apply(x.contiguous(), y)
        return y

    def vjp(self, x, A, B, ybar):
        ops = self.ops
        ybar = ybar.contiguous()
        xbar = torch.empty(self.xl.nelem, dtype=torch.float64, device=x.device)
        self.plan_t.set_mps(B, A)
        self.plan_t.apply(ybar, xbar)
        X, Y = DObj(self.xl, x.contiguous()), DObj(self.yl, ybar)
        Ad, Bd = DObj(self.Al, A.contiguous()), DObj(self.Bl, B.contiguous())
        if self.direction == LEFT:
            Abar = ops.mps_mat(ops.prod(MATT_MPS, X, Bd), Y)          # x^T B_s ybar
            Bbar = ops.prod(MPS_MATT, ops.mat_mps(X, Ad), Y)          # x A_s ybar^T
        else:
            Abar = ops.prod(MPS_MATT, ops.mat_mps(Y, Bd), X)          # ybar B_s x^T
            Bbar = ops.mps_mat(ops.prod(MATT_MPS, Y, Ad), X)          # ybar^T A_s x
        return xbar, ops.onto(Abar, self.Al).data, ops.onto(Bbar, self.Bl).data


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


class TwoSiteHAC:
    """Differentiable H_AC apply: y = H_AC(x; EHL, EHR, HL, HR)."""

    def __init__(self, ctx: Context, AC: BlockLayout, EHL: BlockLayout, EHR: BlockLayout, HL: BlockLayout, HR: BlockLayout,
                 prec: int = FP64):
        self.ctx = ctx
        self.lays = (AC, EHL, EHR, HL, HR)
        self.plan = HACPlan(ctx, AC, EHL, EHR, HL, HR, prec)
        self.ops = BlockOps(ctx, AC.nsym)
        self.d = len(AC.nblk)

    def __call__(self, x, EHL, EHR, HL, HR):
        return _HACFn.apply(self, x, EHL, EHR, HL, HR)

    def forward_raw(self, x, EHL, EHR, HL, HR):
        y = torch.empty(self.lays[0].nelem, dtype=torch.float64, device=x.device)
        self.plan.set_operator(EHL.contiguous(), EHR.contiguous(), HL.contiguous(), HR.contiguous())
        self.plan.apply(x.contiguous(), y)
        return y

    def vjp(self, x, EHL, EHR, HL, HR, ybar):
        ops, d = self.ops, self.d
        ACl, EHLl, EHRl, HLl, HRl = self.lays
        X, Y = DObj(ACl, x.contiguous()), DObj(ACl, ybar.contiguous())
        E_L, E_R = DObj(EHLl, EHL.contiguous()), DObj(EHRl, EHR.contiguous())
        GL, GR = DObj(HLl, HL.contiguous()), DObj(HRl, HR.contiguous())
        real = lambda o: o.data[0] if o.cplx else o.data  # noqa: E731
        # xbar
        xb = ops.add(ops.prod(MATT_MPS, E_L, Y), ops.prod(MPS_MATT, Y, E_R))
        Ye = [ops.entry(Y, i) for i in range(d)]
        Xe = [ops.entry(X, j) for j in range(d)]
        parts = []
        for j in range(d):
            acc = None
            for i in range(d):
                hl, hr = ops.entry(GL, i * d + j), ops.entry(GR, i * d + j)
                for term in (ops.prod(MATT_MPS, hl, Ye[i]), ops.prod(MPS_MATT, Ye[i], hr)):
                    acc = term if acc is None else ops.add(acc, term)
            parts.append(acc)
        xb = ops.add(xb, ops.concat([ops.scale(p, 1.0) for p in parts]))
        xbar = real(ops.onto(xb, ACl))
        # operator gradients
        ehl = real(ops.onto(ops.prod(TM_RIGHT_UNITY, Y, X), EHLl))   # sum_i ybar[i] x[i]^T
        ehr = real(ops.onto(ops.prod(TM_LEFT_UNITY, Y, X), EHRl))    # sum_i x[i]^T ybar[i]
        hl_parts, hr_parts = [], []
        for i in range(d):
            for j in range(d):
                hl_parts.append(ops.onto(ops.prod(MPS_MATT, Ye[i], Xe[j]), ops.entry(GL, i * d + j).lay))   # ybar[i] x[j]^T
                hr_parts.append(ops.onto(ops.prod(MATT_MPS, Xe[j], Ye[i]), ops.entry(GR, i * d + j).lay))   # x[j]^T ybar[i]
        hl = torch.cat([p.data for p in hl_parts])
        hr = torch.cat([p.data for p in hr_parts])
        return xbar.contiguous(), ehl.contiguous(), ehr.contiguous(), hl, hr


class _HACFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h: TwoSiteHAC, x, EHL, EHR, HL, HR):
        ctx.h = h
        ctx.save_for_backward(x, EHL, EHR, HL, HR)
        return h.forward_raw(x, EHL, EHR, HL, HR)

    @staticmethod
    def backward(ctx, ybar):
        x, EHL, EHR, HL, HR = ctx.saved_tensors
        return (None,) + tuple(ctx.h.vjp(x, EHL, EHR, HL, HR, ybar))
