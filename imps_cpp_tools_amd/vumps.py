"""VUMPS ground-state iteration for an N-site unit cell at fixed bond structure.

A thin driver over the library's GPU operations: every contraction, eigensolve, environment solve,
QR and fixed point runs in the HIP library through the C ABI; this module only sequences them and
keeps the device buffers.  The update order follows the outer loop of VUMPS_main.cpp:165-354
(sequential sweep over the unit cell):

  site n (left bond n, right bond n+1; C[n] lives on bond n+1):
    F(n)  = gradient norm at AC = C[n-1] AR[n]          (GradNormRight, VUMPS_helpers.hpp:147-153)
    AC    = lowest eigenvector of H_AC                   (eigs_rs "SA", :186)
    CL,CR = lowest eigenvectors of H_C on bonds n, n+1   (:199-202, sign fixed by the first entry)
    AL[n] = polar(AC CR^T), AR[n] = polar(CL^T AC)       (SVD scheme, :213-220, until the gauge
                                                          error is below svdthresh; polar factors by
                                                          the Newton-Schulz iteration)
    AL[n] = Q(AC,l) Q(CR,l)^T,  AR[n] = Q(CL,r)^T Q(AC,r) (QR polar scheme, :225-232, afterwards)
    environments: LE[n] by the left geometric sum with R = fixed point of the cell TM (trueLR),
    LE[n+1], LE[n+2] by TM steps, RE[n+2] by the right geometric sum, RE[n+1] by a TM step
    (:247-305).
Environments: LE[b] holds the two-site terms left of bond b that do not touch site b
(InvertE_proj on GetHL, VUMPS_helpers.hpp:3-18); RE[b] the terms right of bond b - 1 on sites >= b
(GetHR, :20-36).  H_AC at site n uses LE[n], RE[n+1] and the reduced operators of AL[n-1],
AR[n+1]; H_C on bond b uses LE[b], RE[b], AL[b-1], AR[b].

Energy per site = tr(HL_b R_b) / N with HL_b the one-cell source of the left geometric sum.
Initial state: the reference's random mixed-canonical MPS (randMPS_LR), see VUMPS._random_state.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import (BF16X3, FP32, FP64, LEFT, RIGHT, MAT_MPS, MATT_MPS, MPS_MAT, MPS_MATT, MPS_MPS, MPS_REDOP, REDOP_MPS, TM_LEFT_UNITY,
               TM_RIGHT_UNITY, BlockLayout, Context, EnvPlan, HACPlan, HCPlan, ProdPlan, QRPlan, RedOpPlan, TMPlan, eigsh,
               make_dims, model_d, model_ham, model_nsym, model_periods, mps_structure)
from .streams import side_streams


def lower_sqrt_factor(r: np.ndarray, lay: BlockLayout) -> np.ndarray:
    """Per sector of the symmetric positive semi-definite block-diagonal R (column-major blocks),
    the lower-triangular C with positive diagonal and C C^T = R: C = RQ factor of U sqrt(lam), with
    R = U lam U^T (svd(U, DR, V, R); qr(C, U << sqrt(DR), r) in randMPS_LR)."""
    out = np.zeros(lay.nelem)
    for o, s in zip(lay.offsets, lay.rows):
        m = r[o:o + s * s].reshape(s, s, order="F")
        U, lam, _ = np.linalg.svd(0.5 * (m + m.T))
        _, t = np.linalg.qr((U * np.sqrt(lam)).T)
        t = np.where(np.diag(t) < 0, -1.0, 1.0)[:, None] * t
        out[o:o + s * s] = t.T.reshape(-1, order="F")
    return out


@dataclass
class VumpsStats:
    iteration: int
    energy: float
    F: List[float]
    gradnorm: List[float]
    seconds: float
    eig_iters: int = 0
    env_iters: int = 0


class VUMPS:
    def __init__(self, ctx: Context, model: int, N: int, QN, keys0, sizes0, params, prec: int = FP64, seed: int = 0,
                 eig_tol_max: float = 1e-6, eig_tol_min: float = 5e-16, env_tol: float = 5e-15, env_fac: float = 2.0,
                 krylov: int = 20, eig_maxit: int = 100000, env_krylov: int = 100, env_maxit: int = 20000, polar_start: bool = True,
                 plans: Optional[Dict] = None, mmax: Optional[int] = None, dm_max: int = 20, expand_pause: int = 5,
                 expand_thresh: float = 1e-5, lamthresh: float = 1e-12, concurrent: bool = True):
        """`plans`: the plan cache of another VUMPS over the same model and bond structure, reused
        (plans hold structure only; every use sets its operator data).

        Bond growth (VUMPS_main.cpp:273-306): with `mmax` set (one value, or one per bond), after
        each outer iteration every bond whose total dimension is below mmax grows by ExpandFromH
        with at most min(dm_max, mmax - total) states, when at least `expand_pause` iterations
        passed since its last growth, the gauge error at the start of the iteration is below
        `expand_thresh` and some sector's smallest Schmidt value exceeds `lamthresh` (reference
        defaults: dmmax 20, pause 5, exthresh 1e-5, lamthresh 1e-12).

        `concurrent`: the three eigensolves of a site step (H_AC, H_C on both bonds) read the same
        environments and are independent; they run at once, each on its own stream from its own
        host thread (each solve is a chain of latency-bound Krylov steps, so they overlap).  The
        results are those of the sequential order."""
        import torch
        self.torch = torch
        self.ctx, self.prec, self.N = ctx, prec, N
        self.model, self.QN = model, list(QN)
        self.mmax, self.dm_max, self.expand_pause, self.expand_thresh, self.lamthresh = mmax, dm_max, expand_pause, expand_thresh, lamthresh
        self.expand_ct = [0] * N
        self.concurrent = concurrent
        self._side = None  # (contexts of the two H_C solves, thread pool), made on first use
        self.d, self.nsym = model_d(model), model_nsym(model)
        self.periods = model_periods(model)
        self.bonds = make_dims(model, N, QN, keys0, sizes0)
        self.blay = [BlockLayout.diag(self.nsym, k, s, self.periods) for (k, s) in self.bonds]
        self.slay = [mps_structure(model, N, QN, self.bonds[n], self.bonds[(n + 1) % N]) for n in range(N)]
        self.H = model_ham(model, params)
        self.eig_tol_max, self.eig_tol_min, self.env_tol, self.env_fac, self.krylov = eig_tol_max, eig_tol_min, env_tol, env_fac, krylov
        self.eig_maxit, self.env_krylov, self.env_maxit = eig_maxit, env_krylov, env_maxit
        self.dosvd, self.svdthresh = polar_start, 5e-7  # polar (SVD) scheme until the gauge error is below svdthresh
        self._plans: Dict = {} if plans is None else plans
        self._units: Dict = {}
        self._random_state(np.random.default_rng(seed))
        self.prec_n = np.ones(N)
        self.LE: Dict[int, object] = {}
        self.RE: Dict[int, object] = {}
        self.eig_iters = 0
        self.env_iters = 0
        self._init_environments()

    def _random_state(self, rng):
        """randMPS_LR (MPSBlockUtilities.hpp:1095-1133): left-orthonormal AL[n] from the QR of Gaussian
        tensors, R = the trace-1 dominant right fixed point of the cell TM on bond 0 (TMDominantEig),
        C[N-1] = the lower-triangular factor with C C^T = R, then AR[n] and C[n-1] from the right QR
        of AL[n] C[n] for n = N-1, ..., 0, so that AL[n] C[n] = C[n-1] AR[n] on every site but the
        last QR's leftover factor (discarded, as in the reference)."""
        N = self.N
        self.AL = [self._qr(self._dev(rng.standard_normal(self.slay[n].nelem)), self.slay[n], LEFT)[0] for n in range(N)]
        pr = EnvPlan(self.ctx, RIGHT, self.slay, self.blay[0], FP64)
        pr.set_mps(self.AL)
        R = self._identity(0, 1.0 / float(self.bonds[0][1].sum()))
        pr.fixed_point(R, tol=1e-14, maxit=self.env_maxit)
        C = self._dev(lower_sqrt_factor(R.cpu().numpy(), self.blay[0]))
        self.C, self.AR = [None] * N, [None] * N
        for n in range(N - 1, -1, -1):
            self.C[n] = C
            X, _ = self._prod(MPS_MAT, self.slay[n], self.AL[n], self.blay[self._b(n + 1)], C, ("init", n))
            self.AR[n], C = self._qr(X, self.slay[n], RIGHT)

    # ------------------------------------------------------------------ helpers
    def _dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()

    def _empty(self, n):
        return self.torch.empty(n, dtype=self.torch.float64, device="cuda")

    def _b(self, b):
        return b % self.N

    def _identity(self, bond, scale=1.0):
        lay = self.blay[bond]
        out = np.zeros(lay.nelem)
        for o, s in zip(lay.offsets, lay.rows):
            out[o:o + s * s].reshape(s, s)[np.diag_indices(s)] = scale
        return self._dev(out)

    def _plan(self, key, make):
        p = self._plans.get(key)
        if p is None:
            p = make()
            self._plans[key] = p
        return p

    def _qr(self, A, lay, direction):
        p = self._plan(("qr", id(lay), direction), lambda: QRPlan(self.ctx, lay, direction))
        Q = self._empty(lay.nelem)
        R = self._empty(p.out_layout.nelem)
        p.apply(A, Q, R)
        return Q, R

    def _prod(self, kind, l1, x1, l2, x2, key):
        p = self._plan(("prod", kind, key), lambda: ProdPlan(self.ctx, kind, l1, l2))
        y = self._empty(p.out_layout.nelem)
        p.apply(x1, x2, y)
        return y, p.out_layout

    def _redop(self, A, n, direction):
        p = self._plan(("redop", n, direction), lambda: RedOpPlan(self.ctx, self.H, self.d, self.slay[n], direction))
        y = self._empty(p.out_layout.nelem)
        p.apply(A, y)
        return y, p.out_layout

    def _tm(self, direction, A, n, x, bin_, bout):
        p = self._plan(("tm", direction, n), lambda: TMPlan(self.ctx, direction, self.slay[n], None, self.blay[bin_], self.blay[bout]))
        p.set_mps(A)
        y = self._empty(self.blay[bout].nelem)
        p.apply(x, y)
        return y

    def _optm_left(self, a):
        """ApplyOpTMLeftDiag(H, AL[a] AL[a+1]) on bond a+2 (TMBlockFunctions.hpp:800-806)."""
        a = self._b(a)
        b = self._b(a + 1)
        HL, HLl = self._redop(self.AL[self._b(a)], self._b(a), LEFT)
        W, Wl = self._prod(REDOP_MPS, HLl, HL, self.slay[b], self.AL[b], ("optmL", a))
        y, _ = self._prod(TM_LEFT_UNITY, Wl, W, self.slay[b], self.AL[b], ("optmL2", a))
        return y

    def _optm_right(self, a):
        """ApplyOpTMRightDiag(H, AR[a] AR[a+1]) on bond a (TMBlockFunctions.hpp:842-848)."""
        a = self._b(a)
        HR, HRl = self._redop(self.AR[self._b(a + 1)], self._b(a + 1), RIGHT)
        W, Wl = self._prod(MPS_REDOP, HRl, HR, self.slay[a], self.AR[a], ("optmR", a))
        y, _ = self._prod(TM_RIGHT_UNITY, Wl, W, self.slay[a], self.AR[a], ("optmR2", a))
        return y

    def _get_hl(self, b):
        """GetHL for the cell starting at site b (VUMPS_helpers.hpp:3-18), on bond b."""
        N = self.N
        h = self._optm_left(b - 1)  # term (b-1, b) on bond b+1
        for k in range(1, N):
            s = self._b(b + k)
            h = self._tm(LEFT, self.AL[s], s, h, s, self._b(s + 1)) + self._optm_left(b + k - 1)
        return h

    def _get_hr(self, b):
        """GetHR for the cell starting at site b (VUMPS_helpers.hpp:20-36), on bond b."""
        N = self.N
        h = self._optm_right(b + N - 1)  # term (b+N-1, b+N) on bond b+N-1
        for k in range(N - 2, -1, -1):
            s = self._b(b + k)
            h = self._tm(RIGHT, self.AR[s], s, h, self._b(s + 1), s) + self._optm_right(b + k)
        return h

    def _left_env_job(self, b, tol, side: int = -1):
        """LE[b] = InvertE_proj(AL cell from b, GetHL, 1, R, l) with R the cell fixed point (trueLR).
        The operands are prepared here on the main stream; the returned job runs the fixed point
        and the geometric sum (on the main stream, or on side stream `side`) and returns (LE[b],
        Krylov steps).  The environment plans are FP64 at every working precision: the geometric
        sum's operator 1 - T has the TM gap as its smallest scale, so a low-precision T gains
        nothing on a single right-hand side (each step is latency-bound) and its rounding,
        amplified by 1 / gap, forces extra refinement rounds."""
        b = self._b(b)
        sites = [self._b(b + k) for k in range(self.N)]
        ctx = self.ctx if side < 0 else self._side_contexts()[0][side]
        p = self._plan(("env", LEFT, b, id(ctx)), lambda: EnvPlan(ctx, LEFT, [self.slay[s] for s in sites], self.blay[b], FP64))
        p.set_mps([self.AL[s] for s in sites])
        Cb = self.C[self._b(b - 1)]
        R, _ = self._prod(MPS_MATT, self.blay[b], Cb, self.blay[b], Cb, ("CCt", b))  # C C^T
        # fixed point of T_R over the cell: the right-direction chain of the same AL tensors
        pr = self._plan(("envR_AL", b, id(ctx)), lambda: EnvPlan(ctx, RIGHT, [self.slay[s] for s in sites], self.blay[b], FP64))
        pr.set_mps([self.AL[s] for s in sites])
        HAL = self._get_hl(b)
        one = self._identity(b)
        y = self._empty(self.blay[b].nelem)
        y0 = self.LE.get(b)

        def job():
            pr.fixed_point(R, tol=1e-14, maxit=self.env_maxit, check=False)  # R becomes the fixed point
            it, rr = p.solve(HAL, one, R, y, y0=y0, tol=tol, maxit=self.env_maxit, krylov=self.env_krylov, check=False)
            return y, it
        return job, (HAL, R, b)

    def _right_env_job(self, b, tol, side: int = -1):
        """RE[b] = InvertE_proj(AR cell from b, GetHR, L, 1, r), L = the cell fixed point; as
        _left_env_job."""
        b = self._b(b)
        sites = [self._b(b + k) for k in range(self.N)]
        ctx = self.ctx if side < 0 else self._side_contexts()[0][side]
        p = self._plan(("env", RIGHT, b, id(ctx)), lambda: EnvPlan(ctx, RIGHT, [self.slay[s] for s in sites], self.blay[b], FP64))
        p.set_mps([self.AR[s] for s in sites])
        Cb = self.C[self._b(b - 1)]
        L, _ = self._prod(MATT_MPS, self.blay[b], Cb, self.blay[b], Cb, ("CtC", b))  # C^T C
        pl = self._plan(("envL_AR", b, id(ctx)), lambda: EnvPlan(ctx, LEFT, [self.slay[s] for s in sites], self.blay[b], FP64))
        pl.set_mps([self.AR[s] for s in sites])
        HAR = self._get_hr(b)
        one = self._identity(b)
        y = self._empty(self.blay[b].nelem)
        y0 = self.RE.get(b)

        def job():
            pl.fixed_point(L, tol=1e-14, maxit=self.env_maxit, check=False)
            it, rr = p.solve(HAR, L, one, y, y0=y0, tol=tol, maxit=self.env_maxit, krylov=self.env_krylov, check=False)
            return y, it
        return job

    def _left_env(self, b, tol):
        job, (HAL, R, bb) = self._left_env_job(b, tol)
        y, it = job()
        self.env_iters += it
        e_cell = float(self.torch.dot(HAL, self._transpose_blocks(R, bb)).item())
        return y, e_cell

    def _right_env(self, b, tol):
        y, it = self._right_env_job(b, tol)()
        self.env_iters += it
        return y

    def _envs(self, n, tol):
        """(LE[n], e_cell, RE[n+2]): the left and the right geometric sums of a site step, which are
        independent, at once on two streams when `concurrent`."""
        if not self.concurrent:
            LE, e = self._left_env(n, tol)
            return LE, e, self._right_env(n + 2, tol)
        lj, (HAL, R, bb) = self._left_env_job(n, tol)
        rj = self._right_env_job(n + 2, tol, side=0)
        self.torch.cuda.synchronize()  # operands made on the main stream
        pool = self._side_contexts()[2]
        fr = pool.submit(rj)
        LE, itl = lj()
        RE, itr = fr.result()
        self.torch.cuda.synchronize()
        self.env_iters += itl + itr
        e_cell = float(self.torch.dot(HAL, self._transpose_blocks(R, bb)).item())
        return LE, e_cell, RE

    def _transpose_blocks(self, x, b):
        """Blockwise transpose on the block-diagonal layout of bond b (one gather)."""
        k = ("transpose", b)
        perm = self._units.get(k)
        if perm is None:
            lay = self.blay[b]
            idx = np.empty(lay.nelem, np.int64)
            for o, s in zip(lay.offsets, lay.rows):
                idx[o:o + s * s] = o + np.arange(s * s).reshape(s, s).T.reshape(-1)
            perm = self.torch.from_numpy(idx).cuda()
            self._units[k] = perm
        return x[perm]

    def _init_environments(self):
        N = self.N
        for b in range(N):
            self.LE[b], _ = self._left_env(b, self.env_tol)
            self.RE[b] = self._right_env(b, self.env_tol)

    # ------------------------------------------------------------------ operators
    def _hac_plans(self, n):
        N = self.N
        left, right = self._b(n - 1), self._b(n + 1)
        HL, HLl = self._redop(self.AL[left], left, LEFT)
        HR, HRl = self._redop(self.AR[right], right, RIGHT)
        mk = lambda prec: HACPlan(self.ctx, self.slay[n], self.blay[n], self.blay[self._b(n + 1)], HLl, HRl, prec)  # noqa: E731
        p = self._plan(("hac", n, self.prec), lambda: mk(self.prec))
        ops = (self.LE[n], self.RE[self._b(n + 1)], HL, HR)
        p.set_operator(*ops)
        p64 = None
        if self.prec != FP64:
            p64 = self._plan(("hac", n, FP64), lambda: mk(FP64))
            p64.set_operator(*ops)
        return p, p64

    def _hc_plans(self, b, side: int = -1):
        """H_C plans of bond b on the main context, or (side 0 / 1) on the stream of the left /
        right H_C solve of a concurrent site step."""
        b = self._b(b)
        a = self._b(b - 1)
        HL, HLl = self._redop(self.AL[a], a, LEFT)
        ctx = self.ctx if side < 0 else self._side_contexts()[0][side]
        mk = lambda prec: HCPlan(ctx, self.blay[b], self.blay[b], self.blay[b], HLl, self.slay[b], prec)  # noqa: E731
        p = self._plan(("hc", b, self.prec, side, id(ctx)), lambda: mk(self.prec))
        ops = (self.LE[b], self.RE[b], HL, self.AR[b])
        if side >= 0:
            self.torch.cuda.synchronize()  # the operator data come from the main stream
        p.set_operator(*ops)
        p64 = None
        if self.prec != FP64:
            p64 = self._plan(("hc", b, FP64, side, id(ctx)), lambda: mk(FP64))
            p64.set_operator(*ops)
        if side >= 0:
            # the copies run on the side stream; HL's memory returns to torch's allocator, which
            # only orders it with the main stream
            self.torch.cuda.synchronize()
        return p, p64

    def _side_contexts(self):
        if self._side is None:
            from concurrent.futures import ThreadPoolExecutor
            streams = side_streams(self.torch, self.ctx.device)
            self._side = ([Context(self.ctx.device, st) for st in streams], streams, ThreadPoolExecutor(max_workers=3))
        return self._side

    def close(self):
        """Release the side streams' thread pool (also on garbage collection)."""
        if getattr(self, "_side", None) is not None:
            self._side[2].shutdown(wait=True)
            self._side = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _eigensolves(self, jobs, tol):
        """Lowest eigenvectors, in place, for [(plan, x, plan_fp64), ...]; returns the Krylov step counts."""
        run = lambda j: eigsh(j[0], j[1], tol=tol, maxit=self.eig_maxit, krylov=self.krylov, plan_fp64=j[2], check=False)[1]  # noqa: E731
        if not self.concurrent:
            return [run(j) for j in jobs]
        self.torch.cuda.synchronize()  # inputs made on the main stream
        its = list(self._side_contexts()[2].map(run, jobs))
        self.torch.cuda.synchronize()
        return its

    def gradnorm(self, n, AC, hac):
        """GradNormRight: norm_scaled(ApplyTMmixedRight(H_AC(C AR), NR)) = ||HAC - (sum_s HAC_s AR_s^T) AR|| / sqrt(nsec),
        nsec = number of left sectors of AR[n] with a null space; returns (F, ||B||)."""
        lay = self.slay[n]
        HAC = self._empty(lay.nelem)
        hac.apply(AC, HAC)
        X, Xl = self._prod(TM_RIGHT_UNITY, lay, HAC, lay, self.AR[n], ("gX", n))
        Y, _ = self._prod(MAT_MPS, Xl, X, lay, self.AR[n], ("gY", n))
        B = float(self.torch.linalg.norm(HAC - Y).item())
        cols, rows = {}, {}
        for qi, r, c in zip(map(tuple, lay.qin), lay.rows, lay.cols):
            cols[qi] = cols.get(qi, 0) + int(c)
            rows[qi] = int(r)
        nsec = sum(1 for q in cols if cols[q] > rows[q])
        return (B / np.sqrt(nsec) if nsec else 0.0), B

    def _scaled_identity(self, lay: BlockLayout, scale):
        out = np.zeros(lay.nelem)
        for o, r, sc in zip(lay.offsets, lay.rows, scale):
            out[o:o + r * r].reshape(r, r)[np.diag_indices(r)] = sc
        return self._dev(out)

    def _diagonal(self, lay: BlockLayout, key):
        """(padded gather index, mask, flat diagonal index, sector of each diagonal entry) of the
        square blocks of `lay`, built once per key: G[index] * mask summed over rows 1 is the
        per-block trace in a fixed order."""
        k = ("diag",) + tuple(key)
        v = self._units.get(k)
        if v is None:
            rmax = max([int(r) for r in lay.rows] + [1])
            pad = np.zeros((len(lay.rows), rmax), np.int64)
            mask = np.zeros((len(lay.rows), rmax))
            flat, sec = [], []
            for q, (o, r) in enumerate(zip(lay.offsets, lay.rows)):
                d = o + np.arange(r) * (r + 1)
                pad[q, :r], mask[q, :r] = d, 1.0
                flat.append(d)
                sec.append(np.full(r, q))
            flat = np.concatenate(flat) if flat else np.zeros(0, np.int64)
            sec = np.concatenate(sec) if sec else np.zeros(0, np.int64)
            t = self.torch
            v = (t.from_numpy(pad).cuda(), self._dev(mask), t.from_numpy(flat.astype(np.int64)).cuda(), t.from_numpy(sec.astype(np.int64)).cuda())
            self._units[k] = v
        return v

    def _unit15(self, lay: BlockLayout, key):
        """1.5 x the identity on `lay`, built once per key."""
        k = ("unit15",) + tuple(key)
        u = self._units.get(k)
        if u is None:
            u = 1.5 * self._unit(lay, key)
            self._units[k] = u
        return u

    def _unit(self, lay: BlockLayout, key):
        """Identity on the block-diagonal layout `lay`, built once per key."""
        k = ("unit",) + tuple(key)
        u = self._units.get(k)
        if u is None:
            u = self._scaled_identity(lay, [1.0] * len(lay.rows))
            self._units[k] = u
        return u

    def _polar(self, X, lay, direction, key, tol=1e-14, maxit=200):
        """Isometric polar factor U V^T of the regrouped sector matrices of X (left: rows (s, q - d_s),
        right: the transposed grouping) by the Newton-Schulz iteration X <- X (3 - X^T X) / 2
        (left) or X <- (3 - X X^T) X / 2 (right) after scaling every sector to unit Frobenius norm;
        it stops when X^T X (X X^T) equals the identity to `tol`.  Stands in for the SVD-based
        polar decomposition svd(S, P, X, dir); U P of the reference."""
        t = self.torch
        kind_g = TM_LEFT_UNITY if direction == LEFT else TM_RIGHT_UNITY
        for it in range(maxit):
            G, Gl = self._prod(kind_g, lay, X, lay, X, key + ("G",))
            if it == 0:  # per-sector scaling by the Frobenius norm, sqrt(tr G_q), on the device
                pad, mask, flat, sec = self._diagonal(Gl, key)
                tr = (G[pad] * mask).sum(1)
                scale = t.where(tr > 0, 1.0 / tr.clamp_min(1e-300).sqrt(), t.zeros_like(tr))
                S = t.zeros(Gl.nelem, dtype=t.float64, device=G.device)
                S[flat] = scale[sec]
                X, _ = (self._prod(MPS_MAT, lay, X, Gl, S, key + ("s",)) if direction == LEFT
                        else self._prod(MAT_MPS, Gl, S, lay, X, key + ("s",)))
                continue
            I = self._unit(Gl, key)
            # convergence read every third iteration (one host round trip each); iterations past
            # convergence keep X orthonormal to rounding
            if it % 3 == 1:
                dev = float((G - I).abs().max().item())
                if dev < tol:
                    break
            M = t.sub(self._unit15(Gl, key), G, alpha=0.5)  # 1.5 I - 0.5 G in one kernel
            X, _ = (self._prod(MPS_MAT, lay, X, Gl, M, key + ("m",)) if direction == LEFT
                    else self._prod(MAT_MPS, Gl, M, lay, X, key + ("m",)))
        return X

    # ------------------------------------------------------------------ one outer iteration
    def sweep(self) -> VumpsStats:
        torch = self.torch
        N = self.N
        t0 = time.perf_counter()
        F, G = np.zeros(N), np.zeros(N)
        e_cell = 0.0
        it0, env0 = self.eig_iters, self.env_iters
        exp_prec = float(np.max(self.prec_n))
        for n in range(N):
            left = self._b(n - 1)
            tol = min(max(self.prec_n[n] / 100.0, self.eig_tol_min), self.eig_tol_max)
            hac, hac64 = self._hac_plans(n)
            AC, _ = self._prod(MAT_MPS, self.blay[n], self.C[left], self.slay[n], self.AR[n], ("CAR", n))
            F[n], G[n] = self.gradnorm(n, AC, hac64 or hac)
            CL = self.C[left].clone()
            CR = self.C[n].clone()
            hcl, hcl64 = self._hc_plans(n, 0 if self.concurrent else -1)
            hcr, hcr64 = self._hc_plans(n + 1, 1 if self.concurrent else -1)
            self.eig_iters += sum(self._eigensolves([(hac, AC, hac64), (hcl, CL, hcl64), (hcr, CR, hcr64)], tol))
            CL = CL if float(CL[0].item()) >= 0 else -CL  # sign convention of VUMPS_main.cpp:201-202
            CR = CR if float(CR[0].item()) >= 0 else -CR
            if self.dosvd and self.prec_n[n] < self.svdthresh:
                self.dosvd = False
            if self.dosvd:  # polar factors of AC CR^T and CL^T AC (the SVD scheme, VUMPS_main.cpp:213-220)
                X, _ = self._prod(MPS_MATT, self.slay[n], AC, self.blay[self._b(n + 1)], CR, ("ACCRt", n))
                ALn = self._polar(X, self.slay[n], LEFT, ("pl", n))
                X, _ = self._prod(MATT_MPS, self.blay[n], CL, self.slay[n], AC, ("CLtAC", n))
                ARn = self._polar(X, self.slay[n], RIGHT, ("pr", n))
            else:         # QR scheme (:224-232)
                QLa, _ = self._qr(AC, self.slay[n], LEFT)
                QLc, _ = self._qr(CR, self.blay[self._b(n + 1)], LEFT)
                QRa, _ = self._qr(AC, self.slay[n], RIGHT)
                QRc, _ = self._qr(CL, self.blay[n], RIGHT)
                ALn, _ = self._prod(MPS_MATT, self.slay[n], QLa, self.blay[self._b(n + 1)], QLc, ("ALnew", n))
                ARn, _ = self._prod(MATT_MPS, self.blay[n], QRc, self.slay[n], QRa, ("ARnew", n))
            # precision of the gauge: max(||AC - AL CR||_inf, ||AC - CL AR||_inf)
            ALCR, _ = self._prod(MPS_MAT, self.slay[n], ALn, self.blay[self._b(n + 1)], CR, ("ALCR", n))
            CLAR, _ = self._prod(MAT_MPS, self.blay[n], CL, self.slay[n], ARn, ("CLAR", n))
            self.prec_n[n] = max(float((AC - ALCR).abs().max().item()), float((AC - CLAR).abs().max().item()))
            self.AL[n], self.AR[n] = ALn, ARn
            self.C[left], self.C[n] = CL, CR
            # environments (VUMPS_main.cpp:247-305)
            etol = max(tol / self.env_fac, self.env_tol)
            self.LE[n], e_cell, RE_next = self._envs(n, etol)
            self.LE[self._b(n + 1)] = self._tm(LEFT, self.AL[n], n, self.LE[n], n, self._b(n + 1)) + self._optm_left(n - 1)
            self.LE[self._b(n + 2)] = (self._tm(LEFT, self.AL[self._b(n + 1)], self._b(n + 1), self.LE[self._b(n + 1)], self._b(n + 1),
                                                self._b(n + 2)) + self._optm_left(n))
            self.RE[self._b(n + 2)] = RE_next
            self.RE[self._b(n + 1)] = (self._tm(RIGHT, self.AR[self._b(n + 1)], self._b(n + 1), self.RE[self._b(n + 2)], self._b(n + 2),
                                                self._b(n + 1)) + self._optm_right(n + 1))
        self._maybe_expand(exp_prec)
        torch.cuda.synchronize()
        self.iteration = getattr(self, "iteration", 0) + 1
        return VumpsStats(self.iteration, e_cell / N, list(F), list(G), time.perf_counter() - t0, self.eig_iters - it0,
                          self.env_iters - env0)

    # ------------------------------------------------------------------ bond growth
    def bond_dims(self) -> List[int]:
        return [int(np.sum(sz)) for (_, sz) in self.bonds]

    def expand(self, b: int, dm: int) -> int:
        """ExpandFromH on bond b (between AL[b-1] and AR[b], carrying C[b-1]): at most dm states
        from the two-site residual H (AL C AR) projected on the null spaces (expand.py); the state
        is unchanged (C padded with zeros), AR[b-1] and AL[b] get zero columns / rows for the new
        states, and the environments on bond b are rebuilt from their neighbours as in
        VUMPS_main.cpp:286-301.  Returns the number of states added."""
        from . import i2k
        from .expand import append_columns, append_rows, diag_blocks, expand_from_h, flatten_diag, flatten_mps, grow_bond, mps_blocks
        N = self.N
        if N < 2:
            raise ValueError("bond growth needs a unit cell of at least two sites")
        b = self._b(b)
        l, r = self._b(b - 1), b
        host = lambda t: t.cpu().numpy()  # noqa: E731
        AL_l = mps_blocks(self.slay[l], host(self.AL[l]))
        AR_r = mps_blocks(self.slay[r], host(self.AR[r]))
        AR_l = mps_blocks(self.slay[l], host(self.AR[l]))
        AL_r = mps_blocks(self.slay[r], host(self.AL[r]))
        Cb = diag_blocks(self.blay[b], host(self.C[l]))
        dq = i2k(self.model, N, self.QN)
        add, new_cols, new_rows, _ = expand_from_h(self.H, dq, AL_l, Cb, AR_r, dm, self.periods)
        if not add:
            return 0
        keys, sizes = grow_bond(self.bonds[b][0], self.bonds[b][1], add)
        self.bonds[b] = (keys, sizes)
        self.blay[b] = BlockLayout.diag(self.nsym, keys, sizes, self.periods)
        for n in (l, r):
            self.slay[n] = mps_structure(self.model, N, self.QN, self.bonds[n], self.bonds[self._b(n + 1)])
        self.AL[l] = self._dev(flatten_mps(self.slay[l], append_columns(AL_l, new_cols)))
        self.AR[l] = self._dev(flatten_mps(self.slay[l], AR_l))
        self.AR[r] = self._dev(flatten_mps(self.slay[r], append_rows(AR_r, new_rows)))
        self.AL[r] = self._dev(flatten_mps(self.slay[r], AL_r))
        self.C[l] = self._dev(flatten_diag(self.blay[b], Cb))
        # structures changed: plans and device index tables are rebuilt on demand
        self._plans = {}
        self._units = {}
        # the environments on bond b from the (unchanged) neighbouring bonds
        self.LE[b] = self._tm(LEFT, self.AL[l], l, self.LE[l], l, b) + self._optm_left(l - 1)
        self.RE[b] = self._tm(RIGHT, self.AR[r], r, self.RE[self._b(r + 1)], self._b(r + 1), r) + self._optm_right(r)
        return int(sum(add.values()))

    def truncate(self, b: int, mmax: int, lamthresh: float = 1e-10) -> int:
        """Truncate (MPSBlockUtilities.hpp:846-936) on bond b: the bond is brought to its Schmidt
        basis (C[b-1] = U S V^T per sector; AL[b-1] U, AR[b-1] V, U^T AL[b], V^T AR[b], C = S: the
        state and the isometries are unchanged), then only the Schmidt values above lamthresh and, of
        those, the mmax largest are kept (expand.truncate_sizes): each sector keeps its leading
        states and a sector with none left is dropped.  C is renormalised to unit norm.  AL[b-1] and
        AR[b] stay isometric; AR[b-1] and AL[b] are only approximately so until the next iteration.
        Plans are rebuilt and the environments on bond b recomputed, as after a growth.  Returns the
        number of states removed."""
        from .expand import (cut_left_index, cut_right_index, diag_blocks, flatten_diag, flatten_mps, mps_blocks, rotate_left_index,
                             rotate_right_index, schmidt_basis, truncate_sizes)
        N = self.N
        b = self._b(b)
        l, r = self._b(b - 1), b
        host = lambda t: t.cpu().numpy()  # noqa: E731
        Cb = diag_blocks(self.blay[b], host(self.C[l]))
        U, S, V = schmidt_basis(Cb)
        keep, changed = truncate_sizes(S, mmax, lamthresh)
        if not changed:
            return 0
        removed = int(sum(len(S[q]) - keep[q] for q in S))
        AL_l = cut_right_index(rotate_right_index(mps_blocks(self.slay[l], host(self.AL[l])), U), keep)
        AR_l = cut_right_index(rotate_right_index(mps_blocks(self.slay[l], host(self.AR[l])), V), keep)
        AL_r = cut_left_index(rotate_left_index(mps_blocks(self.slay[r], host(self.AL[r])), U), keep)
        AR_r = cut_left_index(rotate_left_index(mps_blocks(self.slay[r], host(self.AR[r])), V), keep)
        nrm = np.sqrt(sum(float(np.sum(S[q][:keep[q]] ** 2)) for q in S))
        Cn = {q: np.diag(S[q][:keep[q]] / nrm) for q in S if keep[q] > 0}
        qs = sorted(Cn)
        self.bonds[b] = (np.asarray(qs, np.int32).reshape(len(qs), -1), np.asarray([keep[q] for q in qs], np.int32))
        self.blay[b] = BlockLayout.diag(self.nsym, self.bonds[b][0], self.bonds[b][1], self.periods)
        for n in (l, r):
            self.slay[n] = mps_structure(self.model, N, self.QN, self.bonds[n], self.bonds[self._b(n + 1)])
        self.AL[l] = self._dev(flatten_mps(self.slay[l], AL_l))
        self.AR[l] = self._dev(flatten_mps(self.slay[l], AR_l))
        self.AL[r] = self._dev(flatten_mps(self.slay[r], AL_r))
        self.AR[r] = self._dev(flatten_mps(self.slay[r], AR_r))
        self.C[l] = self._dev(flatten_diag(self.blay[b], Cn))
        self._plans = {}
        self._units = {}
        self.LE[b] = self._tm(LEFT, self.AL[l], l, self.LE[l], l, b) + self._optm_left(l - 1)
        self.RE[b] = self._tm(RIGHT, self.AR[r], r, self.RE[self._b(r + 1)], self._b(r + 1), r) + self._optm_right(r)
        self.prec_n[l] = self.prec_n[r] = 1.0  # the next solves start from the loose end of the tolerance range
        return removed

    def _maybe_expand(self, exp_prec: float):
        """The growth policy of VUMPS_main.cpp:273-306, applied to every bond after an iteration."""
        from .expand import schmidt_values
        if self.mmax is None:
            return
        for b in range(self.N):
            total = int(np.sum(self.bonds[b][1]))
            mmax = int(self.mmax[b]) if isinstance(self.mmax, (list, tuple, np.ndarray)) else int(self.mmax)
            grow = self.expand_ct[b] >= self.expand_pause and exp_prec < self.expand_thresh and total < mmax
            if grow:  # the Schmidt values (a host SVD per sector) only when the cheap conditions hold
                lam = schmidt_values(self.blay[b], self.C[self._b(b - 1)].cpu().numpy())
                grow = max((float(v.min()) for v in lam.values() if v.size), default=0.0) > self.lamthresh
            if grow:
                self.expand_ct[b] = 0
                self.expand(b, min(self.dm_max, mmax - total))
            else:
                self.expand_ct[b] += 1

    # ------------------------------------------------------------------ observables
    @staticmethod
    def model_operators(model: int) -> Dict[str, np.ndarray]:
        """The one-site operators of the models' GetObservables, as d x d matrices: XXZ and QIM
        (XYZAbstractModel, Models.cpp:57-145: sx, syi, sz, rx = 2 sx, rz = 2 sz, sp = sx - syi,
        sm = sx + syi), Fermi-Hubbard (:575-635: cup, cdo, cupd, cdod, nup, ndo, sz, n, rx, rz; local
        states empty, down, up, double), quantum Potts (:432-470: X, Z) and Bose-Hubbard (:779-828:
        b, b2, n, nmax)."""
        t, d = model & 0xFF, model_d(model)
        if t in (0, 2):
            sx = np.array([[0.0, 0.5], [0.5, 0.0]])
            syi = np.array([[0.0, -0.5], [0.5, 0.0]])
            sz = np.diag([0.5, -0.5])
            return {"sx": sx, "syi": syi, "sz": sz, "rx": 2 * sx, "rz": 2 * sz, "sp": sx - syi, "sm": sx + syi}
        if t == 1:
            cup, cdo = np.zeros((4, 4)), np.zeros((4, 4))
            cup[0, 2] = cup[1, 3] = 1.0
            cdo[0, 1], cdo[2, 3] = 1.0, -1.0
            nup, ndo = np.diag([0.0, 0.0, 1.0, 1.0]), np.diag([0.0, 1.0, 0.0, 1.0])
            rx = np.zeros((4, 4))
            rx[0, 0], rx[1, 2], rx[2, 1], rx[3, 3] = 1.0, 1.0, 1.0, -1.0
            return {"cup": cup, "cdo": cdo, "cupd": cup.T.copy(), "cdod": cdo.T.copy(), "nup": nup, "ndo": ndo,
                    "sz": 0.5 * (nup - ndo), "n": nup + ndo, "rx": rx, "rz": np.diag([1.0, -1.0, 1.0, -1.0])}
        if t == 3:
            X, Z = np.zeros((d, d)), np.diag([d - 1.0] + [-1.0] * (d - 1))
            X[d - 1, 0] = 1.0
            for i in range(1, d):
                X[i - 1, i] = 1.0
            return {"X": X, "Z": Z}
        if t == 4:
            b = np.zeros((d, d))
            for i in range(1, d):
                b[i - 1, i] = np.sqrt(i)
            nmax = np.zeros((d, d))
            nmax[d - 1, d - 1] = 1.0
            return {"b": b, "b2": b @ b, "n": np.diag(np.arange(d, dtype=float)), "nmax": nmax}
        raise ValueError(f"unknown model {model}")

    @staticmethod
    def _gram(lay, data: np.ndarray) -> np.ndarray:
        """G[s, t] = <A_s | A_t> over the blocks that A_s and A_t share (same ingoing and outgoing
        charges); blocks of different charges are orthogonal, so a charge-changing operator meets
        zeros, as the reference's trace of a charged BlockMat is zero (BlockObj.hpp:1340-1350)."""
        from .expand import mps_blocks
        blk = mps_blocks(lay, data)
        n = len(blk)
        G = np.zeros((n, n))
        for a in range(n):
            for b in range(a, n):
                v = 0.0
                for q, (qo, m) in blk[a].items():
                    e = blk[b].get(q)
                    if e is not None and e[0] == qo:
                        v += float(np.vdot(m, e[1]))
                G[a, b] = G[b, a] = v
        return G

    def observables(self, ops) -> Dict[str, np.ndarray]:
        """MeasureObservables (MPSBlockUtilities.hpp:1017-1073) in the center gauge: for every site n
        of the cell, <O>_n = sum_{s,t} O[s, t] <A_s | A_t> / sum_s <A_s | A_s> with A = AC[n] =
        C[n-1] AR[n] for a one-site operator (d x d) and A = AC[n] AR[n+1] (pair index s = s_l d + s_r)
        for a two-site one (d^2 x d^2, e.g. the Hamiltonian, whose values are the bond energies).
        `ops`: names of the model's operators (model_operators) or a dict {name: matrix}.  Operators
        that change the charges have zero expectation value in these symmetric states."""
        table = self.model_operators(self.model)
        if isinstance(ops, dict):
            items = list(ops.items())
        else:
            items = []
            for name in ops:
                if name not in table:
                    raise ValueError(f"operator {name} not defined (operators of this model: {', '.join(table)})")
                items.append((name, table[name]))
        out = {}
        grams = {}
        for name, O in items:
            O = np.asarray(O, np.float64)
            sites = {self.d: 1, self.d * self.d: 2}.get(O.shape[0])
            if O.shape != (O.shape[0], O.shape[0]) or sites is None:
                raise ValueError(f"operator {name}: shape {O.shape}, expected d x d or d^2 x d^2 (d = {self.d})")
            vals = np.zeros(self.N)
            for n in range(self.N):
                if (sites, n) not in grams:
                    AC, lay = self._prod(MAT_MPS, self.blay[n], self.C[self._b(n - 1)], self.slay[n], self.AR[n], ("obsAC", n))
                    if sites == 2:
                        m = self._b(n + 1)
                        AC, lay = self._prod(MPS_MPS, lay, AC, self.slay[m], self.AR[m], ("obsAA", n))
                    grams[(sites, n)] = self._gram(lay, AC.cpu().numpy())
                G = grams[(sites, n)]
                vals[n] = float(np.sum(O * G) / np.trace(G))
            out[name] = vals
        return out

    # ------------------------------------------------------------------ state files
    def save(self, path: str):
        """saveIMPS (MPSBlockUtilities.hpp:1212-1266) of the unit cell: Schmidt values, C, AL, AR in
        the randMPS_LR cell convention (imps_io.py)."""
        from .expand import schmidt_values
        from .imps_io import write_imps
        N = self.N
        host = [(self.blay[self._b(n + 1)], self.C[n].cpu().numpy()) for n in range(N)]
        lam = [schmidt_values(lay, c) for lay, c in host]
        write_imps(path, lam, host, [(self.slay[n], self.AL[n].cpu().numpy()) for n in range(N)],
                   [(self.slay[n], self.AR[n].cpu().numpy()) for n in range(N)])

    def load(self, path: str):
        """loadIMPS (MPSBlockUtilities.hpp:1268-1325): replace the state by a file's.  The bonds come
        from the stored C sectors; the stored blocks are placed in this model's block structure
        (blocks the file leaves out are zero); environments and plans are rebuilt."""
        from .expand import diag_blocks, flatten_diag, flatten_mps, mps_blocks
        from .imps_io import read_imps
        _, C, AL, AR = read_imps(path, self.nsym, self.periods)
        N = self.N
        if not (len(C) == len(AL) == len(AR) == N):
            raise ValueError(f"state file holds a {len(AL)}-site cell, this VUMPS has {N} sites")
        bonds = [None] * N
        for n in range(N):
            lay = C[n][0]
            bonds[self._b(n + 1)] = (np.array(lay.qin, np.int32), np.array(lay.rows, np.int32))
        blay = [BlockLayout.diag(self.nsym, k, s, self.periods) for (k, s) in bonds]
        slay = [mps_structure(self.model, N, self.QN, bonds[n], bonds[self._b(n + 1)]) for n in range(N)]

        def embed(lay_new, lay_old, data):
            blocks = mps_blocks(lay_old, data)
            want = {(e, tuple(int(v) for v in q)): (int(r), int(c)) for e in range(lay_new.nlist)
                    for q, r, c in zip(lay_new.entry(e).qin, lay_new.entry(e).rows, lay_new.entry(e).cols)}
            for e, ent in enumerate(blocks):
                for q, (_, m) in ent.items():
                    if want.get((e, q)) != m.shape:
                        raise ValueError(f"state file block {q} of physical index {e} does not fit the model's structure")
            return self._dev(flatten_mps(lay_new, blocks))

        self.bonds, self.blay, self.slay = bonds, blay, slay
        self.AL = [embed(slay[n], *AL[n]) for n in range(N)]
        self.AR = [embed(slay[n], *AR[n]) for n in range(N)]
        self.C = [self._dev(flatten_diag(blay[self._b(n + 1)], diag_blocks(*C[n]))) for n in range(N)]
        self._plans, self._units = {}, {}
        self.LE, self.RE = {}, {}
        self.prec_n = np.ones(N)
        self.expand_ct = [0] * N
        self._init_environments()

    def energy(self) -> float:
        b = 0
        _, e_cell = self._left_env(b, self.env_tol)
        return e_cell / self.N
