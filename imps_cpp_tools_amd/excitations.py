"""Quasiparticle excitations: the momentum-k effective Hamiltonian on the GPU.

Restates ApplyHeff (Excitations_helpers.hpp:20-228) with its constants (HeffConstants, :260-347),
the excitation sectors (XDims, :639-678) and the geometric sums InvertE_fac / InvertE_proj_fac
(InvEBlockFunctions.hpp:268-476) over the library's grouped block products.  Every contraction runs
in the HIP library (FP64 products through the C ABI); this module keeps device buffers, sums
objects with different sector sets and runs the GMRES of the geometric sums on device vectors.

Complex objects (the excitation parameters X, B = NL X and everything derived from them) are held
as planar device arrays (2, n): real parts, then imaginary parts.  The MPS tensors and operators
are real, so a complex product is two real launches.  The two-site contractions of ApplyHeff
(ApplyOperator followed by a contraction with AL^T or AR^T) are evaluated through reduced
operators: sum_k AL[k]^T ApplyOperator(AL B, H)[k d + s] = sum_j HL[s][j] B[j] with
HL = ReducedOp(H, AL, l) (MPSBlockUtilities.hpp:39-74), and the mixed forms with B in the middle
through products with the single tensors AL[k]^T / AR[k]^T and the coefficients of H.

The dominant eigenmatrices of the mixed transfer matrix (AL, AR), needed to project the geometric
sums at K = 0 and p = 0 (LRoverlaps, Excitations_helpers.hpp:915-946), are taken in closed form:
with AL C = C AR, LM = C^T and RM = C on bond 0 (eigenvalue 1, tr(LM RM) = 1).
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import (FP64, LEFT, RIGHT, MAT_MPS, MATT_MPS, MPS_MAT, MPS_MATT, MPS_MPS, MPS_REDOP, REDOP_MPS, TM_LEFT_UNITY,
               TM_RIGHT_UNITY, BlockLayout, Context, EnvPlan, ImpsError, ProdPlan, RedOpPlan)


def _lkey(lay: BlockLayout):
    k = getattr(lay, "_key", None)
    if k is None:
        k = (lay.nsym, lay.nblk.tobytes(), lay.qin.tobytes(), lay.qout.tobytes(), lay.rows.tobytes(), lay.cols.tobytes())
        lay._key = k
    return k


def _index(lay: BlockLayout) -> Dict[Tuple[int, tuple], Tuple[tuple, int, int, int]]:
    """(entry, qin) -> (qout, rows, cols, element offset)."""
    idx = getattr(lay, "_idx", None)
    if idx is None:
        idx = {}
        off = lay.offsets
        b = 0
        for e, nb in enumerate(lay.nblk):
            for _ in range(int(nb)):
                idx[(e, tuple(int(v) for v in lay.qin[b]))] = (tuple(int(v) for v in lay.qout[b]), int(lay.rows[b]), int(lay.cols[b]), int(off[b]))
                b += 1
        lay._idx = idx
    return idx


def _layout_from_index(nsym: int, nlist: int, idx) -> BlockLayout:
    nblk, qin, qout, rows, cols = [], [], [], [], []
    for e in range(nlist):
        keys = sorted(q for (ee, q) in idx if ee == e)
        nblk.append(len(keys))
        for q in keys:
            qo, r, c = idx[(e, q)][:3]
            qin.append(q)
            qout.append(qo)
            rows.append(r)
            cols.append(c)
    return BlockLayout(nsym, np.array(nblk, np.int32), np.array(qin, np.int32).reshape(-1, nsym),
                       np.array(qout, np.int32).reshape(-1, nsym), np.array(rows, np.int32), np.array(cols, np.int32))


class DObj:
    """Device block object: a layout and its data, real (n,) or complex planar (2, n)."""
    __slots__ = ("lay", "data")

    def __init__(self, lay: BlockLayout, data):
        self.lay = lay
        self.data = data

    @property
    def cplx(self) -> bool:
        return self.data.dim() == 2


class BlockOps:
    """Block-sparse objects on the device and the reference's block algebra over them: products
    through the library's grouped GEMM (ProdPlan, FP64), sums over the union of sector sets,
    restriction onto a sector set, complex scaling, unit-cell transfer-matrix chains, GMRES for
    the geometric sums and Arnoldi for transfer-matrix spectra."""

    def __init__(self, ctx: Context, nsym: int, d: int = 2, H: Optional[np.ndarray] = None, E0: float = 0.0, tol: float = 1e-14):
        import torch
        self.torch = torch
        self.dev = "cuda"
        self.ctx = ctx
        self.nsym = nsym
        self.d = d
        self.tol = tol
        self.H0 = None if H is None else np.array(H, np.float64) - E0 * np.eye(d * d)
        self._plans: Dict = {}
        self._maps: Dict = {}

    # ---- helpers -----------------------------------------------------------------------------
    def _plan(self, key, make):
        p = self._plans.get(key)
        if p is None:
            p = make()
            self._plans[key] = p
        return p

    def _zeros(self, n, cplx=True):
        t = self.torch
        return t.zeros((2, n) if cplx else (n,), dtype=t.float64, device=self.dev)

    def _transpose_blocks(self, x: DObj):
        t = self.torch.empty_like(x.data)
        for (e, q), (qo, r, c, o) in _index(x.lay).items():
            if x.cplx:
                t[:, o:o + r * c] = x.data[:, o:o + r * c].view(2, c, r).transpose(1, 2).reshape(2, -1)
            else:
                t[o:o + r * c] = x.data[o:o + r * c].view(c, r).t().reshape(-1)
        return t

    def _redop(self, A: DObj, direction):
        p = self._plan(("redop", _lkey(A.lay), direction), lambda: RedOpPlan(self.ctx, self.H0, self.d, A.lay, direction))
        y = self._zeros(p.out_layout.nelem, False)
        p.apply(A.data, y)
        return DObj(p.out_layout, y)

    def prod(self, kind: int, X1: DObj, X2: DObj) -> DObj:
        """Block product through the library (grouped GEMM); complex operands as two launches."""
        if X1.cplx and X2.cplx:
            raise ValueError("product of two complex objects")
        p = self._plan(("prod", kind, _lkey(X1.lay), _lkey(X2.lay)), lambda: ProdPlan(self.ctx, kind, X1.lay, X2.lay))
        cplx = X1.cplx or X2.cplx
        ny = p.out_layout.nelem
        if not (ny and X1.lay.nelem and X2.lay.nelem):
            return DObj(p.out_layout, self._zeros(ny, cplx))
        y = self.torch.empty((2, ny) if cplx else (ny,), dtype=self.torch.float64, device=self.dev)  # fully overwritten
        if X1.cplx:
            p.apply_n(X1.data.contiguous(), X2.data, y, 2)  # real and imaginary parts in one launch
        elif X2.cplx:
            p.apply(X1.data, X2.data[0], y[0])
            p.apply(X1.data, X2.data[1], y[1])
        else:
            p.apply(X1.data, X2.data, y)
        return DObj(p.out_layout, y)

    def scale(self, x: DObj, a: complex) -> DObj:
        """a x as a complex object (x itself when x is complex and a = 1: results are never
        written in place)."""
        ar, ai = float(np.real(a)), float(np.imag(a))
        if x.cplx:
            if ai == 0.0:
                return x if ar == 1.0 else DObj(x.lay, x.data * ar)
            re, im = x.data[0], x.data[1]
            return DObj(x.lay, self.torch.stack((ar * re - ai * im, ar * im + ai * re)))
        if ai == 0.0:
            out = self._zeros(x.lay.nelem, True)
            out[0] = x.data if ar == 1.0 else x.data * ar
            return DObj(x.lay, out)
        return DObj(x.lay, self.torch.stack((ar * x.data, ai * x.data)))

    def _union_host(self, la: BlockLayout, lb: BlockLayout):
        """(union layout, positions of la's elements in it, positions of lb's) on the host."""
        key = ("union_host", _lkey(la), _lkey(lb))
        m = self._maps.get(key)
        if m is None:
            if len(la.nblk) != len(lb.nblk):
                raise ValueError("sum of objects with different list lengths")
            ia_, ib_ = _index(la), _index(lb)
            u = dict(ia_)
            for k, v in ib_.items():
                if k in u and (u[k][0], u[k][1], u[k][2]) != (v[0], v[1], v[2]):
                    raise ValueError("sum of blocks with different shapes or outgoing charges")
                u.setdefault(k, v)
            lu = _layout_from_index(la.nsym, len(la.nblk), u)
            iu = _index(lu)

            def positions(src):
                out = np.zeros(sum(r * c for (_, r, c, _) in src.values()), np.int64)
                for k, (qo, r, c, o) in src.items():
                    out[o:o + r * c] = iu[k][3] + np.arange(r * c)
                return out
            m = (lu, positions(ia_), positions(ib_))
            self._maps[key] = m
        return m

    def _union(self, la: BlockLayout, lb: BlockLayout):
        key = ("union", _lkey(la), _lkey(lb))
        m = self._maps.get(key)
        if m is None:
            lu, pa, pb = self._union_host(la, lb)
            t = self.torch
            m = (lu, t.from_numpy(pa).to(self.dev), t.from_numpy(pb).to(self.dev))
            self._maps[key] = m
        return m

    def _combine_rows(self, key, src, parts, rows) -> DObj:
        """concat over rows of combine(terms): the same layout as the chain of adds (sector sets
        united in term order), computed as one gather of the source elements, a scaling and a sum
        in term order per output element.  src: the sources' data side by side ((2, n) or (n,));
        parts[p] = (layout, element offset in src); rows[r] = [(coefficient, part), ...]."""
        t = self.torch
        m = self._maps.get(key)
        if m is None:
            out_lays, trip, base = [], [], 0
            for terms in rows:
                if not terms:
                    out_lays.append(BlockLayout(self.nsym, [0], np.zeros((0, self.nsym)), np.zeros((0, self.nsym)), [], []))
                    continue
                c0, p0 = terms[0]
                lay = parts[p0][0]
                placed = [(np.arange(lay.nelem), p0, c0)]
                for c, p in terms[1:]:
                    lay, pa, pb = self._union_host(lay, parts[p][0])
                    placed = [(pa[pos], q, cq) for pos, q, cq in placed]
                    placed.append((pb, p, c))
                out_lays.append(lay)
                for pos, q, cq in placed:
                    trip.append((base + pos, parts[q][1] + np.arange(parts[q][0].nelem), np.full(len(pos), float(cq))))
                base += lay.nelem
            lay = BlockLayout.concat(out_lays)
            if trip:
                outs = np.concatenate([a for a, _, _ in trip])
                ins = np.concatenate([b for _, b, _ in trip])
                cs = np.concatenate([c for _, _, c in trip])
                order = np.argsort(outs, kind="stable")  # term order within each output element
                outs, ins, cs = outs[order], ins[order], cs[order]
                start = np.searchsorted(outs, np.arange(base))
                rank = np.arange(len(outs)) - start[outs]
                width = int(rank.max()) + 1
                idx = np.zeros((base, width), np.int64)
                coef = np.zeros((base, width))
                idx[outs, rank] = ins
                coef[outs, rank] = cs
                m = (lay, t.from_numpy(idx).to(self.dev), t.from_numpy(coef).to(self.dev))
            else:
                m = (lay, None, None)
            self._maps[key] = m
        lay, idx, coef = m
        if idx is None:
            return DObj(lay, self._zeros(0, True))
        if src.dim() == 1:
            src = t.stack((src, t.zeros_like(src)))
        return DObj(lay, (src[:, idx] * coef).sum(-1))

    def add(self, a: DObj, b: DObj, alpha=1.0, beta=1.0) -> DObj:
        """alpha a + beta b on the union of the sector sets (BlockMat::operator+=)."""
        sa, sb = self.scale(a, alpha).data, self.scale(b, beta).data
        if _lkey(a.lay) == _lkey(b.lay):
            return DObj(a.lay, sa + sb)
        lu, ia, ib = self._union(a.lay, b.lay)
        out = self._zeros(lu.nelem, True)
        out.index_add_(1, ia, sa)
        out.index_add_(1, ib, sb)
        return DObj(lu, out)

    def onto(self, x: DObj, lay: BlockLayout) -> DObj:
        """x restricted to / zero-extended onto the sector set `lay` (x itself when its structure
        is `lay`: results are never written in place)."""
        if _lkey(x.lay) == _lkey(lay):
            return x
        key = ("onto", _lkey(x.lay), _lkey(lay))
        m = self._maps.get(key)
        if m is None:
            ix, il = _index(x.lay), _index(lay)
            src, dst = [], []
            for k, (qo, r, c, o) in il.items():
                v = ix.get(k)
                if v is not None and (v[0], v[1], v[2]) == (qo, r, c):
                    src.append(np.arange(v[3], v[3] + r * c))
                    dst.append(np.arange(o, o + r * c))
            t = self.torch
            cat = lambda a: t.from_numpy(np.concatenate(a) if a else np.zeros(0, np.int64)).to(self.dev)  # noqa: E731
            m = (cat(src), cat(dst))
            self._maps[key] = m
        out = self._zeros(lay.nelem, x.cplx)
        if x.cplx:
            out[:, m[1]] = x.data[:, m[0]]
        else:
            out[m[1]] = x.data[m[0]]
        return DObj(lay, out)

    def entry(self, x: DObj, e: int) -> DObj:
        a, b = x.lay.entry_range(e)
        return DObj(x.lay.entry(e), x.data[..., a:b])

    def concat(self, parts: Sequence[DObj]) -> DObj:
        cplx = any(p.cplx for p in parts)
        datas = [p.data if p.cplx or not cplx else self.torch.stack((p.data, self.torch.zeros_like(p.data))) for p in parts]
        return DObj(BlockLayout.concat([p.lay for p in parts]), self.torch.cat(datas, dim=-1).contiguous())

    def combine(self, terms: Sequence[Tuple[complex, DObj]]) -> DObj:
        out = None
        for c, x in terms:
            out = self.scale(x, c) if out is None else self.add(out, x, 1.0, c)
        return out

    # ---- restated block operations -----------------------------------------------------------
    def mps_mat(self, A, M):
        return self.prod(MPS_MAT, A, M)

    def mat_mps(self, M, A):
        return self.prod(MAT_MPS, M, A)

    def mps_mps(self, A, B):
        return self.prod(MPS_MPS, A, B)

    def tml_unity(self, A, B):
        return self.prod(TM_LEFT_UNITY, A, B)

    def tmr_unity(self, A, B):
        return self.prod(TM_RIGHT_UNITY, A, B)

    def tml(self, A, B, X):
        return self.tml_unity(self.mat_mps(X, A), B)

    def tmr(self, A, B, X):
        return self.tmr_unity(self.mps_mat(A, X), B)

    def apply_operator(self, AA: DObj, H: np.ndarray) -> DObj:
        """out[row] = sum_col H(row, col) AA[col] (ApplyOperator, MPSBlockUtilities.hpp:18-34)."""
        D = H.shape[0]
        parts = [(AA.lay.entry(e), AA.lay.entry_range(e)[0]) for e in range(D)]
        rows = [[(H[row, col], col) for col in range(D) if H[row, col] != 0.0] for row in range(D)]
        return self._combine_rows(("aop", _lkey(AA.lay), H.tobytes()), AA.data, parts, rows)

    def diag_part(self, x: DObj) -> DObj:
        idx = {k: v for k, v in _index(x.lay).items() if k[1] == v[0]}
        return self.onto(x, _layout_from_index(x.lay.nsym, len(x.lay.nblk), idx))

    def _grid(self, name, T: List[DObj], coef) -> DObj:
        """Row-major d x d grid of combinations of the entries T[k][i]: entry (a, b) is
        sum_{k,i} coef(a, b, k, i) T[k][i] over the nonzero coefficients, in (k, i) order."""
        d = self.d
        parts, off = [], 0
        for Tk in T:
            for i in range(d):
                parts.append((Tk.lay.entry(i), off + Tk.lay.entry_range(i)[0]))
            off += Tk.lay.nelem
        rows = [[(coef(a, b, k, i), k * d + i) for k in range(d) for i in range(d) if coef(a, b, k, i) != 0.0]
                for a in range(d) for b in range(d)]
        src = self.torch.cat([Tk.data for Tk in T], dim=-1)
        return self._combine_rows((name, tuple(_lkey(Tk.lay) for Tk in T), self.H0.tobytes()), src, parts, rows)

    def redop_left_mixed(self, Abra: DObj, Bket: DObj) -> DObj:
        """M[s][j] = sum_{k,i} H(k d + s, i d + j) Abra[k]^T Bket[i] (d x d grid, row-major)."""
        d, H = self.d, self.H0
        T = [self.prod(MATT_MPS, self.entry(Abra, k), Bket) for k in range(d)]  # T[k][i] = Abra[k]^T Bket[i]
        return self._grid("rlm", T, lambda s, j, k, i: H[k * d + s, i * d + j])

    def redop_right_mixed(self, Aket: DObj, Bbra: DObj) -> DObj:
        """M[s][i] = sum_{j,k} H(s d + k, i d + j) Aket[j] Bbra[k]^T (d x d grid, row-major)."""
        d, H = self.d, self.H0
        U = [self.prod(MPS_MATT, Aket, self.entry(Bbra, k)) for k in range(d)]  # U[k][j] = Aket[j] Bbra[k]^T
        return self._grid("rrm", U, lambda s, i, k, j: H[s * d + k, i * d + j])

    # ---- geometric sums ------------------------------------------------------------------------
    def _gmres(self, afun, b, x0, tol):
        """Unrestarted GMRES with complex Givens rotations on complex device vectors (gmres,
        IterativeSolvers.hpp:21-149; same stopping rule |g(j+1)| < tol ||b||).  The Arnoldi vectors
        are orthogonalised by classical Gram-Schmidt applied twice (the reference: modified
        Gram-Schmidt) on a device basis, so that a step makes two batched inner-product passes and
        one host transfer (coefficients and norm) instead of one per basis vector."""
        t = self.torch
        dim = b.numel()
        kry = dim
        bnorm = t.linalg.vector_norm(b).item()
        if bnorm == 0.0:
            return t.zeros_like(b), 0
        r = b - afun(x0)
        rnorm = t.linalg.vector_norm(r).item()
        cap = min(kry + 1, 64)
        V = t.empty((cap, dim), dtype=b.dtype, device=b.device)
        Vc = t.empty_like(V)  # conjugated basis: the inner products read it without conjugating V per step
        V[0] = r / rnorm
        Vc[0] = V[0].conj()
        # Hessenberg matrix and rotations sized by the basis capacity, grown with it
        Hm = np.zeros((cap, cap), complex)
        c = np.zeros(cap, complex)
        s = np.zeros(cap, complex)
        g = np.zeros(cap, complex)
        g[0] = rnorm
        kdim, flag = 0, -1
        for j in range(kry):
            if j + 2 > Hm.shape[0]:  # row j + 1 is written in this step
                n2 = min(kry + 1, 2 * Hm.shape[0])
                H2 = np.zeros((n2, n2), complex)
                H2[:Hm.shape[0], :Hm.shape[1]] = Hm
                Hm = H2
                c, s, g = (np.concatenate([a, np.zeros(n2 - a.size, complex)]) for a in (c, s, g))
            w = afun(V[j])
            Vj, Vcj = V[:j + 1], Vc[:j + 1]
            h1 = t.mv(Vcj, w)
            w = t.addmv(w, Vj.transpose(0, 1), h1, alpha=-1)
            h2 = t.mv(Vcj, w)
            w = t.addmv(w, Vj.transpose(0, 1), h2, alpha=-1)
            cbuf = t.empty(j + 2, dtype=b.dtype, device=b.device)  # coefficients and ||w||, one transfer
            t.add(h1, h2, out=cbuf[:j + 1])
            cbuf[j + 1] = t.linalg.vector_norm(w)
            coef = cbuf.cpu().numpy()
            Hm[:j + 1, j] = coef[:j + 1]
            Hm[j + 1, j] = coef[j + 1].real
            for i in range(j):
                h1_, h2_ = Hm[i, j], Hm[i + 1, j]
                Hm[i, j] = c[i] * h1_ + s[i] * h2_
                Hm[i + 1, j] = -np.conj(s[i]) * h1_ + c[i] * h2_
            beta = np.linalg.norm(Hm[j:j + 2, j])
            if abs(beta) < 2e-16:
                flag = -2
                break
            if j + 1 >= V.shape[0]:
                n2 = min(kry + 1, 2 * V.shape[0])
                grown = t.empty((n2, dim), dtype=b.dtype, device=b.device)
                grown[:V.shape[0]] = V
                V = grown
                grown = t.empty_like(V)
                grown[:Vc.shape[0]] = Vc
                Vc = grown
            t.div(w, complex(Hm[j + 1, j]), out=V[j + 1])
            t.conj_physical(V[j + 1], out=Vc[j + 1])
            habs = abs(Hm[j, j])
            phase = Hm[j, j] / habs if habs > 0 else 1.0
            s[j] = phase * np.conj(Hm[j + 1, j]) / beta
            c[j] = habs / beta
            Hm[j, j] = phase * beta
            Hm[j + 1, j] = 0
            g[j + 1] = -np.conj(s[j]) * g[j]
            g[j] *= c[j]
            kdim += 1
            if abs(g[j + 1]) < tol * bnorm:
                flag = kdim
                break
        x = x0.clone()
        if kdim:
            alpha = np.linalg.solve(np.triu(Hm[:kdim, :kdim]), g[:kdim])
            x = x + t.mv(V[:kdim].transpose(0, 1), t.from_numpy(alpha).to(device=b.device, dtype=b.dtype))
        return x, flag

    def _c(self, x: DObj):
        return self.torch.complex(x.data[0], x.data[1])

    def _p(self, z, lay) -> DObj:
        return DObj(lay, self.torch.stack((z.real, z.imag)).contiguous())

    def _tr(self, a: DObj, b: DObj, device: bool = False):
        """trace(a * b) summed over sectors (a real or complex, b real or complex): sum over the
        matching blocks of A_ij B_ji, as one gather of both operands (cached element pairs); a
        Python complex after one host transfer, or (device=True) a 0-d complex device tensor."""
        t = self.torch
        key = ("tr", _lkey(a.lay), _lkey(b.lay))
        m = self._maps.get(key)
        if m is None:
            ia, ib = _index(a.lay), _index(b.lay)
            xa, xb = [], []
            for (e, q), (qo, r, c, o) in ia.items():
                v = ib.get((e, qo))
                if v is None or v[0] != q:
                    continue
                _, r2, c2, o2 = v
                if (r2, c2) != (c, r):
                    raise ValueError("trace of blocks whose shapes do not match")
                i, j = np.meshgrid(np.arange(r), np.arange(c), indexing="ij")
                xa.append((o + i + j * r).ravel())   # A_ij, column-major
                xb.append((o2 + j + i * r2).ravel())  # B_ji
            cat = lambda parts: t.from_numpy(np.concatenate(parts) if parts else np.zeros(0, np.int64)).to(self.dev)  # noqa: E731
            m = (cat(xa), cat(xb))
            self._maps[key] = m
        ja, jb = m
        A, B = a.data[..., ja], b.data[..., jb]
        if a.cplx and b.cplx:
            re, im = (A[0] * B[0] - A[1] * B[1]).sum(), (A[0] * B[1] + A[1] * B[0]).sum()
        elif a.cplx:
            re, im = (A[0] * B).sum(), (A[1] * B).sum()
        elif b.cplx:
            re, im = (A * B[0]).sum(), (A * B[1]).sum()
        else:
            re, im = (A * B).sum(), t.zeros((), dtype=A.dtype, device=A.device)
        if device:
            return t.complex(re, im)
        v = t.stack((re, im)).cpu().numpy()
        return complex(float(v[0]), float(v[1]))

    def invert_e(self, As, Bs, x: DObj, direction, fac, L: Optional[DObj] = None, R: Optional[DObj] = None, y0=None):
        """InvertE_fac (L, R None) or InvertE_proj_fac of the mixed unit-cell TM (As, Bs)."""
        t = self.torch
        lay = x.lay
        T = (lambda v: self._chain_left(As, Bs, v)) if direction == LEFT else (lambda v: self._chain_right(As, Bs, v))
        proj = L is not None and R is not None
        b = self._c(x)
        if proj:
            if direction == LEFT:
                b = b - self._tr(x, R) * self._c(self.scale(self.onto(L, lay), 1.0))
                rin, rout = R, self.onto(L, lay)
            else:
                b = b - self._tr(L, x) * self._c(self.scale(self.onto(R, lay), 1.0))
                rin, rout = L, self.onto(R, lay)
            routc = self._c(self.scale(rout, 1.0))

        def afun(z):
            xin = self._p(z, lay)
            tx = self._c(self.onto(T(xin), lay))
            out = t.sub(z, tx, alpha=fac)  # z - fac T z
            if proj:  # the trace stays on the device: no host round trip per GMRES step
                tr = self._tr(xin, rin, device=True) if direction == LEFT else self._tr(rin, xin, device=True)
                out = out + tr * routc
            return out
        z0 = self.torch.zeros_like(b) if y0 is None else self._c(y0)
        z, flag = self._gmres(afun, b, z0, self.tol)
        return self._p(z, lay), flag

    def _chain_left(self, As, Bs, X):
        for a, b in zip(As, Bs):
            X = self.tml(a, b, X)
        return X

    def _chain_right(self, As, Bs, X):
        for a, b in zip(reversed(As), reversed(Bs)):
            X = self.tmr(a, b, X)
        return X

    # ---- transfer-matrix spectra ------------------------------------------------------------------
    def tm_sector(self, As, Bs, direction, K) -> BlockLayout:
        """Sector list of the (mixed) TM eigenproblem in charge K: blocks q -> q + K with q over the
        ingoing sectors of A[0] (r) or B[0] (l) (TMmixedEigs, EigsBlockFunctions.hpp:282-322)."""
        def ml(x: DObj):
            return {q: r for (e, q), (qo, r, c, o) in _index(x.lay).items()}
        diml, dimr = ml(As[0]), ml(Bs[0])
        if direction == LEFT:
            diml, dimr = dimr, diml
        K = tuple(int(k) for k in np.atleast_1d(K))
        idx = {}
        for q in sorted(diml):
            qk = tuple(a + b for a, b in zip(q, K))
            if qk in dimr:
                idx[(0, q)] = (qk, diml[q], dimr[qk], 0)
        return _layout_from_index(self.nsym, 1, idx)

    def tm_eigs(self, As, Bs, direction, K, nev: int = 4, tol: float = 1e-12, krylov: int = 40, maxit: int = 100, seed: int = 0):
        """Largest-magnitude eigenvalues of the unit-cell transfer matrix of (As, Bs) in charge sector
        K ("LM" mode of TMEigs / TMmixedEigs, EigsBlockFunctions.hpp:177-391; TMDominantEig :91-172
        is nev = 1 with As = Bs).  Arnoldi with modified Gram-Schmidt applied twice, restarted from
        the sum of the wanted Ritz vectors; converged when every wanted Ritz pair has residual
        <= tol * |theta|.  Returns (eigenvalues sorted by decreasing magnitude, eigenvectors as complex
        objects on the sector list, TM applications)."""
        t = self.torch
        lay = self.tm_sector(As, Bs, direction, K)
        n = lay.nelem
        if n == 0:
            raise ValueError("no transfer-matrix sectors for this charge")
        T = (lambda v: self._chain_left(As, Bs, v)) if direction == LEFT else (lambda v: self._chain_right(As, Bs, v))
        op = lambda z: self._c(self.onto(T(self._p(z, lay)), lay))  # noqa: E731
        m = min(krylov, n)
        gen = t.Generator(device="cpu").manual_seed(seed)
        v = t.complex(t.randn(n, generator=gen, dtype=t.float64), t.randn(n, generator=gen, dtype=t.float64)).to(self.dev)
        nmv = 0
        for _ in range(maxit):
            V = [v / t.linalg.vector_norm(v)]
            Hm = np.zeros((m + 1, m), complex)
            for j in range(m):
                w = op(V[j])
                nmv += 1
                for _ in range(2):
                    for i in range(j + 1):
                        h = t.vdot(V[i], w).item()
                        Hm[i, j] += h
                        w = w - h * V[i]
                Hm[j + 1, j] = t.linalg.vector_norm(w).item()
                if j + 1 < m:
                    if abs(Hm[j + 1, j]) < 1e-14:
                        m = j + 1
                        Hm = Hm[:m + 1, :m]
                        break
                    V.append(w / Hm[j + 1, j])
            theta, S = np.linalg.eig(Hm[:m, :m])
            order = np.argsort(-np.abs(theta))
            theta, S = theta[order], S[:, order]
            k = min(nev, m)
            res = np.abs(Hm[m, m - 1]) * np.abs(S[m - 1, :k])
            Vm = t.stack(V[:m])
            Y = t.from_numpy(S[:, :k].T.copy()).to(self.dev) @ Vm
            if np.all(res <= tol * np.abs(theta[:k])) or m >= n:
                return theta[:k], [self._p(Y[i], lay) for i in range(k)], nmv
            v = Y.sum(dim=0)
        raise ImpsError(2, "transfer-matrix eigensolver did not converge")


class ExcitationEngine(BlockOps):
    """Effective Hamiltonian of the momentum-k excitation ansatz for one ground state.

    AL, AR, NL: device FP64 arrays of the unit cell's tensors with their layouts; C: the bond
    matrices (C[n] on bond n + 1) with layouts; H: dense d^2 x d^2 two-site Hamiltonian (E0 is
    subtracted here); K: excitation charge.  Ground-state energy per site E0 is given by the
    caller.  All work is FP64."""

    def __init__(self, ctx: Context, AL, ALl, AR, ARl, NL, NLl, C, Cl, H: np.ndarray, d: int, E0: float, K, inv_tol: float = 1e-14,
                 recycle: bool = False):
        super().__init__(ctx, ALl[0].nsym, d, H, E0, inv_tol)
        # recycle: start each geometric sum from the previous apply's solution, as
        # Excitations_main.cpp:248,274 does through ApplyHeff's pEBR / pEHBL (otherwise from zero).
        # Inside the eigensolver successive inputs are orthogonal Krylov vectors, so the previous
        # solution is no better a start: same step counts and times measured at D = 512 Hubbard
        self.recycle = recycle
        self._ebr = None
        self._ehbl = None
        self.N = len(AL)
        self.K = tuple(int(k) for k in np.atleast_1d(K))
        self.AL = [DObj(l, a) for a, l in zip(AL, ALl)]
        self.AR = [DObj(l, a) for a, l in zip(AR, ARl)]
        self.NL = [DObj(l, a) for a, l in zip(NL, NLl)]
        self.C = [DObj(l, c) for c, l in zip(C, Cl)]
        self.bond0 = Cl[-1]
        # reduced operators of the single tensors (real)
        self.HLr = [self._redop(a, LEFT) for a in self.AL]
        self.HRr = [self._redop(a, RIGHT) for a in self.AR]
        self.HLtot, self.HRtot = self._heff_constants()
        Cb = self.C[-1]
        self.LM = DObj(Cb.lay, self._transpose_blocks(Cb))
        self.RM = Cb
        self.xdims = self._xdims()

    # ---- constants -----------------------------------------------------------------------------
    def _identity(self, lay: BlockLayout) -> DObj:
        t = np.zeros(lay.nelem)
        for (e, q), (qo, r, c, o) in _index(lay).items():
            t[o:o + r * c].reshape(c, r)[np.diag_indices(min(r, c))] = 1.0
        return DObj(lay, self.torch.from_numpy(t).to(self.dev))

    def _heff_constants(self):
        """HLtot, HRtot (HeffConstants, Excitations_helpers.hpp:260-347), real."""
        N, H = self.N, self.H0
        AL, AR = self.AL, self.AR
        opl = lambda AA: self.diag_part(self.tml_unity(self.apply_operator(AA, H), AA))  # noqa: E731
        opr = lambda AA: self.diag_part(self.tmr_unity(self.apply_operator(AA, H), AA))  # noqa: E731
        real = lambda x: DObj(x.lay, x.data[0].contiguous()) if x.cplx else x  # noqa: E731
        HL = [real(opl(self.mps_mps(AL[-1], AL[0])))]
        for n in range(1, N):
            HL.append(real(self.add(opl(self.mps_mps(AL[n - 1], AL[n])), self.diag_part(self.tml(AL[n], AL[n], HL[-1])))))
        HR = [real(opr(self.mps_mps(AR[-1], AR[0])))]
        for n in range(N - 2, -1, -1):
            HR.insert(0, real(self.add(opr(self.mps_mps(AR[n], AR[n + 1])), self.diag_part(self.tmr(AR[n], AR[n], HR[0])))))
        b0 = self.bond0
        Cb = self.C[-1]
        R = self.prod(MPS_MATT, Cb, Cb)  # C C^T
        L = self.prod(MATT_MPS, Cb, Cb)  # C^T C
        one = self._identity(b0)
        x = self.onto(HL[-1], b0)
        pl = EnvPlan(self.ctx, LEFT, [a.lay for a in AL], b0, FP64)
        pl.set_mps([a.data for a in AL])
        yl = self._zeros(b0.nelem, False)
        pl.solve(x.data, one.data, self.onto(R, b0).data, yl, tol=self.tol, maxit=20000, krylov=100, check=False)
        pr = EnvPlan(self.ctx, RIGHT, [a.lay for a in AR], b0, FP64)
        pr.set_mps([a.data for a in AR])
        yr = self._zeros(b0.nelem, False)
        pr.solve(self.onto(HR[0], b0).data, self.onto(L, b0).data, one.data, yr, tol=self.tol, maxit=20000, krylov=100, check=False)
        EHL = [None] * N
        EHR = [None] * N
        EHL[N - 1] = DObj(b0, yl)
        HLtot = [None] * N
        HLtot[N - 1] = EHL[N - 1]
        for n in range(N - 1):
            EHL[n] = real(self.tml(AL[n], AL[n], EHL[(n - 1) % N]))
            HLtot[n] = real(self.add(HL[n], EHL[n]))
        EHR[0] = DObj(b0, yr)
        HRtot = [None] * N
        HRtot[0] = EHR[0]
        for n in range(N - 1, 0, -1):
            EHR[n] = real(self.tmr(AR[n], AR[n], EHR[(n + 1) % N]))
            HRtot[n] = real(self.add(EHR[n], HR[n]))
        return HLtot, HRtot

    def _xdims(self) -> List[BlockLayout]:
        """Sectors of X[n] (XDims, Excitations_helpers.hpp:639-678)."""
        out = []
        for n in range(self.N):
            ml, mr = {}, {}
            for (e, q), (qo, r, c, o) in _index(self.NL[n].lay).items():
                ml[qo] = c
            for (e, q), (qo, r, c, o) in _index(self.AR[(n + 1) % self.N].lay).items():
                mr[q] = r
            idx = {}
            for q in sorted(ml):
                qk = tuple(a + b for a, b in zip(q, self.K))
                if qk in mr and ml[q] > 0 and mr[qk] > 0:
                    idx[(0, q)] = (qk, ml[q], mr[qk], 0)
            out.append(_layout_from_index(self.nsym, 1, idx))
        return out

    @property
    def mtot(self) -> int:
        return sum(l.nelem for l in self.xdims)

    # ---- ApplyHeff ---------------------------------------------------------------------------
    def _left_sum(self, Bin: List[DObj], ck: complex, proj: bool) -> List[DObj]:
        """The left geometric sum of ApplyHeff: HBL (the Hamiltonian terms left of B) summed
        through (1 - e^{-ik} T_L)^-1 (Excitations_helpers.hpp:108-142)."""
        N, AL, AR, H = self.N, self.AL, self.AR, self.H0
        HBL = [self.add(self.tml_unity(self.mat_mps(self.HLtot[-1], Bin[0]), AL[0]),
                        self.tml_unity(self.apply_operator(self.add(self.mps_mps(AL[-1], Bin[0]), self.mps_mps(Bin[-1], AR[0]), 1.0, ck), H),
                                       self.mps_mps(AL[-1], AL[0])))]
        for n in range(1, N):
            HBL.append(self.add(self.tml_unity(self.add(self.mat_mps(HBL[-1], AR[n]), self.mat_mps(self.HLtot[n - 1], Bin[n])), AL[n]),
                                self.tml_unity(self.apply_operator(self.add(self.mps_mps(AL[n - 1], Bin[n]), self.mps_mps(Bin[n - 1], AR[n])), H),
                                               self.mps_mps(AL[n - 1], AL[n]))))
        y0 = self._ehbl if self.recycle and self._ehbl is not None and _lkey(self._ehbl.lay) == _lkey(HBL[-1].lay) else None
        if proj:  # LM^T = C, RM^T = C^T (bond 0, block diagonal)
            EHBL, _ = self.invert_e(AR, AL, HBL[-1], LEFT, ck, self.RM, self.LM, y0=y0)
        else:
            EHBL, _ = self.invert_e(AR, AL, HBL[-1], LEFT, ck, y0=y0)
        self._ehbl = EHBL
        HBLtot = [self.scale(EHBL, ck)]
        for n in range(N - 1):
            EHBL = self.tml(AR[n], AL[n], EHBL)
            HBLtot.append(self.add(HBL[n], EHBL, 1.0, ck))
        HBLtot = HBLtot[1:] + HBLtot[:1]
        return HBLtot

    def apply(self, X: List[DObj], p: float) -> List[DObj]:
        """H_eff(k) X at momentum p (unit-cell phase p pi N, Excitations_main.cpp:258-259)."""
        N, d = self.N, self.d
        pact = p * math.pi * N
        kfac = complex(math.cos(pact), math.sin(pact))
        ck = kfac.conjugate()
        proj = abs(pact) < 10 * self.tol and all(k == 0 for k in self.K)
        AL, AR, NL = self.AL, self.AR, self.NL
        P = lambda x: x % N  # noqa: E731
        Bin = [self.mps_mat(NL[n], X[n]) for n in range(N)]
        ABR = [self.tmr_unity(Bin[-1], AR[-1])]
        for n in range(N - 1, 0, -1):
            ABR.insert(0, self.tmr_unity(self.add(Bin[n - 1], self.mps_mat(AL[n - 1], ABR[0])), AR[n - 1]))
        y0 = self._ebr if self.recycle and self._ebr is not None and _lkey(self._ebr.lay) == _lkey(ABR[0].lay) else None
        EBR, _ = self.invert_e(AL, AR, ABR[0], RIGHT, kfac, self.LM if proj else None, self.RM if proj else None, y0=y0)
        self._ebr = EBR
        ABRtot = [self.scale(EBR, kfac)]
        for n in range(N - 1, 0, -1):
            EBR = self.tmr(AL[n], AR[n], EBR)
            ABRtot.insert(0, self.add(ABR[n], EBR, 1.0, kfac))
        ABRtot = ABRtot[-1:] + ABRtot[:-1]
        HBLtot = self._left_sum(Bin, ck, proj)
        Bout = []
        for n in range(N):
            b = self.add(self.mat_mps(self.HLtot[P(n - 1)], Bin[n]), self.mps_mat(Bin[n], self.HRtot[P(n + 1)]))
            b = self.add(b, self.prod(REDOP_MPS, self.HLr[P(n - 1)], Bin[n]))
            b = self.add(b, self.prod(MPS_REDOP, self.HRr[P(n + 1)], Bin[n]))
            Bout.append(b)
        # Bin left of Bout
        Bout[0] = self.add(Bout[0], self.mat_mps(HBLtot[-1], AR[0]))
        Bout[0] = self.add(Bout[0], self.prod(REDOP_MPS, self.redop_left_mixed(AL[-1], Bin[-1]), AR[0]), 1.0, ck)
        for n in range(1, N):
            Bout[n] = self.add(Bout[n], self.mat_mps(HBLtot[n - 1], AR[n]))
            Bout[n] = self.add(Bout[n], self.prod(REDOP_MPS, self.redop_left_mixed(AL[n - 1], Bin[n - 1]), AR[n]))
        # Bin right of Bout
        for n in range(N):
            Atmp = self.add(Bin[P(n + 1)], self.mps_mat(AL[P(n + 1)], ABRtot[P(n + 2)]))
            if n == N - 1:
                Atmp = self.scale(Atmp, kfac)
            Y = self.mps_mat(AL[n], ABRtot[P(n + 1)])
            Bout[n] = self.add(Bout[n], self.mat_mps(self.HLtot[P(n - 1)], Y))
            Bout[n] = self.add(Bout[n], self.prod(REDOP_MPS, self.HLr[P(n - 1)], Y))
            Bout[n] = self.add(Bout[n], self.prod(MPS_REDOP, self.redop_right_mixed(Atmp, AR[P(n + 1)]), AL[n]))
        return [self.onto(self.tml_unity(Bout[n], NL[n]), X[n].lay) for n in range(N)]

    # ---- vector interface (eigensolvers) -------------------------------------------------------
    def split(self, z) -> List[DObj]:
        """Complex device vector (mtot,) in XDims order -> X[n] objects."""
        out, pos = [], 0
        for lay in self.xdims:
            n = lay.nelem
            out.append(self._p(z[pos:pos + n], lay))
            pos += n
        return out

    def join(self, X: List[DObj]):
        return self.torch.cat([self._c(x) for x in X])

    def matvec(self, z, p: float):
        return self.join(self.apply(self.split(z), p))

    def lowest(self, p: float, nev: int = 4, tol: float = 1e-10, krylov: int = 60, maxit: int = 50, seed: int = 0):
        """Lowest nev eigenpairs of H_eff(k) (the reference's eigs_n "SR" call, Excitations_main.cpp:279).
        H_eff(k) is Hermitian, so the eigensolver is a restarted Rayleigh-Ritz (Lanczos) iteration with
        full re-orthogonalisation on device vectors: the basis is extended by H v to `krylov` vectors,
        the projected matrix is diagonalised, and the iteration restarts from the lowest nev + 2 Ritz
        vectors plus the residual of the first unconverged one.  Converged when every wanted Ritz pair
        has ||H y - theta y|| <= tol * max(1, |theta|).  Returns (energies, vectors (nev, mtot), matvecs)."""
        t = self.torch
        n = self.mtot
        m = min(krylov, n)
        keep = min(nev + 2, m - 1) if m > 1 else 1
        gen = t.Generator(device="cpu").manual_seed(seed)
        v0 = t.complex(t.randn(n, generator=gen, dtype=t.float64), t.randn(n, generator=gen, dtype=t.float64)).to(self.dev)
        V = [v0 / t.linalg.vector_norm(v0)]
        W = []
        nmv = 0
        theta = None
        for _ in range(maxit):
            while len(V) > len(W):
                W.append(self.matvec(V[len(W)], p))
                nmv += 1
                if len(V) < m:
                    w = W[-1]
                    for _ in range(2):
                        Vm = t.stack(V)
                        w = w - Vm.t() @ (Vm.conj() @ w)
                    nw = t.linalg.vector_norm(w).item()
                    if nw > 1e-12:
                        V.append(w / nw)
            Vm, Wm = t.stack(V), t.stack(W)
            T = (Vm.conj() @ Wm.t()).cpu().numpy()
            T = 0.5 * (T + T.conj().T)
            theta, S = np.linalg.eigh(T)
            St = t.from_numpy(S).to(self.dev)
            Y = St.t() @ Vm
            HY = St.t() @ Wm
            k = min(nev, len(V))
            res = [(t.linalg.vector_norm(HY[i] - float(theta[i]) * Y[i]).item()) for i in range(k)]
            ok = [res[i] <= tol * max(1.0, abs(theta[i])) for i in range(k)]
            if all(ok) or len(V) >= n:
                return theta[:k], Y[:k], nmv
            first = ok.index(False)
            r = HY[first] - float(theta[first]) * Y[first]
            kk = min(keep, len(V))
            V = [Y[i] for i in range(kk)]
            W = [HY[i] for i in range(kk)]
            for _ in range(2):
                Vm = t.stack(V)
                r = r - Vm.t() @ (Vm.conj() @ r)
            V.append(r / t.linalg.vector_norm(r))
        raise ImpsError(2, "excitation eigensolver did not converge")

