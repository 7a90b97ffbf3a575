"""Symmetry transformations of the right-canonical state and the mixed transfer matrix of (AL, AR)
for the excitation program.

apply_transformation restates ApplyTransformation (Excitations_helpers.hpp:680-771): a spin flip
("FS") of the XXZ chain, or a spin flip ("FS") or charge flip ("FC", particle-hole) of the
Fermi-Hubbard chain in its (N, M) representation, applied to AR and to L (the left fixed point of
AR's transfer matrix), with the fermionic signs of the Hubbard flips.  Each physical index s of an
MPS tensor moves to the index whose charge is the flipped charge of s (MPSBlockMat::FlipQN,
MPSBlockMat.hpp:762-787, with the flip of KeyTypes.hpp:156-170), every block keeps its matrix
under flipped in- and outgoing charges (BlockMat::FlipQN, BlockObj.hpp:845-855; BlockDiagMat::FlipQN,
BlockObj.hpp:1740-1749), and the physical indices the reference marks get a sign.  The work is a
gather: the index and sign arrays are made on the host from the layouts, the data move on the
device.

lr_overlaps restates LRoverlaps (Excitations_helpers.hpp:915-946): the dominant left and right
eigenmatrices of the mixed transfer matrix of (AL, AR) in the sectors -K and K, found by the
library's TM Arnoldi (BlockOps.tm_eigs over the grouped block products), and the mean of the two
dominant eigenvalues, whose sign Excitations_main.cpp:177 multiplies into AR[0].  The complex
eigenvectors are made real by the phase of their largest element (the reference keeps the real
part, BlockObj.hpp:612-627) and scaled to tr(LM RM) = 1, the pairing InvertE_proj_fac's projector
assumes (InvEBlockFunctions.hpp:50-118).
"""
from __future__ import annotations

import warnings
from typing import List, Sequence

import numpy as np

from . import LEFT, RIGHT, BlockLayout, Context, ImpsError, i2k

XXZ, FHUB = 0, 1


def _flip_mask(model: int, op: str, N: int):
    """(charge components flipped, per-site sign masks over the target physical index) of
    ApplyTransformation for the (N, M) representation of the Hubbard chain."""
    mt = model & 0xff
    if mt == XXZ:
        if op != "FS":
            raise ValueError(f"ApplyTransformation: transformation {op!r} for XXZ not recognized")
        return [True], [[False, False]] * N
    if mt == FHUB:
        if op == "FS":
            return [False, True], [[False, False, False, True]] * N
        if op == "FC":
            if N % 2 != 0:
                raise ValueError("ApplyTransformation: unit cell size needs to be even for charge flip")
            return [True, False], [[False, False, True, False] if n % 2 == 0 else [True, False, True, True] for n in range(N)]
        raise ValueError(f"ApplyTransformation: transformation {op!r} for FHUB not recognized")
    raise ValueError("ApplyTransformation: transformations for this model not implemented")


def _flip(q: np.ndarray, which) -> np.ndarray:
    out = np.array(q, np.int64, copy=True)
    for i, w in enumerate(which):
        if w:
            out[..., i] = -out[..., i]
    return out


def _entries_map(model: int, N: int, QN: Sequence[int], which, d: int) -> List[int]:
    """Target physical index of each index s: the position of the flipped charge of I2K[s] in I2K
    (the one-site ItoKey of the model)."""
    keys = i2k(model, N, QN).astype(np.int64)
    if keys.shape[0] != d:
        raise ValueError("ApplyTransformation: the MPS layout's physical dimension does not match the model")
    pos = {tuple(int(v) for v in k): s for s, k in enumerate(keys)}
    out = []
    for s in range(d):
        f = tuple(int(v) for v in _flip(keys[s], which))
        if f not in pos:
            raise ValueError(f"MPSBlockMat::FlipQN: key {f} not found in target I2K")
        out.append(pos[f])
    return out


def _gather_plan(lay: BlockLayout, which, dest_of_entry: List[int], entry_sign: List[float]):
    """Flipped layout with its blocks re-sorted by the flipped ingoing charge per (target) entry,
    and the gather index / sign arrays over the elements."""
    nlist = lay.nlist
    starts = np.concatenate([[0], np.cumsum(lay.nblk)]).astype(np.int64)
    off, sz = lay.offsets, lay.sizes
    per_entry = [[] for _ in range(nlist)]
    for e in range(nlist):
        t = dest_of_entry[e]
        for b in range(int(starts[e]), int(starts[e + 1])):
            qi = tuple(int(v) for v in _flip(lay.qin[b], which))
            qo = tuple(int(v) for v in _flip(lay.qout[b], which))
            per_entry[t].append((qi, qo, int(lay.rows[b]), int(lay.cols[b]), int(off[b]), int(sz[b]), entry_sign[t]))
    nblk, qin, qout, rows, cols, idx, sgn = [], [], [], [], [], [], []
    for t in range(nlist):
        blocks = sorted(per_entry[t], key=lambda x: x[0])
        nblk.append(len(blocks))
        for qi, qo, r, c, o, n, sg in blocks:
            qin.append(qi)
            qout.append(qo)
            rows.append(r)
            cols.append(c)
            idx.append(np.arange(o, o + n, dtype=np.int64))
            sgn.append(np.full(n, sg))
    nsym = lay.nsym
    out = BlockLayout(nsym, np.array(nblk, np.int32), np.array(qin, np.int32).reshape(-1, nsym), np.array(qout, np.int32).reshape(-1, nsym),
                      np.array(rows, np.int32), np.array(cols, np.int32), lay.periods)
    index = np.concatenate(idx) if idx else np.zeros(0, np.int64)
    sign = np.concatenate(sgn) if sgn else np.zeros(0)
    return out, index, sign


def mps_transform_plan(model: int, N: int, QN: Sequence[int], op: str, lay: BlockLayout, site: int):
    """(layout, gather index, signs) of ApplyTransformation's MPSFun on the tensor of unit-cell site
    `site`: FlipQN (with the model's component mask) followed by the fermionic signs of the target
    physical indices (Excitations_helpers.hpp:760-767)."""
    which, signs = _flip_mask(model, op, N)
    if lay.nsym != len(which):
        raise ValueError("ApplyTransformation: the layout's charges do not match the model")
    dest = _entries_map(model, N, QN, which, lay.nlist)
    sg = [-1.0 if signs[site][t] else 1.0 for t in range(lay.nlist)]
    return _gather_plan(lay, which, dest, sg)


def bond_transform_plan(model: int, N: int, op: str, lay: BlockLayout):
    """(layout, gather index, signs) of ApplyTransformation's CFun on a bond object (nlist = 1)."""
    which, _ = _flip_mask(model, op, N)
    if lay.nlist != 1 or lay.nsym != len(which):
        raise ValueError("ApplyTransformation: a bond object has nlist = 1 and the model's charges")
    return _gather_plan(lay, which, [0], [1.0])


def _apply_gather(t, index: np.ndarray, sign: np.ndarray):
    import torch
    if index.size == 0:
        return torch.zeros(0, dtype=t.dtype, device=t.device)
    ix = torch.from_numpy(index).to(t.device)
    sg = torch.from_numpy(sign).to(device=t.device, dtype=t.dtype)
    return torch.index_select(t, 0, ix) * sg


def apply_transformation(model: int, N: int, QN: Sequence[int], op: str, AR: Sequence, ARl: Sequence[BlockLayout], L, Ll: BlockLayout):
    """ApplyTransformation(AR, L, op, I2K, model) (Excitations_helpers.hpp:680-771): returns
    (AR', AR' layouts, L', L' layout) with the data gathered on the tensors' device."""
    if len(AR) != N or len(ARl) != N:
        raise ValueError("ApplyTransformation: one tensor per site")
    out, outl = [], []
    for n in range(N):
        lay, idx, sg = mps_transform_plan(model, N, QN, op, ARl[n], n)
        if int(ARl[n].nelem) != idx.size:
            raise ImpsError(1, "ApplyTransformation: gather size mismatch")
        out.append(_apply_gather(AR[n], idx, sg))
        outl.append(lay)
    lay, idx, sg = bond_transform_plan(model, N, op, Ll)
    return out, outl, _apply_gather(L, idx, sg), lay


def lr_overlaps(ctx: Context, AL: Sequence, ALl: Sequence[BlockLayout], AR: Sequence, ARl: Sequence[BlockLayout], K, nev: int = 1,
                tol: float = 1e-14, krylov: int = 40, maxit: int = 200):
    """LRoverlaps (Excitations_helpers.hpp:915-946): ((LM layout, LM), (RM layout, RM), OL) with LM
    the dominant left eigenmatrix of the mixed TM of (AL, AR) in sector -K (rows on AR's bond 0,
    columns on AL's), RM the right one in K (rows on AL's bond, columns on AR's), real FP64 device
    tensors with tr(LM RM) = 1 (when K = 0), and OL the mean of the two dominant eigenvalues' real
    parts.  Warns, as the reference does, when the eigenvalues are complex or differ beyond 10 tol."""
    import torch
    from .excitations import BlockOps, DObj
    nsym = ALl[0].nsym
    ops = BlockOps(ctx, nsym)
    As = [DObj(l, t) for l, t in zip(ALl, AL)]
    Bs = [DObj(l, t) for l, t in zip(ARl, AR)]
    Kq = np.atleast_1d(np.asarray(K, np.int64))
    res = []
    for direction, KK in ((LEFT, -Kq), (RIGHT, Kq)):
        vals, vecs, _ = ops.tm_eigs(As, Bs, direction, KK, nev=max(1, nev), tol=tol, krylov=krylov, maxit=maxit)
        v = vecs[0]
        z = torch.complex(v.data[0], v.data[1])
        k = int(torch.argmax(z.abs()))
        ph = z[k] / z[k].abs()
        zr = (z / ph).real.contiguous()
        zr = zr / torch.linalg.vector_norm(zr)
        res.append((complex(vals[0]), v.lay, zr))
    (wl, lml, lm), (wr, rml, rm) = res
    lim = 10 * max(tol, 1e-12)
    if abs(wl.imag) > lim * max(1.0, abs(wl)) or abs(wr.imag) > lim * max(1.0, abs(wr)):
        warnings.warn(f"LRoverlaps: dominant overlap between AL and AR is complex ({wl}, {wr})")
    if abs(wl.real - wr.real) > lim * max(1.0, abs(wl)):
        warnings.warn(f"LRoverlaps: dominant left and right overlaps differ by {wl.real - wr.real}")
    if all(int(q) == 0 for q in Kq):
        # tr(LM RM): LM (AR x AL) times RM (AL x AR), summed over the charge-diagonal blocks
        s = torch.zeros((), dtype=torch.float64, device=lm.device)
        ri = {tuple(int(x) for x in rml.qin[b]): b for b in range(len(rml.rows))}
        loff, roff = lml.offsets, rml.offsets
        for b in range(len(lml.rows)):
            q = tuple(int(x) for x in lml.qin[b])
            if q not in ri:
                continue
            c = ri[q]
            r1, c1 = int(lml.rows[b]), int(lml.cols[b])
            A = lm[loff[b]:loff[b] + r1 * c1].view(c1, r1).t()
            r2, c2 = int(rml.rows[c]), int(rml.cols[c])
            B = rm[roff[c]:roff[c] + r2 * c2].view(c2, r2).t()
            s = s + (A * B.t()).sum()
        sv = float(s)
        if sv == 0.0:
            raise ImpsError(1, "LRoverlaps: LM and RM are orthogonal (tr(LM RM) = 0)")
        rm = rm / sv
    return (lml, lm), (rml, rm), 0.5 * (wl.real + wr.real)


def transformed_state(ctx: Context, model: int, N: int, QN: Sequence[int], AL: Sequence, ALl: Sequence[BlockLayout], AR: Sequence,
                      ARl: Sequence[BlockLayout], C: Sequence, Cl: Sequence[BlockLayout], K, op: str = "", rel_shift: int = 0,
                      nev: int = 1, tol: float = 1e-14):
    """The right half of the excitation program's state (Excitations_main.cpp:142-177): L_n = C_n^T C_n
    per bond, ApplyTransformation(AR, L, op) when op is given, the relative shift of AR and L by
    rel_shift sites (MPSBlockUtilities.hpp:1153-1168: site n takes site n - rel_shift), L = the
    last bond's, LRoverlaps in charge K, and AR[0] times the sign of the overlap.  Returns a dict
    with AR, ARl (the new right tensors and layouts), L, LM, RM ((layout, device tensor) pairs for
    HeffPlan.set_state), Lvec (every bond's L, (layout, tensor), as saveALRN stores it) and OL."""
    from . import MATT_MPS, ProdPlan
    import torch
    Lvec = []
    for b in range(N):
        P = ProdPlan(ctx, MATT_MPS, Cl[b], Cl[b])
        Lb = torch.zeros(P.out_layout.nelem, dtype=torch.float64, device=C[b].device)
        P.apply(C[b], C[b], Lb)
        Lvec.append((P.out_layout, Lb))
    ARn, ARln = list(AR), list(ARl)
    if op:
        tL = []
        for n, (lay, t) in enumerate(Lvec):
            if n == 0:
                ARn, ARln, t2, lay2 = apply_transformation(model, N, QN, op, ARn, ARln, t, lay)
            else:
                lay2, idx, sg = bond_transform_plan(model, N, op, lay)
                t2 = _apply_gather(t, idx, sg)
            tL.append((lay2, t2))
        Lvec = tL
    s = rel_shift % N
    if s:
        ARn = [ARn[(n - s) % N] for n in range(N)]
        ARln = [ARln[(n - s) % N] for n in range(N)]
        Lvec = [Lvec[(n - s) % N] for n in range(N)]
    LM, RM, OL = lr_overlaps(ctx, AL, ALl, ARn, ARln, K, nev=nev, tol=tol)
    if OL < 0:
        ARn[0] = -ARn[0]
    return dict(AR=ARn, ARl=ARln, L=Lvec[N - 1], LM=LM, RM=RM, OL=OL, Lvec=Lvec)
