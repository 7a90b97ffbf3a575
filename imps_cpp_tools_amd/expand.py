"""Bond-dimension growth of an MPS bond from the two-site Hamiltonian (ExpandFromH,
MPSBlockUtilities.hpp:407-608) and Schmidt-value bookkeeping for the expansion policy of the VUMPS
outer loop (VUMPS_main.cpp:273-306).

The bond b between AL (left site, right bond b) and AR (right site, left bond b) carries C.  With
phi = H (AL C AR) the two-site residual direction, every charge Q of the bond is a "center" sector:
its dense split matrix Phi_Q has rows (s_l, q_l) with q_l + dq(s_l) = Q and columns (s_r, q_r) with
q_r = Q + dq(s_r), blocks in ascending physical index (GetDenseSplitMat, MPSBlockMat.hpp:531-571).
For a charge already on the bond, Phi_Q is projected on the null spaces of the regrouped AL and AR
(M_Q = NL^T Phi_Q NR^T, skipped when either side is already full rank); a new charge takes Phi_Q as
is.  The singular values of all sectors compete for at most dm_max new states; each sector grows
by the count it wins.  AL gains the columns NL U (existing) or U (new), AR the rows V^T NR or V^T,
and C is padded with zeros, so the state AL C AR is unchanged.  Ties between equal singular values
are broken by the sort order (the reference's std::sort leaves that order open).

This runs on the host (NumPy SVDs of sector matrices of at most a few hundred rows), as the
reference's LAPACK calls do; it is called every few outer iterations, not per Krylov step.
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np

from . import BlockLayout

Key = Tuple[int, ...]


def mps_blocks(lay: BlockLayout, data: np.ndarray) -> List[Dict[Key, Tuple[Key, np.ndarray]]]:
    """Per physical index s: {q_in: (q_out, block)} of an MPS tensor (column-major blocks)."""
    out: List[Dict[Key, Tuple[Key, np.ndarray]]] = []
    b = 0
    offs = lay.offsets
    for e in range(lay.nlist):
        ent = {}
        for j in range(b, b + int(lay.nblk[e])):
            r, c = int(lay.rows[j]), int(lay.cols[j])
            ent[tuple(int(v) for v in lay.qin[j])] = (tuple(int(v) for v in lay.qout[j]),
                                                      data[offs[j]:offs[j] + r * c].reshape(r, c, order="F"))
        out.append(ent)
        b += int(lay.nblk[e])
    return out


def diag_blocks(lay: BlockLayout, data: np.ndarray) -> Dict[Key, np.ndarray]:
    """{q: block} of a block-diagonal bond matrix."""
    return {tuple(int(v) for v in q): data[o:o + int(r) * int(c)].reshape(int(r), int(c), order="F")
            for q, o, r, c in zip(lay.qin, lay.offsets, lay.rows, lay.cols)}


def flatten_mps(lay: BlockLayout, blocks) -> np.ndarray:
    """Inverse of mps_blocks on `lay`; blocks missing from `blocks` are zero, present ones are
    copied into the top-left corner of the layout's block."""
    out = np.zeros(lay.nelem)
    b = 0
    offs = lay.offsets
    for e in range(lay.nlist):
        for j in range(b, b + int(lay.nblk[e])):
            src = blocks[e].get(tuple(int(v) for v in lay.qin[j]))
            if src is None:
                continue
            r, c = int(lay.rows[j]), int(lay.cols[j])
            m = np.zeros((r, c))
            m[:src[1].shape[0], :src[1].shape[1]] = src[1]
            out[offs[j]:offs[j] + r * c] = m.reshape(-1, order="F")
        b += int(lay.nblk[e])
    return out


def flatten_diag(lay: BlockLayout, blocks: Dict[Key, np.ndarray]) -> np.ndarray:
    """Block-diagonal bond matrix on `lay`, old blocks in the top-left corner, zeros elsewhere."""
    out = np.zeros(lay.nelem)
    for q, o, r in zip(lay.qin, lay.offsets, lay.rows):
        src = blocks.get(tuple(int(v) for v in q))
        if src is None:
            continue
        r = int(r)
        m = np.zeros((r, r))
        m[:src.shape[0], :src.shape[1]] = src
        out[o:o + r * r] = m.reshape(-1, order="F")
    return out


def schmidt_values(lay: BlockLayout, data: np.ndarray) -> Dict[Key, np.ndarray]:
    """Singular values of each sector block of the bond matrix C (svd(C), VUMPS_main.cpp:207-208)."""
    return {q: np.linalg.svd(m, compute_uv=False) for q, m in diag_blocks(lay, data).items()}


def _add(a: Key, b, periods=None) -> Key:
    """Charge sum, reduced per Z(N) charge (period N > 1) to [0, N)."""
    if periods is None:
        return tuple(int(x) + int(y) for x, y in zip(a, b))
    return tuple((int(x) + int(y)) % int(p) if p > 1 else int(x) + int(y) for x, y, p in zip(a, b, periods))


def _null_rows(A: np.ndarray) -> np.ndarray:
    """Orthonormal basis (columns) of the orthogonal complement of the columns of the isometry A."""
    R, m = A.shape
    if R <= m:
        return np.zeros((R, 0))
    U, _, _ = np.linalg.svd(A, full_matrices=True)
    return U[:, m:]


def expand_from_h(H: np.ndarray, dq: np.ndarray, AL, C: Dict[Key, np.ndarray], AR, dm_max: int, periods=None):
    """One ExpandFromH step on the bond of C (host data from mps_blocks / diag_blocks).

    H: two-site Hamiltonian (d^2 x d^2, pair index s = s_l d + s_r); dq[s]: charge shift of
    physical index s; AL, AR: per-s block dicts of the left-orthonormal left tensor and the
    right-orthonormal right tensor; C: {Q: block} of the bond.  Returns (dm, new_cols, new_rows, lam):
    dm = {Q: added states} (in ascending charge order), new_cols[s] = {q_l: columns appended to
    AL's block (q_l -> Q)}, new_rows[s] = {q_r_in: rows appended to AR's block (Q -> q_r)} and lam =
    {Q: all singular values of the sector's matrix}.  periods: the symmetry group (None: U(1))."""
    d = len(AL)
    dq = [tuple(int(v) for v in row) for row in np.asarray(dq).reshape(d, -1)]
    # left and right bond sector sizes
    ml, mr = {}, {}
    for s in range(d):
        for ql, (_, blk) in AL[s].items():
            ml[ql] = blk.shape[0]
        for Q, (qr, blk) in AR[s].items():
            mr[qr] = blk.shape[1]
    # two-site state psi[(s_l, s_r)][q_l] = (q_r, AL C AR) and phi = H psi (ApplyOperator)
    psi: Dict[Tuple[int, int], Dict[Key, Tuple[Key, np.ndarray]]] = {}
    for sl in range(d):
        for ql, (Q, a) in AL[sl].items():
            if Q not in C:
                continue
            aC = a @ C[Q]
            for sr in range(d):
                e = AR[sr].get(Q)
                if e is not None:
                    psi.setdefault((sl, sr), {})[ql] = (e[0], aC @ e[1])
    phi: Dict[Tuple[int, int], Dict[Key, Tuple[Key, np.ndarray]]] = {}
    for sl in range(d):
        for sr in range(d):
            s = sl * d + sr
            for tl in range(d):
                for tr in range(d):
                    h = H[s, tl * d + tr]
                    if h == 0.0 or (tl, tr) not in psi:
                        continue
                    for ql, (qr, m) in psi[(tl, tr)].items():
                        # H conserves charge: the pair (sl, sr) maps ql to the same qr
                        if _add(_add(ql, dq[sl], periods), dq[sr], periods) != qr or qr not in mr:
                            continue
                        ent = phi.setdefault((sl, sr), {})
                        if ql in ent:
                            ent[ql] = (qr, ent[ql][1] + h * m)
                        else:
                            ent[ql] = (qr, h * m)
    # center sectors of phi, rows (s_l, q_l) and columns (s_r, q_r) in ascending physical index
    centers: Dict[Key, Dict[str, Dict[int, Key]]] = {}
    for (sl, sr), ent in phi.items():
        for ql, (qr, _) in ent.items():
            Q = _add(ql, dq[sl], periods)
            c = centers.setdefault(Q, {"l": {}, "r": {}})
            c["l"][sl] = ql
            c["r"][sr] = qr
    cand = []  # (Q, isnew, U, lam, V, row spans, col spans)
    for Q in sorted(centers):
        c = centers[Q]
        isnew = Q not in C
        if isnew:
            rows = sorted(c["l"].items())
            cols = sorted(c["r"].items())
        else:
            # the regrouped AL (rows (s, q_l) of every block q_l -> Q) and AR (columns (s, q_r))
            rows = sorted((s, ql) for s in range(d) for ql, (q2, _) in AL[s].items() if q2 == Q)
            cols = sorted((s, AR[s][Q][0]) for s in range(d) if Q in AR[s])
        rsp, r0 = {}, 0
        for s, ql in rows:
            rsp[s] = (ql, r0, r0 + ml[ql])
            r0 += ml[ql]
        csp, c0 = {}, 0
        for s, qr in cols:
            csp[s] = (qr, c0, c0 + mr[qr])
            c0 += mr[qr]
        Phi = np.zeros((r0, c0))
        for (sl, sr), ent in phi.items():
            if sl not in rsp or sr not in csp:
                continue
            ql, a0, a1 = rsp[sl]
            e = ent.get(ql)
            if e is None:
                continue
            qr, b0, b1 = csp[sr]
            if e[0] == qr:
                Phi[a0:a1, b0:b1] += e[1]
        if isnew:
            Mq, NL, NR = Phi, None, None
        else:
            A = np.zeros((r0, C[Q].shape[0]))
            for s, (ql, a0, a1) in rsp.items():
                A[a0:a1, :] = AL[s][ql][1]
            B = np.zeros((C[Q].shape[1], c0))
            for s, (qr, b0, b1) in csp.items():
                B[:, b0:b1] = AR[s][Q][1]
            NL = _null_rows(A)
            NR = _null_rows(B.T).T
            if NL.shape[1] == 0 or NR.shape[0] == 0:
                continue  # already full rank on one side: nothing to add (Nullspace(.., true) marks)
            Mq = NL.T @ Phi @ NR.T
        if Mq.size == 0:
            continue
        U, lam, Vt = np.linalg.svd(Mq, full_matrices=False)
        cand.append((Q, isnew, U, lam, Vt.T, rsp, csp, NL, NR))
    # at most dm_max states in total, the largest singular values first
    counts = [len(x[3]) for x in cand]
    if sum(counts) > dm_max:
        allv = [(lam_v, i) for i, x in enumerate(cand) for lam_v in x[3]]
        order = sorted(range(len(allv)), key=lambda k: -allv[k][0])[:dm_max]
        counts = [0] * len(cand)
        for k in order:
            counts[allv[k][1]] += 1
    dm: Dict[Key, int] = {}
    new_cols: List[Dict[Key, np.ndarray]] = [dict() for _ in range(d)]
    new_rows: List[Dict[Key, np.ndarray]] = [dict() for _ in range(d)]
    lam_all: Dict[Key, np.ndarray] = {}
    for (Q, isnew, U, lam, V, rsp, csp, NL, NR), k in zip(cand, counts):
        lam_all[Q] = lam
        if k == 0:
            continue
        dm[Q] = k
        Lc = U[:, :k] if isnew else NL @ U[:, :k]
        Rr = V[:, :k].T if isnew else V[:, :k].T @ NR
        for s, (ql, a0, a1) in rsp.items():
            new_cols[s][ql] = Lc[a0:a1, :]
        for s, (qr, b0, b1) in csp.items():
            new_rows[s][Q] = Rr[:, b0:b1]
    return dm, new_cols, new_rows, lam_all


def grow_bond(keys: np.ndarray, sizes: np.ndarray, dm: Dict[Key, int]):
    """Bond (keys, sizes) with the added states, new charges inserted in ascending key order."""
    cur = {tuple(int(v) for v in k): int(s) for k, s in zip(np.asarray(keys).reshape(len(sizes), -1), sizes)}
    for q, k in dm.items():
        cur[q] = cur.get(q, 0) + int(k)
    qs = sorted(cur)
    return np.asarray(qs, np.int32).reshape(len(qs), -1), np.asarray([cur[q] for q in qs], np.int32)


def append_columns(AL, new_cols):
    """AL blocks (q_l -> Q) with the new columns appended; a block for a new charge Q is created
    from its columns alone (rows of q_l)."""
    out = [dict(a) for a in AL]
    for s, ent in enumerate(new_cols):
        for ql, cols in ent.items():
            if ql in out[s]:
                Q, blk = out[s][ql]
                out[s][ql] = (Q, np.hstack([blk, cols]))
            else:
                out[s][ql] = (None, cols)  # new charge: the block is the new columns
    return out


def append_rows(AR, new_rows):
    """AR blocks (Q -> q_r) with the new rows appended (new charges: the rows alone)."""
    out = [dict(a) for a in AR]
    for s, ent in enumerate(new_rows):
        for Q, rows in ent.items():
            if Q in out[s]:
                qr, blk = out[s][Q]
                out[s][Q] = (qr, np.vstack([blk, rows]))
            else:
                out[s][Q] = (None, rows)
    return out


# ---- Truncate (MPSBlockUtilities.hpp:846-936) --------------------------------------------------
def truncate_sizes(lam: Dict[Key, np.ndarray], mmax: int, lamthresh: float = 1e-10) -> Tuple[Dict[Key, int], bool]:
    """Truncate's selection (MPSBlockUtilities.hpp:846-906): of the Schmidt values on a bond
    ({Q: values in descending order}), those above lamthresh survive, and when more than mmax do,
    the mmax largest of them.  Returns ({Q: kept count}, changed).  A sector keeps its leading
    values, so the count is all that is needed (0: the sector goes).  Ties at the mmax cut are kept
    in ascending charge order, then by position (the reference's std::sort leaves that order open)."""
    vals = []
    small = False
    for q in sorted(lam):
        for i, v in enumerate(np.asarray(lam[q], float)):
            if v > lamthresh:
                vals.append((-float(v), q, i))
            else:
                small = True
    big = len(vals) > mmax
    keep = {q: 0 for q in lam}
    if big:
        vals.sort()
        vals = vals[:mmax]
    for _, q, _ in vals:
        keep[q] += 1
    return keep, bool(big or small)


def schmidt_basis(C: Dict[Key, np.ndarray]):
    """Per sector C_Q = U_Q diag(S_Q) V_Q^T (singular values descending): ({Q: U}, {Q: S}, {Q: V})."""
    U, S, V = {}, {}, {}
    for q, m in C.items():
        u, s, vt = np.linalg.svd(m)
        U[q], S[q], V[q] = u, s, vt.T
    return U, S, V


def rotate_right_index(A, R: Dict[Key, np.ndarray]):
    """A's blocks (q_in -> Q) times R_Q on the right (A from mps_blocks; Q the block's out charge)."""
    return [{qi: (Q, blk @ R[Q]) for qi, (Q, blk) in ent.items()} for ent in A]


def rotate_left_index(A, R: Dict[Key, np.ndarray]):
    """R_Q^T times A's blocks (Q -> q_out) on the left."""
    return [{Q: (qo, R[Q].T @ blk) for Q, (qo, blk) in ent.items()} for ent in A]


def cut_right_index(A, keep: Dict[Key, int]):
    """A's blocks (q_in -> Q) with their first keep[Q] columns; blocks into a dropped sector go."""
    return [{qi: (Q, blk[:, :keep[Q]]) for qi, (Q, blk) in ent.items() if keep.get(Q, 0) > 0} for ent in A]


def cut_left_index(A, keep: Dict[Key, int]):
    """A's blocks (Q -> q_out) with their first keep[Q] rows; blocks out of a dropped sector go."""
    return [{Q: (qo, blk[:keep[Q], :]) for Q, (qo, blk) in ent.items() if keep.get(Q, 0) > 0} for ent in A]
