"""Left null spaces of MPS tensors on the GPU (Nullspace(A, l), MPSBlockUtilities.hpp:247-340).

For a left-orthonormal A (sum_s A_s^T A_s = 1), the regrouped matrix of right sector q has R_q
rows (s, q - d_s) and m_q orthonormal columns; its orthogonal complement has dimension R_q - m_q.
NL holds an orthonormal basis of it in A's block structure, with right bond dimension R_q - m_q
(sectors without a complement are left out, as the reference leaves them out).  The basis is the
isometric polar factor of (1 - A A^T)^2 G for a seeded Gaussian G with NL's structure: the same
space as the reference's null() basis, from which it differs by an orthogonal rotation of the
columns; every quantity the excitation operator takes from NL is invariant under that rotation.
The polar factor comes from the Newton-Schulz iteration X <- X (3 - X^T X) / 2 after scaling each
sector to unit Frobenius norm (the complement sectors, up to 240 x 180 at D = 1024 Hubbard, do not
fit the LDS-resident sector QR).  All products run in the library (FP64).
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from . import MPS_MAT, TM_LEFT_UNITY, BlockLayout, Context, ProdPlan


def left_nullspace_layout(lay: BlockLayout) -> BlockLayout:
    """Block structure of the left null space of an MPS tensor with layout `lay`."""
    R, m = {}, {}
    for qo, r, c in zip(map(tuple, lay.qout), lay.rows, lay.cols):
        R[qo] = R.get(qo, 0) + int(r)
        if m.setdefault(qo, int(c)) != int(c):
            raise ValueError("inconsistent right sector size in MPS layout")
    null = {q: R[q] - m[q] for q in R if R[q] > m[q]}
    nblk, qin, qout, rows, cols = [], [], [], [], []
    b = 0
    for e in range(lay.nlist):
        n = 0
        for j in range(b, b + int(lay.nblk[e])):
            qo = tuple(lay.qout[j])
            if qo in null:
                qin.append(lay.qin[j]), qout.append(lay.qout[j]), rows.append(int(lay.rows[j])), cols.append(null[qo])
                n += 1
        nblk.append(n)
        b += int(lay.nblk[e])
    nb = sum(nblk)
    return BlockLayout(lay.nsym, nblk, np.asarray(qin, np.int32).reshape(nb, lay.nsym), np.asarray(qout, np.int32).reshape(nb, lay.nsym),
                       rows, cols, lay.periods)


def left_nullspace(ctx: Context, A, lay: BlockLayout, seed: int = 0) -> Tuple[object, BlockLayout]:
    """(NL, layout): orthonormal left null space of the left-orthonormal device tensor A."""
    import torch
    nl = left_nullspace_layout(lay)
    if nl.nelem == 0:
        return torch.empty(0, dtype=torch.float64, device=A.device), nl
    rng = np.random.default_rng(seed)
    G = torch.from_numpy(rng.standard_normal(nl.nelem)).to(A.device)
    proj = ProdPlan(ctx, TM_LEFT_UNITY, nl, lay)          # P = sum_s A_s^T G_s
    back = ProdPlan(ctx, MPS_MAT, lay, proj.out_layout)  # Y_s = A_s P
    if not back.out_layout == nl:
        raise ValueError("null-space projection does not reproduce the null-space structure")
    P = torch.empty(proj.out_layout.nelem, dtype=torch.float64, device=A.device)
    Y = torch.empty(nl.nelem, dtype=torch.float64, device=A.device)

    def project(X):
        for _ in range(2):  # twice: the second pass removes what rounding left of A's range
            proj.apply(X, A, P)
            back.apply(A, P, Y)
            X -= Y
        return X

    # The polar iteration lifts the small singular values of the projected Gaussian (by up to
    # ~100 at D = 1024), and with them the rounding it adds along A's range; a second projection
    # and a polish (one or two steps from an almost orthonormal start) bring that to rounding level.
    X = _left_polar(ctx, project(G), nl)
    return _left_polar(ctx, project(X), nl), nl


def _left_polar(ctx: Context, X, lay: BlockLayout, tol: float = 1e-14, maxit: int = 100):
    """Isometric polar factor (per right sector, rows regrouped) of the full-column-rank X."""
    import torch
    gram = ProdPlan(ctx, TM_LEFT_UNITY, lay, lay)       # G = sum_s X_s^T X_s, square per sector
    gl = gram.out_layout
    mul = ProdPlan(ctx, MPS_MAT, lay, gl)                # X_s M
    eye = np.zeros(gl.nelem)
    for o, r in zip(gl.offsets, gl.rows):
        eye[o:o + r * r].reshape(r, r)[np.diag_indices(r)] = 1.0
    I = torch.from_numpy(eye).to(X.device)
    G = torch.empty(gl.nelem, dtype=torch.float64, device=X.device)
    Y = torch.empty_like(X)
    gram.apply(X, X, G)
    scale = np.zeros(gl.nelem)
    Gh = G.cpu().numpy()
    for o, r in zip(gl.offsets, gl.rows):
        tr = float(np.trace(Gh[o:o + r * r].reshape(r, r)))
        scale[o:o + r * r].reshape(r, r)[np.diag_indices(r)] = 1.0 / np.sqrt(tr) if tr > 0 else 0.0
    mul.apply(X, torch.from_numpy(scale).to(X.device), Y)
    X = Y.clone()
    for _ in range(maxit):
        gram.apply(X, X, G)
        if float((G - I).abs().max().item()) < tol:
            return X
        mul.apply(X, 1.5 * I - 0.5 * G, Y)
        X, Y = Y, X
    raise RuntimeError("null space: Newton-Schulz polar iteration did not converge")
