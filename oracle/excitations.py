"""Complex FP64 CPU restatement of the quasiparticle-excitation effective Hamiltonian.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU path, never by the product.

Block objects are plain Python: a BlockMat is a dict {qin: (qout, ndarray)} with charge keys as
tuples (std::map order = tuple order, KeyTypes.hpp:19-58); an MPS tensor is a list of BlockMats
over the physical index.  Entries are complex (numpy complex128) or real.  Each function cites the
reference code it restates:

  products, sums          BlockObj.hpp:2557-2625 (operator*), :1044-1084 (+=, union of sectors);
                          MPSBlockMat.hpp:934-1043
  mixed TMs (unity / X)   TMBlockFunctions.hpp:259-533; unit-cell chains :542-788
  ApplyOperator           MPSBlockUtilities.hpp:18-34; ApplyOpTM*  TMBlockFunctions.hpp:797-864
  gmres                   IterativeSolvers.hpp:21-149 (unrestarted, MGS, complex Givens)
  InvertE_fac / _proj_fac InvEBlockFunctions.hpp:52-125, 166-197, 268-476
  HeffConstants           Excitations_helpers.hpp:260-347
  XDims                   Excitations_helpers.hpp:639-678
  LRoverlaps              Excitations_helpers.hpp:915-946 (TMmixedEigs, EigsBlockFunctions.hpp:282-391)
  TM spectra (dense)      TMEigs / TMmixedEigs, EigsBlockFunctions.hpp:177-391
  ApplyHeff               Excitations_helpers.hpp:20-228
  ApplyTransformation     Excitations_helpers.hpp:680-771 (FlipQN MPSBlockMat.hpp:762-787,
                          BlockObj.hpp:845-855, 1740-1749; FlipK KeyTypes.hpp:156-170)

Deviations: GMRES starts from y0 = 0 when no guess is given (the reference draws a random y0;
only the iterate sequence differs); the dominant mixed-TM eigenvectors LM, RM are obtained by a
dense eigensolve (the reference uses ARPACK) and scaled so that tr(LM RM) = 1, the pairing that
InvEfun_proj_fac's trace(x*rho) forms (the reference leaves ARPACK's normalisation; the projected
quantities do not depend on it for excitation vectors).
Parity of whole ApplyHeff iterate sequences against the reference is unpinned (no reference
fixtures exist).  tests/test_excitations_oracle.py pins this module by: its block algebra (mixed
TMs, MPS products, ApplyOperator, operator TMs) against the separately pinned C++ oracle; the
explicit geometric series sum_m (fac T)^m x for InvertE_fac; the reduction of InvertE_proj_fac at
fac = 1 to the C++ oracle's InvertE_proj; Hermiticity of the assembled H_eff(k) and positivity of
its spectrum above the ground state; the fixed-point eigenvalue 1 of the dense TM spectra.
tests/test_excitations_pins.py pins apply_heff itself: closed-form one-particle bands at bond
dimension 1 (XXZ magnon, Hubbard hole and doublon); its quadratic form <X|H_eff(k)|X> against the
second derivative of the energy density of the perturbed uniform state, computed from dense
symmetry-free transfer matrices, at K = 0 (projected and not) and K = +-2 for k in {0, pi/2, pi}
per cell; the reference's trace checks of HeffConstants and the LRoverlaps closed form (C^T, C);
and it checks that these pins fail when any single term of apply_heff is left out (`omit`).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import numpy as np

Key = Tuple[int, ...]
BM = Dict[Key, Tuple[Key, np.ndarray]]


def key_add(a: Key, b: Key) -> Key:
    return tuple(x + y for x, y in zip(a, b))


def key_sub(a: Key, b: Key) -> Key:
    return tuple(x - y for x, y in zip(a, b))


# ---- conversions from/to the C++ oracle's block objects (oracle.pyoracle.OB) -------------------
def from_ob(ob) -> List[BM]:
    out: List[BM] = [dict() for _ in range(len(ob.nblk))]
    for e, qi, qo, m in ob.blocks():
        out[e][tuple(int(v) for v in qi)] = (tuple(int(v) for v in qo), np.array(m))
    return out


def to_ob(lists: List[BM], nsym: int, like=None):
    """Real block objects in the oracle's OB form (imaginary parts must be zero)."""
    from oracle.pyoracle import OB
    nblk, qin, qout, rows, cols, data = [], [], [], [], [], []
    for bm in lists:
        nblk.append(len(bm))
        for q in sorted(bm):
            qo, m = bm[q]
            qin.append(q)
            qout.append(qo)
            rows.append(m.shape[0])
            cols.append(m.shape[1])
            data.append(np.real(m).reshape(-1, order="F"))
    return OB(nsym, np.array(nblk, np.int32), np.array(qin, np.int32).reshape(-1, nsym), np.array(qout, np.int32).reshape(-1, nsym),
              np.array(rows, np.int32), np.array(cols, np.int32), np.concatenate(data) if data else np.zeros(0))


def dims(bm: BM):
    """Sizes vector (qin, qout, rows, cols) in map order (GetSizesVector, BlockObj.hpp:699-707)."""
    return [(q, bm[q][0], bm[q][1].shape[0], bm[q][1].shape[1]) for q in sorted(bm)]


def vectorize(bm: BM) -> np.ndarray:
    """Concatenation of column-major blocks in map order (Vectorize, BlockObj.hpp:764-780)."""
    parts = [bm[q][1].reshape(-1, order="F") for q in sorted(bm)]
    return np.concatenate(parts) if parts else np.zeros(0)


def from_vec(v: np.ndarray, dv) -> BM:
    out: BM = {}
    pos = 0
    for q, qo, r, c in dv:
        out[q] = (qo, np.asarray(v[pos:pos + r * c]).reshape((r, c), order="F").copy())
        pos += r * c
    return out


def vectorize_list(xs: List[BM]) -> np.ndarray:
    return np.concatenate([vectorize(x) for x in xs])


def from_vec_list(v: np.ndarray, dvs) -> List[BM]:
    out, pos = [], 0
    for dv in dvs:
        n = sum(r * c for _, _, r, c in dv)
        out.append(from_vec(v[pos:pos + n], dv))
        pos += n
    return out


# ---- block algebra -------------------------------------------------------------------------
def bm_add(a: BM, b: BM, alpha=1.0, beta=1.0) -> BM:
    """alpha*a + beta*b with the union of sectors (BlockMat::operator+=, BlockObj.hpp:1044-1084)."""
    out = {q: (qo, alpha * m) for q, (qo, m) in a.items()}
    for q, (qo, m) in b.items():
        if q in out:
            if out[q][0] != qo:
                raise ValueError("sum of blocks with different outgoing charges")
            out[q] = (qo, out[q][1] + beta * m)
        else:
            out[q] = (qo, beta * m)
    return out


def bm_scale(a: BM, s) -> BM:
    return {q: (qo, s * m) for q, (qo, m) in a.items()}


def bm_t(a: BM) -> BM:
    return {qo: (q, m.T) for q, (qo, m) in a.items()}


def bm_mul(l: BM, r: BM) -> BM:
    """BlockMat product: lhs block q -> q' meets the rhs block with Qin = q' (operator*)."""
    out: BM = {}
    for q, (qo, a) in l.items():
        if qo in r:
            qo2, b = r[qo]
            if q in out:
                out[q] = (qo2, out[q][1] + a @ b)
            else:
                out[q] = (qo2, a @ b)
    return out


def dK(bm: BM) -> Optional[Key]:
    for q, (qo, _) in bm.items():
        return key_sub(qo, q)
    return None


def mps_add(a, b, alpha=1.0, beta=1.0):
    return [bm_add(x, y, alpha, beta) for x, y in zip(a, b)]


def mps_scale(a, s):
    return [bm_scale(x, s) for x in a]


def mps_mat(A, M):
    return [bm_mul(a, M) for a in A]


def mat_mps(M, A):
    return [bm_mul(M, a) for a in A]


def mps_mps(A, B):
    """Two-site tensor A[i] B[j] at index i*d + j (MPSBlockMat.hpp:934-1043)."""
    return [bm_mul(a, b) for a in A for b in B]


def tml_unity(A, B) -> BM:
    """ApplyTMmixedLeft(A, B) = sum_s B_s^T A_s, keyed by Qout(B_s) (TMBlockFunctions.hpp:308-333)."""
    out: BM = {}
    for a_s, b_s in zip(A, B):
        if not a_s or not b_s:
            continue
        for q, (qo, a) in a_s.items():
            if q in b_s:
                qb, b = b_s[q]
                t = b.T @ a
                if qb in out:
                    out[qb] = (out[qb][0], out[qb][1] + t)
                else:
                    out[qb] = (qo, t)
    return out


def tmr_unity(A, B) -> BM:
    """ApplyTMmixedRight(A, B) = sum_s A_s B_s^T, keyed by Qin(A_s) (TMBlockFunctions.hpp:445-470)."""
    out: BM = {}
    for a_s, b_s in zip(A, B):
        if not a_s or not b_s:
            continue
        dkb = dK(b_s)
        for q, (qo, a) in a_s.items():
            qb = key_sub(qo, dkb)
            if qb in b_s:
                _, b = b_s[qb]
                t = a @ b.T
                if q in out:
                    out[q] = (out[q][0], out[q][1] + t)
                else:
                    out[q] = (qb, t)
    return out


def tml(A, B, X: BM) -> BM:
    """ApplyTMmixedLeft(A, B, X) = sum_s B_s^T X A_s (TMBlockFunctions.hpp:336-342)."""
    return tml_unity(mat_mps(X, A), B)


def tmr(A, B, X: BM) -> BM:
    """ApplyTMmixedRight(A, B, X) = sum_s A_s X B_s^T (TMBlockFunctions.hpp:476-482)."""
    return tmr_unity(mps_mat(A, X), B)


def tml_chain(As, Bs, X):
    """Unit-cell chain, first site first (TMBlockFunctions.hpp:651-664)."""
    for a, b in zip(As, Bs):
        X = tml(a, b, X)
    return X


def tmr_chain(As, Bs, X):
    """Unit-cell chain, last site first (TMBlockFunctions.hpp:732-746)."""
    for a, b in zip(reversed(As), reversed(Bs)):
        X = tmr(a, b, X)
    return X


def diag_part(bm: BM) -> BM:
    return {q: (qo, m) for q, (qo, m) in bm.items() if qo == q}


def apply_operator(AA, H: np.ndarray):
    """out[row] += H(row, col) * AA[col] over the nonzeros of H (ApplyOperator, MPSBlockUtilities.hpp:18-34)."""
    D = H.shape[0]
    out = [dict() for _ in range(D)]
    for col in range(D):
        for row in range(D):
            h = H[row, col]
            if h != 0.0:
                out[row] = bm_add(out[row], AA[col], 1.0, h)
    return out


def op_tml_gen(H, A, B):
    """ApplyOpTMLeftGen(H, A, B) = ApplyTMmixedLeft(ApplyOperator(A, H), B) (TMBlockFunctions.hpp:822-828)."""
    return tml_unity(apply_operator(A, H), B)


def op_tml_diag(H, AA):
    """ApplyOpTMLeftDiag (TMBlockFunctions.hpp:797-803): diagonal sectors of sum_s AA_s^T (H AA)_s."""
    return diag_part(tml_unity(apply_operator(AA, H), AA))


def op_tmr_diag(H, AA):
    """ApplyOpTMRightDiag (TMBlockFunctions.hpp:834-840)."""
    return diag_part(tmr_unity(apply_operator(AA, H), AA))


def tr_prod(a: BM, b: BM):
    """trace(a*b) summed over sectors."""
    return sum(np.trace(m) for _, (_, m) in bm_mul(a, b).items())


def eye(dm: BM) -> BM:
    return {q: (q, np.eye(m.shape[0])) for q, (_, m) in dm.items()}


# ---- GMRES (IterativeSolvers.hpp:21-149) -----------------------------------------------------
def gmres(Afun, b: np.ndarray, x0: np.ndarray, tol: float, maxit: int = 0):
    """Unrestarted GMRES with modified Gram-Schmidt and complex Givens rotations; stops when the
    residual estimate |gamma(j+1)| < tol*||b||.  Returns (x, flag) with flag = steps or < 0."""
    dim = b.size
    maxit = dim if maxit == 0 or maxit > dim else maxit
    kry = min(dim, maxit)
    dt = np.result_type(b.dtype, x0.dtype, np.float64)
    r = b - Afun(x0)
    x = x0.astype(dt).copy()
    bnorm = np.linalg.norm(b)
    rnorm = np.linalg.norm(r)
    if bnorm == 0.0:
        return np.zeros(dim, dt), 0
    V = np.zeros((dim, kry + 1), dt)
    Hm = np.zeros((kry + 1, kry + 1), dt)
    c = np.zeros(kry + 1, dt)
    s = np.zeros(kry + 1, dt)
    g = np.zeros(kry + 1, dt)
    g[0] = rnorm
    V[:, 0] = r / rnorm
    kdim, flag = 0, -1
    for j in range(kry):
        w = Afun(V[:, j]).astype(dt)
        for i in range(j + 1):
            Hm[i, j] = np.vdot(V[:, i], w)
            w = w - Hm[i, j] * V[:, i]
        Hm[j + 1, j] = np.linalg.norm(w)
        for i in range(j):
            h1, h2 = Hm[i, j], Hm[i + 1, j]
            Hm[i, j] = c[i] * h1 + s[i] * h2
            Hm[i + 1, j] = -np.conj(s[i]) * h1 + c[i] * h2
        beta = np.linalg.norm(Hm[j:j + 2, j])
        if abs(beta) < 2e-16:
            flag = -2
            break
        V[:, j + 1] = w / Hm[j + 1, j] if Hm[j + 1, j] != 0 else 0.0  # exact breakdown: the next vector is not used
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
    if kdim:
        alpha = np.linalg.solve(np.triu(Hm[:kdim, :kdim]), g[:kdim])
        x = x + V[:, :kdim] @ alpha
    return x, flag


# ---- geometric sums (InvEBlockFunctions.hpp) -------------------------------------------------
def _map_onto(tx: BM, dv) -> np.ndarray:
    """Vectorize tx on the sector list dv (sectors of tx outside dv are dropped, missing ones zero)."""
    parts = []
    for q, qo, r, c in dv:
        if q in tx and tx[q][0] == qo:
            parts.append(tx[q][1].reshape(-1, order="F"))
        else:
            parts.append(np.zeros(r * c))
    return np.concatenate(parts) if parts else np.zeros(0)


def invert_e_fac(As, Bs, x: BM, direction: int, fac, tol=1e-14, y0: Optional[BM] = None):
    """Solves (y|[1 - fac T] = (x| (l) or [1 - fac T]|y) = |x) (r) with T the mixed unit-cell TM of
    (As, Bs) (InvertE_fac, InvEBlockFunctions.hpp:436-476; Afun = InvEfun_fac :166-197)."""
    dv = dims(x)
    T = (lambda v: tml_chain(As, Bs, v)) if direction == 0 else (lambda v: tmr_chain(As, Bs, v))
    Afun = lambda v: v - fac * _map_onto(T(from_vec(v, dv)), dv)  # noqa: E731
    b = vectorize(x)
    v0 = vectorize(y0) if y0 is not None and dims(y0) == dv else np.zeros(b.size, np.result_type(b, complex(fac)))
    y, flag = gmres(Afun, b.astype(np.result_type(b, complex(fac))), v0, tol)
    return from_vec(y, dv), flag


def invert_e_proj_fac(As, Bs, x: BM, L: BM, R: BM, direction: int, fac, tol=1e-14, y0: Optional[BM] = None):
    """Solves (y|[1 - fac T + |R)(L|] = (x|[1 - |R)(L|] (l) or the mirrored right equation
    (InvertE_proj_fac, InvEBlockFunctions.hpp:268-331; Afun = InvEfun_proj_fac :52-125)."""
    dv = dims(x)
    b = vectorize(x).astype(np.complex128)
    if direction == 0:
        b = b - tr_prod(x, R) * _map_onto(L, dv)
        T = lambda v: tml_chain(As, Bs, v)  # noqa: E731
        rin, rout = R, L
    else:
        b = b - tr_prod(L, x) * _map_onto(R, dv)
        T = lambda v: tmr_chain(As, Bs, v)  # noqa: E731
        rin, rout = L, R

    def Afun(v):
        xin = from_vec(v, dv)
        tr = tr_prod(xin, rin) if direction == 0 else tr_prod(rin, xin)
        return v + tr * _map_onto(rout, dv) - fac * _map_onto(T(xin), dv)
    v0 = vectorize(y0).astype(np.complex128) if y0 is not None and dims(y0) == dv else np.zeros(b.size, np.complex128)
    y, flag = gmres(Afun, b, v0, tol)
    return from_vec(y, dv), flag


# ---- excitation set-up -----------------------------------------------------------------------
def heff_bond_sums(AL, AR, H):
    """The unit-cell bond-energy sums HL[n], HR[n] of HeffConstants before the environment solves
    (Excitations_helpers.hpp:281-303).  With E0 subtracted from H, tr(HL[N-1] R) and tr(L HR[0])
    vanish: the reference's own check that the bond energies average to E0 (:287-291, :299-303)."""
    N = len(AL)
    HL = [op_tml_diag(H, mps_mps(AL[-1], AL[0]))]
    for n in range(1, N):
        HL.append(bm_add(op_tml_diag(H, mps_mps(AL[n - 1], AL[n])), diag_part(tml(AL[n], AL[n], HL[-1]))))
    HR = [op_tmr_diag(H, mps_mps(AR[-1], AR[0]))]
    for n in range(N - 2, -1, -1):
        HR.insert(0, bm_add(op_tmr_diag(H, mps_mps(AR[n], AR[n + 1])), diag_part(tmr(AR[n], AR[n], HR[0]))))
    return HL, HR


def heff_constants(AL, AR, L: BM, R: BM, H, tol=1e-14):
    """HLtot, HRtot (HeffConstants, Excitations_helpers.hpp:260-347); H must have E0 subtracted."""
    N = len(AL)
    HL, HR = heff_bond_sums(AL, AR, H)
    EHL = [None] * N
    EHR = [None] * N
    EHL[N - 1], _ = invert_e_proj_fac(AL, AL, HL[-1], eye(R), R, 0, 1.0, tol)
    EHL[N - 1] = {q: (qo, np.real(m)) for q, (qo, m) in EHL[N - 1].items()}
    HLtot = [None] * N
    HLtot[N - 1] = EHL[N - 1]
    for n in range(N - 1):
        EHL[n] = tml(AL[n], AL[n], EHL[(n - 1) % N])
        HLtot[n] = bm_add(HL[n], EHL[n])
    EHR[0], _ = invert_e_proj_fac(AR, AR, HR[0], L, eye(L), 1, 1.0, tol)
    EHR[0] = {q: (qo, np.real(m)) for q, (qo, m) in EHR[0].items()}
    HRtot = [None] * N
    HRtot[0] = EHR[0]
    for n in range(N - 1, 0, -1):
        EHR[n] = tmr(AR[n], AR[n], EHR[(n + 1) % N])
        HRtot[n] = bm_add(EHR[n], HR[n])
    return HLtot, HRtot


def xdims(NL, AR, K: Key):
    """Sector list of the excitation parameters X[n]: Qout(NL[n]) -> same + K, with sizes from NL's
    right and AR[n+1]'s left dimensions (XDims, Excitations_helpers.hpp:639-678)."""
    N = len(NL)
    out = []
    for n in range(N):
        ml = {}
        for bm in NL[n]:
            for q, (qo, m) in bm.items():
                ml[qo] = m.shape[1]
        mr = {}
        for bm in AR[(n + 1) % N]:
            for q, (qo, m) in bm.items():
                mr[q] = m.shape[0]
        dv = []
        for q in sorted(ml):
            qk = key_add(q, K)
            if qk in mr and ml[q] > 0 and mr[qk] > 0:
                dv.append((q, qk, ml[q], mr[qk]))
        out.append(dv)
    return out


def _tm_sector_dims(A, B, direction, K: Key):
    """Sectors of the mixed-TM eigenproblem in charge K (TMmixedEigs, EigsBlockFunctions.hpp:282-322)."""
    def ml_of(M):
        d = {}
        for bm in M[0]:
            for q, (qo, m) in bm.items():
                d[q] = m.shape[0]
        return d
    diml, dimr = ml_of(A), ml_of(B)
    if direction == 0:
        diml, dimr = dimr, diml
    dv = []
    for q in sorted(diml):
        qk = key_add(q, K)
        if qk in dimr:
            dv.append((q, qk, diml[q], dimr[qk]))
    return dv


def tm_eigs_dense(As, Bs, direction: int, K: Key):
    """All eigenvalues (sorted by decreasing magnitude) of the unit-cell mixed TM of (As, Bs) on the
    charge-K sector list of TMmixedEigs (EigsBlockFunctions.hpp:282-391), by a dense eigensolve."""
    dv = _tm_sector_dims(As, Bs, direction, K)
    n = sum(r * c for *_, r, c in dv)
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        X = from_vec(e, dv)
        T = tml_chain(As, Bs, X) if direction == 0 else tmr_chain(As, Bs, X)
        M[:, j] = np.real(_map_onto(T, dv))
    w = np.linalg.eigvals(M)
    return w[np.argsort(-np.abs(w))]


def lr_overlaps(AL, AR, K: Key):
    """Dominant left/right eigenmatrices LM, RM of the mixed TM (AL, AR) in sectors -K / K and the
    mean dominant eigenvalue (LRoverlaps, Excitations_helpers.hpp:915-946), by dense eigensolve."""
    mK = tuple(-k for k in K)
    res = []
    for direction, KK in ((0, mK), (1, K)):
        dv = _tm_sector_dims(AL, AR, direction, KK)
        n = sum(r * c for *_, r, c in dv)
        M = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1.0
            X = from_vec(e, dv)
            T = tml_chain(AL, AR, X) if direction == 0 else tmr_chain(AL, AR, X)
            M[:, j] = np.real(_map_onto(T, dv))
        w, V = np.linalg.eig(M)
        i = int(np.argmax(np.abs(w)))
        v = np.real(V[:, i] * np.exp(-1j * np.angle(V[np.argmax(np.abs(V[:, i])), i])))
        res.append((w[i], from_vec(v / np.linalg.norm(v), dv)))
    (wl, LM), (wr, RM) = res
    s = tr_prod(LM, RM)  # the pairing of the projectors: tr(LM RM) (InvEfun_proj_fac's trace(x*rho))
    RM = bm_scale(RM, 1.0 / s)
    return LM, RM, 0.5 * (np.real(wl) + np.real(wr))


# ---- ApplyTransformation (Excitations_helpers.hpp:680-771) -------------------------------------
def transformation_masks(model: str, op: str, N: int):
    """(which charge components flip, per-site sign flags over the physical index) of
    ApplyTransformation: XXZ "FS" flips the one U(1) charge, no signs (:693-701); FHUB in the (N, M)
    representation (FHUB_NM_REP, the reference's build) "FS" flips M with a sign on |3> (:704-716),
    "FC" flips N with the signs {2} on even and {0, 2, 3} on odd sites, N even (:717-737)."""
    if model == "XXZ" and op == "FS":
        return (True,), [(False, False)] * N
    if model == "FHUB" and op == "FS":
        return (False, True), [(False, False, False, True)] * N
    if model == "FHUB" and op == "FC":
        if N % 2:
            raise ValueError("unit cell size needs to be even for charge flip")
        return (True, False), [((False, False, True, False) if n % 2 == 0 else (True, False, True, True)) for n in range(N)]
    raise ValueError(f"no transformation {op} for {model}")


def flip_key(q: Key, which) -> Key:
    """FlipK (KeyTypes.hpp:156-170): negate the masked components."""
    return tuple(-x if w else x for x, w in zip(q, which))


def flip_bm(bm: BM, which) -> BM:
    """BlockMat::FlipQN / BlockDiagMat::FlipQN: same matrices under flipped charges."""
    return {flip_key(q, which): (flip_key(qo, which), m) for q, (qo, m) in sorted(bm.items())}


def apply_transformation(AR, L: BM, I2K, which, signs):
    """AR -> FlipQN(I2K, which) then the sign flags per site; L -> FlipQN(which).  Entry i of an
    MPS tensor (charge dK = Qout - Qin, I2K[i] when empty) moves to the position of its flipped
    charge in I2K."""
    order = {tuple(k): i for i, k in enumerate(I2K)}
    out = []
    for n, A in enumerate(AR):
        B = [dict() for _ in A]
        for i, bm in enumerate(A):
            k = dK(bm)
            k = tuple(I2K[i]) if k is None else k
            f = flip_key(k, which)
            if f not in order:
                raise ValueError(f"key {f} not found in target I2K")
            B[order[f]] = flip_bm(bm, which)
        for s, neg in enumerate(signs[n]):
            if neg:
                B[s] = bm_scale(B[s], -1.0)
        out.append(B)
    return out, flip_bm(L, which)


# ---- ApplyHeff (Excitations_helpers.hpp:20-228) -----------------------------------------------
HEFF_TERMS = ("ABR_chain", "ABRtot_prop", "HBL_env", "HBL_op_left", "HBL_op_right", "HBL_chain", "HBLtot_prop",
              "on_HL", "on_HR", "on_HAB", "on_HBA", "left_HBLtot", "left_wrap", "left_inner",
              "right_HL", "right_HAA", "right_HAB")


def apply_heff(X: List[BM], kfac, AL, AR, NL, LM: Optional[BM], RM: Optional[BM], H, HLtot, HRtot, tol=1e-14, omit=()):
    """Effective Hamiltonian of the momentum-k excitation ansatz on the parameters X[n]
    (B[n] = NL[n] X[n]); returns X_out in the sector structure of X.

    `omit` (test infrastructure): names from HEFF_TERMS whose contribution is left out, so that a
    test can check that its pins notice the loss of any single term."""
    for t in omit:
        if t not in HEFF_TERMS:
            raise ValueError(f"unknown ApplyHeff term {t}")
    keep = lambda t: t not in omit  # noqa: E731
    zero = lambda bm: {q: (qo, 0.0 * m) for q, (qo, m) in bm.items()}  # noqa: E731
    use = lambda t, bm: bm if keep(t) else zero(bm)  # noqa: E731
    usem = lambda t, A: A if keep(t) else [zero(a) for a in A]  # noqa: E731
    N = len(X)
    proj = LM is not None and RM is not None
    ck = np.conj(kfac)
    P = lambda x: x % N  # noqa: E731
    Bin = [mps_mat(NL[n], X[n]) for n in range(N)]
    # terms where Bin is right of Bout
    ABR = [tmr_unity(Bin[-1], AR[-1])]
    for n in range(N - 1, 0, -1):
        ABR.insert(0, tmr_unity(mps_add(Bin[n - 1], usem("ABR_chain", mps_mat(AL[n - 1], ABR[0]))), AR[n - 1]))
    if proj:
        EBR, _ = invert_e_proj_fac(AL, AR, ABR[0], LM, RM, 1, kfac, tol)
    else:
        EBR, _ = invert_e_fac(AL, AR, ABR[0], 1, kfac, tol)
    ABRtot = [bm_scale(EBR, kfac)]
    for n in range(N - 1, 0, -1):
        EBR = tmr(AL[n], AR[n], EBR)
        ABRtot.insert(0, bm_add(ABR[n], use("ABRtot_prop", bm_scale(EBR, kfac))))
    ABRtot = ABRtot[-1:] + ABRtot[:-1]  # shift(ABRtot, 1): last element to the front
    # terms where Bin is left of Bout (H left of or on Bin)
    HBL = [bm_add(use("HBL_env", tml_unity(mat_mps(HLtot[-1], Bin[0]), AL[0])),
                  op_tml_gen(H, mps_add(usem("HBL_op_left", mps_mps(AL[-1], Bin[0])),
                                        usem("HBL_op_right", mps_scale(mps_mps(Bin[-1], AR[0]), ck))), mps_mps(AL[-1], AL[0])))]
    for n in range(1, N):
        HBL.append(bm_add(tml_unity(mps_add(usem("HBL_chain", mat_mps(HBL[-1], AR[n])), usem("HBL_env", mat_mps(HLtot[n - 1], Bin[n]))),
                                    AL[n]),
                          op_tml_gen(H, mps_add(usem("HBL_op_left", mps_mps(AL[n - 1], Bin[n])),
                                                usem("HBL_op_right", mps_mps(Bin[n - 1], AR[n]))), mps_mps(AL[n - 1], AL[n]))))
    LMt = bm_t(LM) if proj else None
    RMt = bm_t(RM) if proj else None
    if proj:
        EHBL, _ = invert_e_proj_fac(AR, AL, HBL[-1], LMt, RMt, 0, ck, tol)
    else:
        EHBL, _ = invert_e_fac(AR, AL, HBL[-1], 0, ck, tol)
    HBLtot = [bm_scale(EHBL, ck)]
    for n in range(N - 1):
        EHBL = tml(AR[n], AL[n], EHBL)
        HBLtot.append(bm_add(HBL[n], use("HBLtot_prop", bm_scale(EHBL, ck))))
    HBLtot = HBLtot[1:] + HBLtot[:1]  # shift(HBLtot, -1): first element to the back
    d = len(AL[0])
    # Bin and Bout on top of each other
    Bout = [mps_add(usem("on_HL", mat_mps(HLtot[P(n - 1)], Bin[n])), usem("on_HR", mps_mat(Bin[n], HRtot[P(n + 1)]))) for n in range(N)]
    for n in range(N):
        HAB = apply_operator(mps_mps(AL[P(n - 1)], Bin[n]), H)
        HBA = apply_operator(mps_mps(Bin[n], AR[P(n + 1)]), H)
        for s in range(d):
            for k in range(d):
                Bout[n][s] = bm_add(Bout[n][s], use("on_HAB", bm_mul(bm_t(AL[P(n - 1)][k]), HAB[k * d + s])))
                Bout[n][s] = bm_add(Bout[n][s], use("on_HBA", bm_mul(HBA[s * d + k], bm_t(AR[P(n + 1)][k]))))
    # Bin left of Bout
    Bout[0] = mps_add(Bout[0], usem("left_HBLtot", mat_mps(HBLtot[-1], AR[0])))
    HBA = apply_operator(mps_mps(Bin[-1], AR[0]), H)
    for s in range(d):
        for k in range(d):
            Bout[0][s] = bm_add(Bout[0][s], use("left_wrap", bm_scale(bm_mul(bm_t(AL[-1][k]), HBA[k * d + s]), ck)))
    for n in range(1, N):
        Bout[n] = mps_add(Bout[n], usem("left_HBLtot", mat_mps(HBLtot[n - 1], AR[n])))
        HBA = apply_operator(mps_mps(Bin[n - 1], AR[n]), H)
        for s in range(d):
            for k in range(d):
                Bout[n][s] = bm_add(Bout[n][s], use("left_inner", bm_mul(bm_t(AL[n - 1][k]), HBA[k * d + s])))
    # Bin right of Bout
    for n in range(N):
        Atmp = mps_add(Bin[P(n + 1)], mps_mat(AL[P(n + 1)], ABRtot[P(n + 2)]))
        if n == N - 1:
            Atmp = mps_scale(Atmp, kfac)
        Bout[n] = mps_add(Bout[n], usem("right_HL", mat_mps(HLtot[P(n - 1)], mps_mat(AL[n], ABRtot[P(n + 1)]))))
        HAA = apply_operator(mps_mps(AL[P(n - 1)], mps_mat(AL[n], ABRtot[P(n + 1)])), H)
        HAB = apply_operator(mps_mps(AL[n], Atmp), H)
        for s in range(d):
            for k in range(d):
                Bout[n][s] = bm_add(Bout[n][s], use("right_HAA", bm_mul(bm_t(AL[P(n - 1)][k]), HAA[k * d + s])))
                Bout[n][s] = bm_add(Bout[n][s], use("right_HAB", bm_mul(HAB[s * d + k], bm_t(AR[P(n + 1)][k]))))
    # X_out = NL^T B_out, kept on the structure of X
    out = []
    for n in range(N):
        t = tml_unity(Bout[n], NL[n])
        dv = dims(X[n])
        out.append(from_vec(_map_onto(t, dv), dv))
    return out


# ---- excitation observables (Excitations_helpers.hpp:350-600, 774-826) ------------------------------
def opeff_constants(AL, AR, L: BM, R: BM, O, tol=1e-14):
    """OPLtot, OPRtot of a one-site operator O (d x d, its expectation subtracted) and the
    reference's two checks tr(OPL R), tr(L OPR) (OPeffConstants, Excitations_helpers.hpp:531-600):
    OPL[n] = sum over the cell's sites m <= n of O at m carried to bond n + 1 by the left TMs, OPR
    the mirror; the sums over all other cells by the projected geometric sums (InvertE_proj)."""
    N = len(AL)
    OPL = [op_tml_diag(O, AL[0])]
    for n in range(1, N):
        OPL.append(bm_add(op_tml_diag(O, AL[n]), diag_part(tml(AL[n], AL[n], OPL[-1]))))
    OPR = [op_tmr_diag(O, AR[-1])]
    for n in range(N - 2, -1, -1):
        OPR.insert(0, bm_add(op_tmr_diag(O, AR[n]), diag_part(tmr(AR[n], AR[n], OPR[0]))))
    checks = (tr_prod(OPL[-1], R), tr_prod(L, OPR[0]))
    real = lambda bm: {q: (qo, np.real(m)) for q, (qo, m) in bm.items()}  # noqa: E731
    EOPL = [None] * N
    EOPR = [None] * N
    OPLtot = [None] * N
    OPRtot = [None] * N
    EOPL[N - 1] = real(invert_e_proj_fac(AL, AL, OPL[-1], eye(R), R, 0, 1.0, tol)[0])
    OPLtot[N - 1] = EOPL[N - 1]
    for n in range(N - 1):
        EOPL[n] = tml(AL[n], AL[n], EOPL[(n - 1) % N])
        OPLtot[n] = bm_add(OPL[n], EOPL[n])
    EOPR[0] = real(invert_e_proj_fac(AR, AR, OPR[0], L, eye(L), 1, 1.0, tol)[0])
    OPRtot[0] = EOPR[0]
    for n in range(N - 1, 0, -1):
        EOPR[n] = tmr(AR[n], AR[n], EOPR[(n + 1) % N])
        OPRtot[n] = bm_add(EOPR[n], OPR[n])
    return OPLtot, OPRtot, checks


def _norm_inf(bm: BM) -> float:
    return max((float(np.abs(m).sum(axis=1).max()) for _, (_, m) in bm.items() if m.size), default=0.0)


OPEFF_TERMS = ("ABR_chain", "ABRtot_prop", "OPBL_env", "OPBL_op", "OPBL_chain", "EOPBL", "on_OPL", "on_OPR", "on_O", "left_OPBL",
               "right_OPL", "right_O")


def apply_opeff(X: List[BM], kfac, AL, AR, NL, LM: Optional[BM], RM: Optional[BM], O, OPLtot, OPRtot, tol=1e-14, omit=()):
    """The momentum-k effective operator of a one-site observable O (expectation subtracted) on the
    excitation parameters X[n] (ApplyOPeff, Excitations_helpers.hpp:350-528): the terms with Bin
    right of Bout (the same ABR / EBR sums as ApplyHeff), Bin left of Bout (OPBL and its geometric
    sum EOPBL, skipped when OPBL's norm is below 10 tol as the reference does), and both on the same
    site (OPLtot B + B OPRtot + O B); returns X_out in the structure of X.  `omit` (test
    infrastructure): names from OPEFF_TERMS left out, so that a test can check its pins see each."""
    for t in omit:
        if t not in OPEFF_TERMS:
            raise ValueError(f"unknown ApplyOPeff term {t}")
    zero = lambda bm: {q: (qo, 0.0 * m) for q, (qo, m) in bm.items()}  # noqa: E731
    use = lambda t, bm: zero(bm) if t in omit else bm  # noqa: E731
    usem = lambda t, A: [zero(a) for a in A] if t in omit else A  # noqa: E731
    N = len(X)
    proj = LM is not None and RM is not None
    ck = np.conj(kfac)
    P = lambda x: x % N  # noqa: E731
    Bin = [mps_mat(NL[n], X[n]) for n in range(N)]
    ABR = [tmr_unity(Bin[-1], AR[-1])]
    for n in range(N - 1, 0, -1):
        ABR.insert(0, tmr_unity(mps_add(Bin[n - 1], usem("ABR_chain", mps_mat(AL[n - 1], ABR[0]))), AR[n - 1]))
    if proj:
        EBR, _ = invert_e_proj_fac(AL, AR, ABR[0], LM, RM, 1, kfac, tol)
    else:
        EBR, _ = invert_e_fac(AL, AR, ABR[0], 1, kfac, tol)
    ABRtot = [bm_scale(EBR, kfac)]
    for n in range(N - 1, 0, -1):
        EBR = tmr(AL[n], AR[n], EBR)
        ABRtot.insert(0, bm_add(ABR[n], use("ABRtot_prop", bm_scale(EBR, kfac))))
    ABRtot = ABRtot[-1:] + ABRtot[:-1]  # shift(ABRtot, 1)
    OPBL = [bm_add(use("OPBL_env", tml_unity(mat_mps(OPLtot[-1], Bin[0]), AL[0])), use("OPBL_op", op_tml_gen(O, Bin[0], AL[0])))]
    for n in range(1, N):
        OPBL.append(bm_add(tml_unity(mps_add(usem("OPBL_chain", mat_mps(OPBL[-1], AR[n])), usem("OPBL_env", mat_mps(OPLtot[n - 1], Bin[n]))),
                                     AL[n]),
                           use("OPBL_op", op_tml_gen(O, Bin[n], AL[n]))))
    have = _norm_inf(OPBL[-1]) > 10 * tol and "EOPBL" not in omit
    EOPBL = None
    if have:
        if proj:
            EOPBL, _ = invert_e_proj_fac(AR, AL, OPBL[-1], bm_t(LM), bm_t(RM), 0, ck, tol)
        else:
            EOPBL, _ = invert_e_fac(AR, AL, OPBL[-1], 0, ck, tol)
    # Bin and Bout on the same site
    Bout = [mps_add(mps_add(usem("on_OPL", mat_mps(OPLtot[P(n - 1)], Bin[n])), usem("on_OPR", mps_mat(Bin[n], OPRtot[P(n + 1)]))),
                    usem("on_O", apply_operator(Bin[n], O)))
            for n in range(N)]
    # Bin left of Bout
    if have:
        Bout[0] = mps_add(Bout[0], mat_mps(bm_scale(EOPBL, ck), AR[0]))
    for n in range(1, N):
        tmp = use("left_OPBL", OPBL[n - 1])
        if have:
            EOPBL = tml(AR[n - 1], AL[n - 1], EOPBL)
            tmp = bm_add(tmp, bm_scale(EOPBL, ck))
        Bout[n] = mps_add(Bout[n], mat_mps(tmp, AR[n]))
    # Bin right of Bout
    for n in range(N):
        Bout[n] = mps_add(Bout[n], usem("right_OPL", mat_mps(OPLtot[P(n - 1)], mps_mat(AL[n], ABRtot[P(n + 1)]))))
        Bout[n] = mps_add(Bout[n], usem("right_O", mps_mat(apply_operator(AL[n], O), ABRtot[P(n + 1)])))
    out = []
    for n in range(N):
        t = tml_unity(Bout[n], NL[n])
        dv = dims(X[n])
        out.append(from_vec(_map_onto(t, dv), dv))
    return out


def measure_excitation(X: List[BM], kfac, AL, AR, NL, LM, RM, O, OPLtot, OPRtot, tol=1e-14, omit=()) -> np.ndarray:
    """Per site <X[n] | (O_eff X)[n]> (MeasureExcitations, Excitations_helpers.hpp:774-826, on a
    given X; the reference draws a random normalised one)."""
    Y = apply_opeff(X, kfac, AL, AR, NL, LM, RM, O, OPLtot, OPRtot, tol, omit)
    return np.array([np.vdot(vectorize(x), vectorize(y)) for x, y in zip(X, Y)])
