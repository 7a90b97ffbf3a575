"""Independent pins of the excitation oracle's ApplyHeff (oracle/excitations.py), on the CPU.

1. Closed forms at bond dimension 1, where the excitation ansatz is exact:
   * the one-magnon band of the ferromagnetic XXZ chain in the reference's convention
     (Models.cpp:57-93: H = -J (SxSx + SySy) - Delta SzSz - hz/2 (Sz x 1 + 1 x Sz)) above the fully
     polarised state: eps(k) = Delta + hz - J cos k;
   * the hole and the doublon of the fully polarised Fermi-Hubbard chain (Models.cpp:543-572; no
     down electrons, so no fermion signs): eps(k) = U/2 +- 2 t cos k.
   With the two-site cell and the reference's unit-cell phase p_act = p pi N (Excitations_main.cpp:
   258-259), H_eff(p) holds the two bands at per-site momenta p pi and p pi + pi.  The bookkeeping's
   |m| < N rule (Models.cpp:222) excludes full polarisation, so the states are built directly; the
   charges are the conserved ones relabelled per site (a constant shift per physical state).
2. The quadratic form <X|H_eff(k)|X> against the second derivative of the energy density of the
   perturbed uniform state.  With G[n] = NL[n] X[n] C[n]^-1 (the tangent vector in the left gauge,
   C[n] on the bond right of site n), the state whose cell c on a q-cell supercell carries
   AL[n] + eps e^{ikc} G[n] (k = 2 pi m / q) has, per cell,
        e(eps) = e0 + |eps|^2 <X|H_eff(k)|X> + Re(c eps^2) + O(eps^3),
   where c (two insertions against the ground state) is removed by averaging eps over {r, ir, -r,
   -ir}.  e(eps) is computed from dense symmetry-free transfer matrices and their dominant
   eigenvectors (numpy), a route that shares nothing with the block-sparse algebra of ApplyHeff.
3. The reference's own consistency checks: HeffConstants' tr(HL R) = tr(L HR) = 0 with E0
   subtracted (Excitations_helpers.hpp:287-303) and LRoverlaps' equal, real dominant overlaps of the
   mixed transfer matrix (:937-942), whose eigenmatrices are C^T and C in the K = 0 sector.
4. Each of these pins fails when any single term of ApplyHeff is left out (oracle `omit` hook).
"""
import functools

import numpy as np
import pytest

from oracle import excitations as ex
from oracle import pyoracle as orc
import excitation_state as es

XXZ, FHUB = 0, 1


# ---- bond-dimension-1 polarised states -----------------------------------------------------------
def _product_state(d, nsym, up, charges):
    """AL = AR = |up> on every site of a two-site cell, NL the other local states, C = 1."""
    z = tuple([0] * nsym)
    one = np.ones((1, 1))
    A = [dict() for _ in range(d)]
    A[up] = {z: (z, one)}
    NL = [dict() for _ in range(d)]
    for s in range(d):
        if s != up:
            NL[s] = {z: (tuple(charges[s]), one)}
    C = {z: (z, one)}
    return [A, A], [A, A], [NL, NL], [C, C]


def _heff_dense(AL, AR, NL, C, H0, K, p, omit=()):
    HLtot, HRtot = ex.heff_constants(AL, AR, C[-1], C[-1], H0)
    xd = ex.xdims(NL, AR, K)
    n = sum(r * c for dv in xd for *_, r, c in dv)
    kfac = np.exp(1j * p * np.pi * 2)
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        M[:, j] = ex.vectorize_list(ex.apply_heff(ex.from_vec_list(e, xd), kfac, AL, AR, NL, None, None, H0, HLtot, HRtot, omit=omit))
    return M


def _magnon_case(J=1.0, Delta=1.3, hz=0.3):
    H = np.array(orc.model_ham(XXZ, [J, Delta, hz]))
    AL, AR, NL, C = _product_state(2, 1, 0, [(0,), (-2,)])
    E0 = -Delta / 4 - hz / 2  # per bond: <up up| H |up up>
    return AL, AR, NL, C, H - E0 * np.eye(4), (2,)


def _hubbard_case(K, t=1.0, U=4.0):
    H = np.array(orc.model_ham(FHUB, [t, U, 0.0, 0.0, 0.0]))
    # local states 0, down, up, up-down; charges of the (N, M) representation shifted by -(0, 1)
    AL, AR, NL, C = _product_state(4, 2, 2, [(-1, -1), (0, -2), (0, 0), (1, -1)])
    E0 = -U / 4  # per bond: U/2 ((-1/4) + (-1/4))
    return AL, AR, NL, C, H - E0 * np.eye(16), K


@pytest.mark.parametrize("p", [0.0, 0.17, 0.5, 0.8, 1.0])
def test_magnon_dispersion(p):
    J, Delta, hz = 1.0, 1.3, 0.3
    AL, AR, NL, C, H0, K = _magnon_case(J, Delta, hz)
    M = _heff_dense(AL, AR, NL, C, H0, K, p)
    assert M.shape == (2, 2)
    w = np.sort(np.linalg.eigvalsh(0.5 * (M + M.conj().T)))
    ref = np.sort([Delta + hz - J * np.cos(p * np.pi), Delta + hz - J * np.cos(p * np.pi + np.pi)])
    assert np.abs(M - M.conj().T).max() < 1e-14
    assert np.abs(w - ref).max() < 1e-13


@pytest.mark.parametrize("K", [(1, 1), (-1, 1)], ids=["hole", "doublon"])
@pytest.mark.parametrize("p", [0.0, 0.23, 0.5, 1.0])
def test_hubbard_hole_and_doublon_dispersion(K, p):
    t, U = 1.0, 4.0
    AL, AR, NL, C, H0, K = _hubbard_case(K, t, U)
    M = _heff_dense(AL, AR, NL, C, H0, K, p)
    assert M.shape == (2, 2)
    w = np.sort(np.linalg.eigvalsh(0.5 * (M + M.conj().T)))
    c = 2 * t * np.cos(p * np.pi)
    assert np.abs(w - np.sort([U / 2 + c, U / 2 - c])).max() < 1e-13


# ---- energy density of the perturbed uniform state ------------------------------------------------
def _bond_maps(s):
    """Dense offsets of the sectors on bonds 0 and 1 of the two-site cell (ascending keys)."""
    maps = []
    for n in range(2):
        keys = {}
        for bm in s["AL"][n]:
            for q, (qo, m) in bm.items():
                keys[q] = m.shape[0]
        off, pos = {}, 0
        for q in sorted(keys):
            off[q] = pos
            pos += keys[q]
        maps.append((off, pos))
    return maps


def _dense_site(A, lmap, rmap, d):
    (loff, Dl), (roff, Dr) = lmap, rmap
    out = np.zeros((d, Dl, Dr), complex)
    for sidx, bm in enumerate(A):
        for q, (qo, m) in bm.items():
            out[sidx, loff[q]:loff[q] + m.shape[0], roff[qo]:roff[qo] + m.shape[1]] = m
    return out


def _dense_bm(bm, lmap, rmap):
    (loff, Dl), (roff, Dr) = lmap, rmap
    out = np.zeros((Dl, Dr))
    for q, (qo, m) in bm.items():
        out[loff[q]:loff[q] + m.shape[0], roff[qo]:roff[qo] + m.shape[1]] = np.real(m)
    return out


def _site_tm(T):
    return sum(np.kron(T[s], T[s].conj()) for s in range(T.shape[0]))


def _bond_tm(Ta, Tb, H0, d):
    out = 0
    for i2 in range(d):
        for j2 in range(d):
            ket = sum(H0[i2 * d + j2, i * d + j] * (Ta[i] @ Tb[j]) for i in range(d) for j in range(d))
            out = out + np.kron(ket, (Ta[i2] @ Tb[j2]).conj())
    return out


def _energy_per_cell(T, H0, d):
    """Energy per two-site cell of the uniform state with periodic site tensors T (dense TMs)."""
    L = len(T)
    M = [_site_tm(t) for t in T]
    E = functools.reduce(np.matmul, M)
    w, V = np.linalg.eig(E)
    i = int(np.argmax(np.abs(w)))
    lam, r = w[i], V[:, i]
    wl, Vl = np.linalg.eig(E.T)
    l = Vl[:, int(np.argmax(np.abs(wl)))]
    norm = l @ r
    tot = 0.0
    for b in range(L):
        left = functools.reduce(np.matmul, M[:b], np.eye(M[0].shape[0]))
        if b < L - 1:
            right = functools.reduce(np.matmul, M[b + 2:], np.eye(M[b + 1].shape[1])) if b + 2 < L else np.eye(M[-1].shape[1])
            tot += (l @ left @ _bond_tm(T[b], T[b + 1], H0, d) @ right @ r) / (lam * norm)
        else:
            right = functools.reduce(np.matmul, M[1:])
            tot += (l @ left @ _bond_tm(T[b], T[0], H0, d) @ right @ r) / (lam * lam * norm)
    return tot / (L // 2)


def _hessian_form(s, X, K, q, m):
    """<X|H_eff(k)|X> at k = 2 pi m / q from the energy of the perturbed uniform state."""
    d = s["d"]
    maps = _bond_maps(s)
    AL = [_dense_site(s["AL"][n], maps[n], maps[(n + 1) % 2], d) for n in range(2)]
    G = []
    for n in range(2):
        B = ex.mps_mat(s["NL"][n], X[n])
        Bd = _dense_site(B, maps[n], maps[(n + 1) % 2], d)  # B maps bond n to bond n + 1 (charges shifted by K)
        Cd = _dense_bm(s["C"][n], maps[(n + 1) % 2], maps[(n + 1) % 2])
        G.append(np.einsum("sab,bc->sac", Bd, np.linalg.inv(Cd)))
    k = 2 * np.pi * m / q

    def e(eps):
        T = [AL[n] + eps * np.exp(1j * k * c) * G[n] for c in range(q) for n in range(2)]
        return _energy_per_cell(T, s["H0"], d)

    e0 = e(0.0)

    def f(r):
        return np.real(sum(e(z * r) for z in (1, 1j, -1, -1j)) - 4 * e0) / (4 * r * r)
    r = 1e-3  # Richardson step: the quartic terms (~r^2) and rounding (~1e-16 / r^2) stay below 1e-9 here
    return (4 * f(r / 2) - f(r)) / 3, e0


# A small gapped ground state whose bond matrices are well conditioned (smallest Schmidt value 0.03;
# the tangent map G = B C^-1 amplifies by its inverse): XXZ Delta = -1.5 (reference convention),
# bond 0 sectors [1, 2, 1], converged to F < 1e-12.
HESS_STATE = (-1.5, (1, 2, 1), 7)
# cases (K, p, q, m): momentum k = 2 pi m / q per cell, the reference's p = k / (pi N) = m / q
HESS_CASES = [((2,), 0.0, 1, 0), ((2,), 0.5, 2, 1), ((2,), 0.25, 4, 1), ((-2,), 0.25, 4, 1), ((0,), 0.0, 1, 0), ((0,), 0.5, 2, 1),
              ((0,), 0.25, 4, 1)]


@functools.lru_cache(maxsize=None)
def _hess_reference(K, p, q, m, seed):
    s = es.state(*HESS_STATE)
    xd = ex.xdims(s["NL"], s["AR"], K)
    n = sum(r * c for dv in xd for *_, r, c in dv)
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z /= np.linalg.norm(z)
    h, e0 = _hessian_form(s, ex.from_vec_list(z, xd), K, q, m)
    return z, h, e0


def _oracle_form(K, p, z, omit=()):
    s = es.state(*HESS_STATE)
    xd = ex.xdims(s["NL"], s["AR"], K)
    proj = abs(p) < 1e-13 and all(k == 0 for k in K)
    LM, RM = (None, None)
    if proj:
        LM, RM, _ = ex.lr_overlaps(s["AL"], s["AR"], K)
    kfac = np.exp(1j * p * np.pi * s["N"])
    y = ex.vectorize_list(ex.apply_heff(ex.from_vec_list(z, xd), kfac, s["AL"], s["AR"], s["NL"], LM, RM, s["H0"], s["HLtot"],
                                        s["HRtot"], omit=omit))
    return np.vdot(z, y)


@pytest.mark.parametrize("K,p,q,m", HESS_CASES)
def test_heff_quadratic_form_is_energy_hessian(K, p, q, m):
    z, h, e0 = _hess_reference(K, p, q, m, 11)
    assert abs(e0) < 1e-12  # E0 subtracted: the ground state has zero energy per cell
    got = _oracle_form(K, p, z)
    assert abs(got.imag) < 1e-12
    assert abs(got.real - h) < 1e-8 * max(1.0, abs(h)), (got, h)


# ---- the reference's own consistency checks ---------------------------------------------------------
@pytest.mark.parametrize("st", [None, HESS_STATE], ids=["D16", "D4"])
def test_heff_constants_traces_vanish(st):
    s = es.state() if st is None else es.state(*st)
    HL, HR = ex.heff_bond_sums(s["AL"], s["AR"], s["H0"])
    # the reference warns above 1e-14 (it only prints); E0 of these states, converged to F < 1e-12,
    # is good to about 1e-13
    assert abs(ex.tr_prod(HL[-1], s["R"])) < 1e-12
    assert abs(ex.tr_prod(s["L"], HR[0])) < 1e-12
    # without the subtraction the traces are the cell energy 2 E0
    HLr, HRr = ex.heff_bond_sums(s["AL"], s["AR"], s["H"])
    assert abs(ex.tr_prod(HLr[-1], s["R"]) - 2 * s["E0"]) < 1e-12
    assert abs(ex.tr_prod(s["L"], HRr[0]) - 2 * s["E0"]) < 1e-12


def test_lr_overlaps_closed_form():
    s = es.state()
    LM, RM, ol = ex.lr_overlaps(s["AL"], s["AR"], (0,))
    assert abs(abs(ol) - 1.0) < 1e-12
    Cb = s["C"][-1]
    # eigenmatrices of the mixed TM at K = 0: C^T (left) and C (right), up to one common factor
    a = ex.vectorize(LM) @ ex.vectorize(ex.bm_t(Cb)) / (ex.vectorize(ex.bm_t(Cb)) @ ex.vectorize(ex.bm_t(Cb)))
    assert np.abs(ex.vectorize(LM) - a * ex.vectorize(ex.bm_t(Cb))).max() < 1e-11
    b = ex.vectorize(RM) @ ex.vectorize(Cb) / (ex.vectorize(Cb) @ ex.vectorize(Cb))
    assert np.abs(ex.vectorize(RM) - b * ex.vectorize(Cb)).max() < 1e-11
    assert abs(a * b * ex.tr_prod(ex.bm_t(Cb), Cb) - 1.0) < 1e-12
    # left and right dominant overlaps are the same real number (LRoverlaps' check)
    for K in ((0,), (2,)):
        wl = ex.tm_eigs_dense(s["AL"], s["AR"], 0, tuple(-k for k in K))[0]
        wr = ex.tm_eigs_dense(s["AL"], s["AR"], 1, K)[0]
        assert abs(wl - wr) < 1e-11 and abs(np.imag(wl)) < 1e-12


# ---- every term is seen by a pin ----------------------------------------------------------------------
def _pins_pass(omit):
    for K, p, q, m in HESS_CASES:
        z, h, _ = _hess_reference(K, p, q, m, 11)
        got = _oracle_form(K, p, z, omit)
        if abs(got - h) > 1e-8 * max(1.0, abs(h)):
            return False
    J, Delta, hz = 1.0, 1.3, 0.3
    for p in (0.17, 0.8):
        M = _heff_dense(*_magnon_case(J, Delta, hz), p, omit)
        w = np.sort(np.linalg.eigvals(M).real)
        ref = np.sort([Delta + hz - J * np.cos(p * np.pi), Delta + hz - J * np.cos(p * np.pi + np.pi)])
        if np.abs(M - M.conj().T).max() > 1e-14 or np.abs(w - ref).max() > 1e-13:
            return False
    return True


@pytest.mark.parametrize("term", ex.HEFF_TERMS)
def test_pins_notice_a_dropped_term(term):
    assert _pins_pass(())
    assert not _pins_pass((term,)), f"no pin notices the loss of {term}"
