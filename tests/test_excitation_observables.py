"""Pins of the oracle's excitation observables (OPeffConstants, ApplyOPeff, MeasureExcitations;
Excitations_helpers.hpp:350-600, 774-826), on the CPU.

1. Closed forms at bond dimension 1: on the fully polarised XXZ chain one magnon lowers the total
   Sz by one, so (Sz - <Sz>)_eff(k) = -1 at every momentum; on the polarised Hubbard chain (one up
   electron per site) a hole changes n by -1 and nup by -1, a doublon n and ndo by +1.
2. The quadratic form <X|O_eff(k)|X> against the second derivative of the observable's density in
   the perturbed uniform state (the construction of the H_eff pin with a one-site operator; dense
   symmetry-free transfer matrices, theta-averaging over {r, ir, -r, -ir} removes the linear and
   holomorphic parts).  Observable: the double occupancy of a converged D = 4 Hubbard state (U(1) x
   U(1), smallest Schmidt value 0.011) at K in {(0,0), (1,1), (-1,1), (0,2), (2,0)} and k in {0, pi/2,
   pi} per cell; every one of ApplyOPeff's terms changes at least one of these forms by 1e-4 or more,
   and the pin fails when any single term is dropped.  (A conserved charge is a weak probe: for Sz on
   the XXZ chain the form is exactly the excitation's charge, and several terms vanish identically.)
3. OPeffConstants' checks: tr(OPL R) = tr(L OPR) = 0 with the expectation subtracted, and the cell's
   expectation N <O> without; O_eff(k) Hermitian.
"""
import functools

import numpy as np
import pytest

from oracle import excitations as ex
import excitation_state as es
import test_excitations_pins as tp

SZ = np.diag([0.5, -0.5])


def _opeff_dense(AL, AR, NL, C, O, K, p, L=None, R=None):
    L = C[-1] if L is None else L
    R = C[-1] if R is None else R
    OPLtot, OPRtot, _ = ex.opeff_constants(AL, AR, L, R, O)
    xd = ex.xdims(NL, AR, K)
    n = sum(r * c for dv in xd for *_, r, c in dv)
    kfac = np.exp(1j * p * np.pi * len(AL))
    proj = abs(kfac - 1) < 1e-13 and all(k == 0 for k in K)
    LM, RM = (ex.lr_overlaps(AL, AR, K)[:2]) if proj else (None, None)
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        M[:, j] = ex.vectorize_list(ex.apply_opeff(ex.from_vec_list(e, xd), kfac, AL, AR, NL, LM, RM, O, OPLtot, OPRtot))
    return M


@pytest.mark.parametrize("p", [0.0, 0.17, 0.5, 1.0])
def test_magnon_changes_sz_by_one(p):
    AL, AR, NL, C, _, K = tp._magnon_case()
    M = _opeff_dense(AL, AR, NL, C, SZ - 0.5 * np.eye(2), K, p)
    assert M.shape == (2, 2)
    np.testing.assert_allclose(M, -np.eye(2), atol=1e-13)


@pytest.mark.parametrize("K,dn,dup,ddo", [((1, 1), -1, -1, 0), ((-1, 1), 1, 0, 1)], ids=["hole", "doublon"])
@pytest.mark.parametrize("p", [0.0, 0.23, 1.0])
def test_hubbard_hole_and_doublon_charges(K, dn, dup, ddo, p):
    AL, AR, NL, C, _, K = tp._hubbard_case(K)
    for diag, ref in (([0, 1, 1, 2], dn), ([0, 0, 1, 1], dup), ([0, 1, 0, 1], ddo)):
        O = np.diag(np.asarray(diag, float))
        O = O - O[2, 2] * np.eye(4)  # <O> of the polarised state (one up electron, local state 2)
        M = _opeff_dense(AL, AR, NL, C, O, K, p)
        np.testing.assert_allclose(M, ref * np.eye(M.shape[0]), atol=1e-13)


# ---- the quadratic form as a second derivative ---------------------------------------------------
def _site_op_tm(T, O):
    d = T.shape[0]
    return sum(np.kron(sum(O[s2, s] * T[s] for s in range(d)), T[s2].conj()) for s2 in range(d))


def _obs_per_cell(T, O, ncell_sites=2):
    """Sum of <O> over one two-site cell of the uniform state with periodic site tensors T."""
    L = len(T)
    M = [tp._site_tm(t) for t in T]
    E = functools.reduce(np.matmul, M)
    w, V = np.linalg.eig(E)
    i = int(np.argmax(np.abs(w)))
    lam, r = w[i], V[:, i]
    wl, Vl = np.linalg.eig(E.T)
    l = Vl[:, int(np.argmax(np.abs(wl)))]
    norm = l @ r
    tot = 0.0
    eye = np.eye(M[0].shape[0])
    for b in range(L):
        left = functools.reduce(np.matmul, M[:b], eye)
        right = functools.reduce(np.matmul, M[b + 1:], np.eye(M[-1].shape[1]))
        tot += (l @ left @ _site_op_tm(T[b], O) @ right @ r) / (lam * norm)
    return tot / (L // ncell_sites)


def _obs_hessian(s, X, O, q, m):
    d = s["d"]
    maps = tp._bond_maps(s)
    AL = [tp._dense_site(s["AL"][n], maps[n], maps[(n + 1) % 2], d) for n in range(2)]
    G = []
    for n in range(2):
        Bd = tp._dense_site(ex.mps_mat(s["NL"][n], X[n]), maps[n], maps[(n + 1) % 2], d)
        Cd = tp._dense_bm(s["C"][n], maps[(n + 1) % 2], maps[(n + 1) % 2])
        G.append(np.einsum("sab,bc->sac", Bd, np.linalg.inv(Cd)))
    k = 2 * np.pi * m / q

    def e(eps):
        return _obs_per_cell([AL[n] + eps * np.exp(1j * k * c) * G[n] for c in range(q) for n in range(2)], O)

    e0 = e(0.0)

    def f(r):
        return np.real(sum(e(z * r) for z in (1, 1j, -1, -1j)) - 4 * e0) / (4 * r * r)
    r = 1e-3
    return (4 * f(r / 2) - f(r)) / 3, e0


def _expectation(s, O):
    """Mean of <O> over the cell's sites from AL and R (the reference's MeasureObservables on AL)."""
    OPL, _, (tl, _) = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], O)
    return np.real(tl) / s["N"]


DOCC = np.diag([0.0, 0.0, 0.0, 1.0])
OBS_STATE = (2, 1, 5, 4.0)  # hubbard_state(nside, peak, seed, U): D = 4 over 4 sectors
OBS_CASES = [(K, p, q, m) for K in ((0, 0), (1, 1), (-1, 1), (0, 2), (2, 0)) for (p, q, m) in ((0.0, 1, 0), (0.5, 2, 1), (0.25, 4, 1))]


@functools.lru_cache(maxsize=None)
def _obs_state():
    s = es.hubbard_state(*OBS_STATE)
    O = DOCC - _expectation(s, DOCC) * np.eye(4)
    return s, O


_HC = {}


def _hess_case(K, p, q, m):
    if (K, p, q, m) in _HC:
        return _HC[(K, p, q, m)]
    s, O = _obs_state()
    xd = ex.xdims(s["NL"], s["AR"], K)
    n = sum(r * c for dv in xd for *_, r, c in dv)
    rng = np.random.default_rng(12)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    z /= np.linalg.norm(z)
    X = ex.from_vec_list(z, xd)
    h, e0 = _obs_hessian(s, X, O, q, m)
    OPLtot, OPRtot, _ = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], O)
    proj = abs(p) < 1e-13 and all(k == 0 for k in K)
    LM, RM = (ex.lr_overlaps(s["AL"], s["AR"], K)[:2]) if proj else (None, None)
    _HC[(K, p, q, m)] = (s, O, X, h, e0, OPLtot, OPRtot, LM, RM)
    return _HC[(K, p, q, m)]


def _form(case, omit=()):
    K, p, q, m = case
    s, O, X, h, e0, OPLtot, OPRtot, LM, RM = _hess_case(*case)
    return ex.measure_excitation(X, np.exp(1j * p * np.pi * 2), s["AL"], s["AR"], s["NL"], LM, RM, O, OPLtot, OPRtot, omit=omit).sum()


@pytest.mark.parametrize("K,p,q,m", OBS_CASES)
def test_opeff_quadratic_form_is_observable_hessian(K, p, q, m):
    s, O, X, h, e0, *_ = _hess_case(K, p, q, m)
    assert abs(e0) < 1e-12  # expectation subtracted
    got = _form((K, p, q, m))
    assert abs(got.imag) < 1e-12
    assert abs(got.real - h) < 1e-8 * max(1.0, abs(h)), (got, h)


@pytest.mark.parametrize("K", [(0,), (2,), (-2,)])
def test_conserved_charge_form_is_the_excitation_charge(K):
    """Sz - <Sz> on the XXZ chain: <X|O_eff(k)|X> = -K/2 |X|^2 (K counts 2 Sz, with the sign of the
    charge convention), at every momentum."""
    s = es.state(*tp.HESS_STATE)
    O = SZ - _expectation(s, SZ) * np.eye(2)
    M = _opeff_dense(s["AL"], s["AR"], s["NL"], s["C"], O, K, 0.3, s["L"], s["R"])
    np.testing.assert_allclose(M, -0.5 * K[0] * np.eye(M.shape[0]), atol=1e-10)


def test_opeff_constants_checks_and_hermiticity():
    s = es.state()
    ez = _expectation(s, SZ)
    _, _, (tl, tr) = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], SZ)
    assert abs(tl - 2 * ez) < 1e-12 and abs(tr - 2 * ez) < 1e-12  # the cell's expectation without the subtraction
    O = SZ - ez * np.eye(2)
    _, _, (tl, tr) = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], O)
    assert abs(tl) < 1e-12 and abs(tr) < 1e-12
    for K, p in (((0,), 0.3), ((2,), 0.0), ((0,), 0.0)):
        M = _opeff_dense(s["AL"], s["AR"], s["NL"], s["C"], O, K, p, s["L"], s["R"])
        assert np.abs(M - M.conj().T).max() < 1e-11


# ---- every term is seen by a pin -------------------------------------------------------------------
def _obs_pins_pass(omit):
    for case in OBS_CASES:
        h = _hess_case(*case)[3]
        if abs(_form(case, omit) - h) > 1e-8 * max(1.0, abs(h)):
            return False
    return True


@pytest.mark.parametrize("term", ex.OPEFF_TERMS)
def test_observable_pins_notice_a_dropped_term(term):
    assert _obs_pins_pass(())
    assert not _obs_pins_pass((term,)), f"no pin notices the loss of {term}"
