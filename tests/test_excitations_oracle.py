"""Pins of the complex excitation oracle (oracle/excitations.py) on the CPU.

  * its block algebra against the pinned C++ oracle on real data (mixed TMs with and without X,
    MPS products, ApplyOperator, operator TMs);
  * InvertE_fac against the explicit geometric series sum_m (fac T)^m x where it converges
    (charge K != 0: the mixed TM has spectral radius < 1 there);
  * InvertE_proj_fac at fac = 1 with A = B against the C++ oracle's InvertE_proj (the identity the
    reference checks in Excitations_helpers.hpp:855-877);
  * Hermiticity of the assembled H_eff(k) and positivity of its spectrum (E0 subtracted).
"""
import numpy as np
import pytest

from oracle import excitations as ex
from oracle import pyoracle as orc
import excitation_state as es


def _ob_vec(ob):
    return np.asarray(ob.data)


def test_block_algebra_matches_cpp_oracle():
    s = es.state()
    AL, AR = s["AL"], s["AR"]
    ALob, ARob = s["ALob"], s["ARob"]
    rng = np.random.default_rng(5)
    # a random BlockMat on bond 0 with charge shift 0 (the structure of C)
    Cob = s["Cob"][-1]
    Xob = orc.OB.like(Cob, rng.standard_normal(Cob.nelem))
    # mixed TMs with X (mixed left: sum_s B_s^T X A_s with X on bond 0; right: sum_s A_s X B_s^T
    # with X on bond 1)
    C1 = s["Cob"][0]
    X1ob = orc.OB.like(C1, rng.standard_normal(C1.nelem))
    for direction, fn, xo in ((0, ex.tml, Xob), (1, ex.tmr, X1ob)):
        A, B = (ALob[0], ARob[0])
        got = fn(ex.from_ob(A), ex.from_ob(B), ex.from_ob(xo)[0])
        assert got
        lay = ex.to_ob([got], 1)
        ref = orc.tm_mixed(direction, A, B, xo, lay)
        assert np.abs(ex.vectorize(got) - _ob_vec(ref)).max() < 1e-13
    # MPS products
    got = ex.mps_mat(AL[0], ex.from_ob(X1ob)[0])
    ref = orc.mps_mat(ALob[0], X1ob)
    assert got[0]
    assert np.abs(np.concatenate([ex.vectorize(b) for b in got]) - _ob_vec(ref)).max() < 1e-13
    got = ex.mps_mps(AL[0], AL[1])
    ref = orc.mps_mul(ALob[0], ALob[1])
    assert np.abs(np.concatenate([ex.vectorize(b) for b in got]) - _ob_vec(ref)).max() < 1e-13
    # ApplyOperator and the diagonal operator TMs
    H = s["H"]
    got = ex.apply_operator(got, H)
    ref = orc.apply_operator(ref, H)
    assert np.abs(np.concatenate([ex.vectorize(b) for b in got]) - _ob_vec(ref)).max() < 1e-13
    AA = orc.mps_mul(ALob[1], ALob[0])
    got = ex.op_tml_diag(H, ex.from_ob(AA))
    ref = orc.op_tm_diag(0, H, 2, AA)
    assert np.abs(ex.vectorize(got) - _ob_vec(ref)).max() < 1e-13


def test_invert_e_fac_is_the_geometric_series():
    s = es.state()
    AL, AR, NL = s["AL"], s["AR"], s["NL"]
    K = (2,)
    xd = ex.xdims(NL, AR, K)
    rng = np.random.default_rng(6)
    # an ABR-type right vector in sector K: tmr_unity(NL X, AR)
    X = ex.from_vec_list(rng.standard_normal(sum(r * c for dv in xd for *_, r, c in dv)) + 0j, xd)
    x = ex.tmr_unity(ex.mps_mat(NL[0], X[0]), AR[0])
    fac = np.exp(1j * 0.37 * np.pi * 2)
    y, flag = ex.invert_e_fac(AL, AR, x, 1, fac, 1e-14)
    assert flag > 0
    # sum_m (fac T)^m x on the sector list of x
    dv = ex.dims(x)
    term, total = x, ex.vectorize(x).astype(complex)
    for m in range(2000):
        term = ex.from_vec(fac * ex._map_onto(ex.tmr_chain(AL, AR, term), dv), dv)
        v = ex.vectorize(term)
        total = total + v
        if np.linalg.norm(v) < 1e-17 * np.linalg.norm(total):
            break
    assert np.linalg.norm(ex.vectorize(y) - total) / np.linalg.norm(total) < 1e-11


def test_invert_e_proj_fac_reduces_to_invert_e_proj():
    s = es.state()
    rng = np.random.default_rng(7)
    R = s["R"]
    Rob = ex.to_ob([R], 1)
    xob = orc.OB.like(Rob, rng.standard_normal(Rob.nelem))
    one = ex.eye(R)
    # 1e-13: on this 44-dimensional system the unrestarted GMRES of both oracles reaches its full
    # Krylov dimension at 1e-14 without meeting the round-off-limited residual test
    y1, flag = ex.invert_e_proj_fac(s["AL"], s["AL"], ex.from_ob(xob)[0], one, R, 0, 1.0, 1e-13)
    assert flag > 0
    y2, _ = orc.invert_e_proj(0, s["ALob"], xob, ex.to_ob([one], 1), Rob, tol=1e-13)
    assert np.linalg.norm(ex.vectorize(y1) - _ob_vec(y2)) / np.linalg.norm(_ob_vec(y2)) < 1e-10


@pytest.mark.parametrize("K,p", [((2,), 0.3), ((0,), 0.0), ((0,), 0.5)])
def test_heff_hermitian_and_positive(K, p):
    M = es.oracle_heff_matrix(K, p)
    assert np.abs(M - M.conj().T).max() <= 1e-10 * np.abs(M).max()
    w = np.linalg.eigvalsh(0.5 * (M + M.conj().T))
    assert w[0] > 0  # excitations lie above the ground state (E0 subtracted)


def test_tm_spectrum_dense_fixed_point():
    """The dense TM spectrum has the isometries' fixed-point eigenvalue 1 (left-normalised AL, left
    and right chains alike) and the mixed (AL, AR) TM of the converged state has it as well."""
    s = es.state()
    for direction in (0, 1):
        w = ex.tm_eigs_dense(s["AL"], s["AL"], direction, (0,))
        assert abs(w[0] - 1.0) < 1e-12 and abs(w[1]) < 1.0
        wm = ex.tm_eigs_dense(s["AL"], s["AR"], direction, (0,))
        assert abs(abs(wm[0]) - 1.0) < 1e-10
