"""Quasiparticle excitations on the GPU against the oracle (ApplyHeff, HeffConstants, eigenvalues).

State: the converged D = 16 anisotropic Heisenberg chain of excitation_state.py (oracle VUMPS).
The GPU engine gets the same AL, AR, C and null spaces NL, computes its own constants (reduced
operators, HLtot / HRtot through the library's environment solver) and applies H_eff(k) through
the library's block products.  Tolerances: H_eff X relative error 1e-10 (FP64 throughout; the
geometric sums are solved to 1e-14); the lowest four energies 1e-9 absolute against a dense
eigensolve of the oracle's H_eff(k)."""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
import excitation_state as es

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def _dev(ob):
    return torch.from_numpy(np.ascontiguousarray(ob.data, np.float64)).cuda()


_ENG = {}


def _new_engine(K, **kw):
    from imps_cpp_tools_amd.excitations import ExcitationEngine
    s = es.state()
    return ExcitationEngine(im.Context(0), [_dev(a) for a in s["ALob"]], [_lay(a) for a in s["ALob"]], [_dev(a) for a in s["ARob"]],
                            [_lay(a) for a in s["ARob"]], [_dev(a) for a in s["NLob"]], [_lay(a) for a in s["NLob"]],
                            [_dev(c) for c in s["Cob"]], [_lay(c) for c in s["Cob"]], s["H"], s["d"], s["E0"], K, **kw)


def engine(K):
    """The shared engine of a charge."""
    if K not in _ENG:
        _ENG[K] = _new_engine(K)
    return _ENG[K]


@pytest.mark.parametrize("K", [(0,), (2,)], ids=["K0", "K2"])
def test_heff_constants(K):
    s = es.state()
    eng = engine(K)
    from oracle import excitations as ex
    for n in range(s["N"]):
        for got, ref in ((eng.HLtot[n], s["HLtot"][n]), (eng.HRtot[n], s["HRtot"][n])):
            r = ex.vectorize(ref)  # same sector order on both sides (ascending charge)
            g = got.data.cpu().numpy()
            assert g.shape == r.shape
            assert np.abs(g - r).max() <= 1e-11 * max(1.0, np.abs(r).max())


@pytest.mark.parametrize("K", [(0,), (2,)], ids=["K0", "K2"])
@pytest.mark.parametrize("p", [0.0, 0.3, 1.0])
def test_apply_heff_parity(K, p):
    eng = engine(K)
    rng = np.random.default_rng(11)
    z = rng.standard_normal(eng.mtot) + 1j * rng.standard_normal(eng.mtot)
    ref = es.oracle_heff(K, p, z)
    got = eng.matvec(torch.from_numpy(z).cuda(), p).cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"K {K} p {p}: relerr {err:.2e}", flush=True)
    assert err < 1e-10


@pytest.mark.parametrize("K", [(0,), (2,)], ids=["K0", "K2"])
def test_heff_hermitian(K):
    """<y, H x> = conj(<x, H y>) for H_eff(k) at a generic momentum."""
    eng = engine(K)
    rng = np.random.default_rng(12)
    x = torch.from_numpy(rng.standard_normal(eng.mtot) + 1j * rng.standard_normal(eng.mtot)).cuda()
    y = torch.from_numpy(rng.standard_normal(eng.mtot) + 1j * rng.standard_normal(eng.mtot)).cuda()
    hx, hy = eng.matvec(x, 0.37), eng.matvec(y, 0.37)
    a, b = torch.vdot(y, hx).item(), torch.vdot(x, hy).item()
    assert abs(a - b.conjugate()) <= 1e-11 * abs(a)


@pytest.mark.parametrize("K", [(0,), (2,)], ids=["K0", "K2"])
@pytest.mark.parametrize("p", [0.0, 0.3])
def test_excitation_energies(K, p):
    eng = engine(K)
    M = es.oracle_heff_matrix(K, p)
    ref = np.linalg.eigvalsh(0.5 * (M + M.conj().T))[:4]
    w, Y, nmv = eng.lowest(p, nev=4, tol=1e-10, krylov=30)
    print(f"K {K} p {p}: {w} ({nmv} applies)", flush=True)
    assert np.abs(w - ref).max() < 1e-9


@pytest.mark.parametrize("K", [(0,), (2,)], ids=["K0", "K2"])
def test_recycled_sums(K):
    """recycle=True (each geometric sum starts from the previous apply's solution, as the
    reference's excitation program does): applies to new vectors keep the oracle parity and the
    lowest energies stay those of the dense H_eff(k)."""
    eng = _new_engine(K, recycle=True)
    rng = np.random.default_rng(14)
    for p in (0.3, 0.3, 0.0):
        z = rng.standard_normal(eng.mtot) + 1j * rng.standard_normal(eng.mtot)
        ref = es.oracle_heff(K, p, z)
        got = eng.matvec(torch.from_numpy(z).cuda(), p).cpu().numpy()
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-10, p
    M = es.oracle_heff_matrix(K, 0.3)
    ref = np.linalg.eigvalsh(0.5 * (M + M.conj().T))[:4]
    w, _, _ = eng.lowest(0.3, nev=4, tol=1e-10, krylov=30)
    assert np.abs(w - ref).max() < 1e-9


@pytest.mark.parametrize("mixed", [False, True], ids=["AL", "AL_AR"])
@pytest.mark.parametrize("direction", [0, 1], ids=["l", "r"])
@pytest.mark.parametrize("K", [(0,), (2,)], ids=["K0", "K2"])
def test_tm_eigs(mixed, direction, K):
    """Largest-magnitude TM eigenvalues in a charge sector (TMEigs / TMmixedEigs) against a dense
    eigensolve of the oracle's transfer matrix."""
    from oracle import excitations as ex
    s = es.state()
    eng = engine((0,))
    As = eng.AL
    Bs = eng.AR if mixed else eng.AL
    ref = ex.tm_eigs_dense(s["AL"], s["AR"] if mixed else s["AL"], direction, K)
    w, Y, nmv = eng.tm_eigs(As, Bs, direction, K, nev=4, tol=1e-12, krylov=30)
    assert np.allclose(np.abs(w), np.abs(ref[:4]), rtol=0, atol=1e-10)
    assert max(min(abs(x - ref)) for x in w) < 1e-10
    if not mixed and K == (0,):
        assert abs(w[0] - 1.0) < 1e-12  # the isometry's fixed point


def test_energies_with_gpu_nullspace():
    """The excitation energies do not depend on the null-space basis: the engine built on the GPU's
    Nullspace (a different orthonormal basis of the same space) gives the dense oracle's energies."""
    from imps_cpp_tools_amd.excitations import ExcitationEngine
    from imps_cpp_tools_amd.nullspace import left_nullspace
    s = es.state()
    ctx = im.Context(0)
    AL = [_dev(a) for a in s["ALob"]]
    NLs = [left_nullspace(ctx, a, _lay(ob), seed=9) for a, ob in zip(AL, s["ALob"])]
    for (nl, nll), ob in zip(NLs, s["NLob"]):
        assert nll == _lay(ob)
    K, p = (2,), 0.3
    eng = ExcitationEngine(ctx, AL, [_lay(a) for a in s["ALob"]], [_dev(a) for a in s["ARob"]], [_lay(a) for a in s["ARob"]],
                           [a for a, _ in NLs], [l for _, l in NLs], [_dev(c) for c in s["Cob"]], [_lay(c) for c in s["Cob"]],
                           s["H"], s["d"], s["E0"], K)
    M = es.oracle_heff_matrix(K, p)
    ref = np.linalg.eigvalsh(0.5 * (M + M.conj().T))[:4]
    w, _, _ = eng.lowest(p, nev=4, tol=1e-10, krylov=30)
    assert np.abs(w - ref).max() < 1e-9
