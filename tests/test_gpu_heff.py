"""The C-ABI excitation engine (imps_heff_*) on the GPU against the oracle.

States (oracle VUMPS, CPU): the converged D = 16 anisotropic Heisenberg chain of excitation_state.py
(U(1), d = 2) and a D = 16 Fermi-Hubbard U(1)xU(1) state over 16 sectors (d = 4).  The engine gets
the same AL, AR, C and null spaces NL, computes its own constants and applies H_eff(k) for several
momenta in one batched call.  At K = 0 the geometric sums are projected wherever e^{i p pi N} = 1
(the reference projects at p = 0 only and leaves p pi N = 2 pi, where 1 - T is just as singular, to
an ill-conditioned solve).  Tolerances (FP64 throughout, geometric sums solved to 1e-14):
H_eff X relative error 1e-10 per momentum; the lowest four energies 1e-9 absolute against a dense
eigensolve of the oracle's H_eff(k); Hermiticity 1e-11."""
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


_CTX = []
_PLANS = {}


def ctx():
    if not _CTX:
        _CTX.append(im.Context(0))
    return _CTX[0]


def plan(which, K, nmom, prec=im.FP64, inv_tol=1e-14):
    key = (which, K, nmom, prec, inv_tol)
    if key not in _PLANS:
        s = es.state() if which == "xxz" else es.hubbard_state()
        print(f"heff plan {which} K={K} nmom={nmom} prec={prec}", flush=True)
        P = im.HeffPlan(ctx(), [_lay(a) for a in s["ALob"]], [_lay(a) for a in s["ARob"]], [_lay(a) for a in s["NLob"]],
                        [_lay(c) for c in s["Cob"]], s["H"], s["E0"], K, nmom, prec=prec)
        P.set_state([_dev(a) for a in s["ALob"]], [_dev(a) for a in s["ARob"]], [_dev(a) for a in s["NLob"]], [_dev(c) for c in s["Cob"]],
                    inv_tol=inv_tol)
        _PLANS[key] = (P, s)
    return _PLANS[key]


CASES = [("xxz", (0,)), ("xxz", (2,)), ("hub", (0, 0)), ("hub", (1, 1)), ("hub", (0, 2))]
MOMENTA = [0.0, 0.3, 1.0, 0.5]


@pytest.mark.parametrize("which,K", CASES)
def test_xdims_and_constants_check(which, K):
    P, s = plan(which, K, len(MOMENTA))
    from oracle import excitations as ex
    xd = ex.xdims(s["NL"], s["AR"], K)
    ref = [(q, qo, r, c) for dv in xd for (q, qo, r, c) in dv]
    got = []
    b = 0
    for n, nb in enumerate(P.xdims.nblk):
        for _ in range(int(nb)):
            got.append((tuple(int(v) for v in P.xdims.qin[b]), tuple(int(v) for v in P.xdims.qout[b]), int(P.xdims.rows[b]),
                        int(P.xdims.cols[b])))
            b += 1
    assert got == ref
    trl, trr, its = P.check()
    print(f"{which} K={K}: tr(HL R) {trl:.2e} tr(L HR) {trr:.2e} env applications {its}", flush=True)
    assert abs(trl) < 1e-12 and abs(trr) < 1e-12  # the reference's HeffConstants check (it warns above 1e-14)


@pytest.mark.parametrize("which,K", CASES)
def test_apply_parity_batched_momenta(which, K):
    P, s = plan(which, K, len(MOMENTA))
    rng = np.random.default_rng(21)
    z = rng.standard_normal((len(MOMENTA), P.mtot)) + 1j * rng.standard_normal((len(MOMENTA), P.mtot))
    x = torch.from_numpy(z).cuda()
    y = torch.empty_like(x)
    steps, rr = P.apply(MOMENTA, x, y)
    got = y.cpu().numpy()
    for t, p in enumerate(MOMENTA):
        ref = es.oracle_heff(K, p, z[t], s, proj_at_unit_phase=True)
        err = np.linalg.norm(got[t] - ref) / np.linalg.norm(ref)
        print(f"{which} K={K} p={p}: relerr {err:.2e} (gmres {steps} steps, relres {rr:.1e})", flush=True)
        assert err < 1e-10


@pytest.mark.parametrize("which,K", [("xxz", (2,)), ("hub", (0, 0))])
def test_apply_hermitian(which, K):
    P, s = plan(which, K, len(MOMENTA))
    rng = np.random.default_rng(22)
    shape = (len(MOMENTA), P.mtot)
    x = torch.from_numpy(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).cuda()
    w = torch.from_numpy(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).cuda()
    hx, hw = torch.empty_like(x), torch.empty_like(x)
    P.apply(MOMENTA, x, hx)
    P.apply(MOMENTA, w, hw)
    for t in range(len(MOMENTA)):
        a, b = torch.vdot(w[t], hx[t]).item(), torch.vdot(x[t], hw[t]).item()
        assert abs(a - b.conjugate()) <= 1e-11 * abs(a)


def test_apply_repeatable():
    """Two applies of the same vectors give bitwise equal results (fixed reduction orders)."""
    P, s = plan("hub", (1, 1), len(MOMENTA))
    rng = np.random.default_rng(23)
    x = torch.from_numpy(rng.standard_normal((len(MOMENTA), P.mtot)) + 1j * rng.standard_normal((len(MOMENTA), P.mtot))).cuda()
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    P.apply(MOMENTA, x, y1)
    P.apply(MOMENTA, x, y2)
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("which,K", [("xxz", (0,)), ("xxz", (2,)), ("hub", (0, 0)), ("hub", (1, 1))])
def test_lowest_energies_batched(which, K):
    moms = [0.0, 0.3]
    P, s = plan(which, K, len(moms))
    ev, vec, res, na = P.eigs(moms, nev=4, tol=1e-10, vectors=True)
    for t, p in enumerate(moms):
        M = es.oracle_heff_matrix(K, p, s, proj_at_unit_phase=True)
        ref = np.linalg.eigvalsh(0.5 * (M + M.conj().T))[:4]
        print(f"{which} K={K} p={p}: {ev[t]} vs {ref} ({na} batched applies, res {res[t].max():.1e})", flush=True)
        assert np.abs(ev[t] - ref).max() < 1e-9
        V = vec[t].cpu().numpy()
        # Ritz vectors: orthonormal and eigenvectors of the oracle's H_eff
        assert np.abs(V.conj() @ V.T - np.eye(4)).max() < 1e-10
        R = M @ V.T - V.T * ev[t]
        assert np.abs(R).max() < 1e-8


@pytest.mark.parametrize("prec", [im.FP32, im.BF16X3], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("which,K", [("xxz", (2,)), ("hub", (0, 0)), ("hub", (1, 1))])
def test_mixed_apply_refined_to_fp64(which, K, prec):
    """Low-precision TM chains in the geometric sums' Arnoldi steps, FP64 residual at every restart:
    the sums reach the FP64 tolerance (1e-12 here), so H_eff X matches the oracle as the FP64 path does."""
    P, s = plan(which, K, len(MOMENTA), prec=prec, inv_tol=1e-12)
    rng = np.random.default_rng(24)
    z = rng.standard_normal((len(MOMENTA), P.mtot)) + 1j * rng.standard_normal((len(MOMENTA), P.mtot))
    x = torch.from_numpy(z).cuda()
    y = torch.empty_like(x)
    steps, rr = P.apply(MOMENTA, x, y)
    got = y.cpu().numpy()
    assert rr < 1e-12
    for t, p in enumerate(MOMENTA):
        ref = es.oracle_heff(K, p, z[t], s, proj_at_unit_phase=True)
        err = np.linalg.norm(got[t] - ref) / np.linalg.norm(ref)
        print(f"{which} K={K} p={p} prec={prec}: relerr {err:.2e} (gmres {steps} steps, relres {rr:.1e})", flush=True)
        assert err < 1e-10


@pytest.mark.parametrize("which,K", [("xxz", (0,)), ("hub", (0, 0))])
def test_mixed_lowest_energies_equal_fp64(which, K):
    moms = [0.0, 0.3]
    P64, s = plan(which, K, len(moms))
    P32, _ = plan(which, K, len(moms), prec=im.FP32, inv_tol=1e-14)
    e64, _, _, _ = P64.eigs(moms, nev=4, tol=1e-10)
    e32, _, _, na = P32.eigs(moms, nev=4, tol=1e-10)
    print(f"{which} K={K}: FP32-refined {e32.ravel()} vs FP64 {e64.ravel()} ({na} applies)", flush=True)
    assert np.abs(e32 - e64).max() < 1e-9
