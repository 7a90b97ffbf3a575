"""Excitations between AL and a transformed or shifted AR on the GPU against the oracle.

The product prepares the right half of the excitation program on the device
(imps_cpp_tools_amd.transform.transformed_state: L = C^T C by a library product, ApplyTransformation
as a gather, the relative shift, LRoverlaps by the library's TM Arnoldi, the sign of the overlap
into AR[0]) and hands L, LM, RM to the C-ABI engine (imps_heff_set_state_lr).  The oracle does the
same with its own restatements (tests/excitation_state.py transformed_state: dense eigensolves).
States: the D = 16 XXZ chain at Delta = -1.5 (spin flip FS: overlap 0.70 per cell in K = 0, -0.88
in K = -2; shift by one site: 0.76 in K = -1) and the D = 16 Hubbard state (spin flip FS: -1 in
K = (0, 0), so AR[0] changes sign and the sums at p = 0 are projected; charge flip FC: -2.9e-4 in
K = (2, 0)).  Tolerances: transformed tensors bitwise; L 1e-14; the overlap 1e-10 (relative 1e-5
for the ill-conditioned 2.9e-4); LM and RM 1e-9 (up to their common sign); H_eff X relative error
1e-10 per momentum; the lowest energies 1e-9."""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
import excitation_state as es
from oracle import excitations as ex

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = [("xxz", "FS", (0,), 0), ("xxz", "FS", (-2,), 0), ("xxz", "", (-1,), 1), ("hub", "FS", (0, 0), 0), ("hub", "FC", (2, 0), 0)]
IDS = ["xxz-FS-K0", "xxz-FS-K-2", "xxz-shift1-K-1", "hub-FS-K00", "hub-FC-K20"]
MOMENTA = [0.0, 0.3, 1.0, 0.5]

_CTX = []
_PREP = {}
_PLANS = {}


def ctx():
    if not _CTX:
        _CTX.append(im.Context(0))
    return _CTX[0]


def _lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def _dev(ob):
    return torch.from_numpy(np.ascontiguousarray(ob.data, np.float64)).cuda()


def prepared(which, op, K, shift):
    key = (which, op, K, shift)
    if key not in _PREP:
        from imps_cpp_tools_amd import transform as tr
        s = es.state() if which == "xxz" else es.hubbard_state()
        model, _, QN = es.MODELS[which]
        print(f"transformed state {which} op={op!r} K={K} shift={shift}", flush=True)
        st = tr.transformed_state(ctx(), model, s["N"], QN, [_dev(a) for a in s["ALob"]], [_lay(a) for a in s["ALob"]],
                                  [_dev(a) for a in s["ARob"]], [_lay(a) for a in s["ARob"]], [_dev(c) for c in s["Cob"]],
                                  [_lay(c) for c in s["Cob"]], K, op=op, rel_shift=shift)
        _PREP[key] = (st, s)
    return _PREP[key]


def plan(which, op, K, shift, nmom):
    key = (which, op, K, shift, nmom)
    if key not in _PLANS:
        st, s = prepared(which, op, K, shift)
        P = im.HeffPlan(ctx(), [_lay(a) for a in s["ALob"]], st["ARl"], [_lay(a) for a in s["NLob"]], [_lay(c) for c in s["Cob"]], s["H"],
                        s["E0"], K, nmom)
        P.set_state([_dev(a) for a in s["ALob"]], st["AR"], [_dev(a) for a in s["NLob"]], [_dev(c) for c in s["Cob"]], inv_tol=1e-14,
                    L=st["L"], LM=st["LM"], RM=st["RM"])
        _PLANS[key] = P
    return _PLANS[key]


def _objs(lay, t):
    from oracle import pyoracle as orc
    return ex.from_ob(orc.OB.like(lay, t.cpu().numpy()))


@pytest.mark.parametrize("which,op,K,shift", CASES, ids=IDS)
def test_transformed_state_equals_oracle(which, op, K, shift):
    st, _ = prepared(which, op, K, shift)
    ref = es.transformed_state(which, op, K, shift)
    for n in range(ref["N"]):
        got = _objs(st["ARl"][n], st["AR"][n])
        for a, b in zip(got, ref["AR"][n]):
            assert set(a) == set(b)
            for q in a:
                assert a[q][0] == b[q][0] and np.array_equal(a[q][1], b[q][1])
    Lg = _objs(*st["L"])[0]
    assert set(Lg) == set(ref["L"])
    assert max(np.abs(Lg[q][1] - ref["L"][q][1]).max() for q in Lg) < 1e-14
    print(f"{which} {op} K={K}: OL {st['OL']:.12f} (oracle {ref['OL']:.12f})", flush=True)
    # the charge-flipped Hubbard overlap 2.9e-4 is an eigenvalue of a strongly non-normal TM: there
    # its left and right Arnoldi values already differ by 3e-10, so it is compared relatively
    assert abs(st["OL"] - ref["OL"]) < (1e-10 if abs(ref["OL"]) > 1e-2 else 1e-5 * abs(ref["OL"]))


@pytest.mark.parametrize("which,op,K,shift", [c for c in CASES if all(k == 0 for k in c[2])],
                         ids=[i for c, i in zip(CASES, IDS) if all(k == 0 for k in c[2])])
def test_lr_overlaps_equal_oracle(which, op, K, shift):
    st, _ = prepared(which, op, K, shift)
    ref = es.transformed_state(which, op, K, shift)
    LM, RM, _ = ex.lr_overlaps(ref["AL"], ref["AR"], K)
    lm, rm = _objs(*st["LM"])[0], _objs(*st["RM"])[0]

    def err(a, b, sg):
        return max(np.abs(sg * a[q][1] - b[q][1]).max() for q in b) / max(np.abs(b[q][1]).max() for q in b)
    e = min(max(err(lm, LM, sg), err(rm, RM, sg)) for sg in (1.0, -1.0))
    assert e < 1e-9


@pytest.mark.parametrize("which,op,K,shift", CASES, ids=IDS)
def test_apply_parity_transformed(which, op, K, shift):
    P = plan(which, op, K, shift, len(MOMENTA))
    ref = es.transformed_state(which, op, K, shift)
    trl, trr, _ = P.check()
    assert abs(trl) < 1e-12 and abs(trr) < 1e-12
    rng = np.random.default_rng(31)
    z = rng.standard_normal((len(MOMENTA), P.mtot)) + 1j * rng.standard_normal((len(MOMENTA), P.mtot))
    x = torch.from_numpy(z).cuda()
    y = torch.empty_like(x)
    steps, rr = P.apply(MOMENTA, x, y)
    got = y.cpu().numpy()
    for t, p in enumerate(MOMENTA):
        r = es.oracle_heff(K, p, z[t], ref, proj_at_unit_phase=True)
        e = np.linalg.norm(got[t] - r) / np.linalg.norm(r)
        print(f"{which} {op} K={K} shift={shift} p={p}: relerr {e:.2e} (gmres {steps}, relres {rr:.1e})", flush=True)
        assert e < 1e-10


@pytest.mark.parametrize("which,op,K,shift", [CASES[0], CASES[3]], ids=[IDS[0], IDS[3]])
def test_lowest_energies_transformed(which, op, K, shift):
    moms = [0.0, 0.3]
    P = plan(which, op, K, shift, len(moms))
    ref = es.transformed_state(which, op, K, shift)
    ev, _, rs, na = P.eigs(moms, nev=2, tol=1e-12)
    for t, p in enumerate(moms):
        M = es.oracle_heff_matrix(K, p, ref, proj_at_unit_phase=True)
        w = np.sort(np.linalg.eigvalsh(0.5 * (M + M.conj().T)))[:2]
        print(f"{which} {op} K={K} p={p}: {ev[t]} oracle {w}", flush=True)
        assert np.abs(ev[t] - w).max() < 1e-9


def test_general_bond_input_equals_closed_form():
    """The untransformed state through imps_heff_set_state_lr (L = C^T C by a product, LM and RM by
    the TM Arnoldi) gives the closed-form engine's H_eff."""
    from imps_cpp_tools_amd import transform as tr
    s = es.state()
    K = (0,)
    st = tr.transformed_state(ctx(), 0, s["N"], [0], [_dev(a) for a in s["ALob"]], [_lay(a) for a in s["ALob"]],
                              [_dev(a) for a in s["ARob"]], [_lay(a) for a in s["ARob"]], [_dev(c) for c in s["Cob"]],
                              [_lay(c) for c in s["Cob"]], K)
    assert abs(st["OL"] - 1.0) < 1e-10
    args = ([_lay(a) for a in s["ALob"]], [_lay(a) for a in s["ARob"]], [_lay(a) for a in s["NLob"]], [_lay(c) for c in s["Cob"]], s["H"],
            s["E0"], K, len(MOMENTA))
    A, B = im.HeffPlan(ctx(), *args), im.HeffPlan(ctx(), *args)
    tensors = ([_dev(a) for a in s["ALob"]], [_dev(a) for a in s["ARob"]], [_dev(a) for a in s["NLob"]], [_dev(c) for c in s["Cob"]])
    A.set_state(*tensors, inv_tol=1e-14)
    B.set_state(*tensors, inv_tol=1e-14, L=st["L"], LM=st["LM"], RM=st["RM"])
    rng = np.random.default_rng(32)
    x = torch.from_numpy(rng.standard_normal((len(MOMENTA), A.mtot)) + 1j * rng.standard_normal((len(MOMENTA), A.mtot))).cuda()
    ya, yb = torch.empty_like(x), torch.empty_like(x)
    A.apply(MOMENTA, x, ya)
    B.apply(MOMENTA, x, yb)
    e = (torch.linalg.vector_norm(ya - yb, dim=1) / torch.linalg.vector_norm(ya, dim=1)).max().item()
    assert e < 1e-10


def test_bond_input_rejects_foreign_blocks():
    """A bond object with a block outside the bonds of AL and AR is an error, not a silent drop."""
    s = es.state()
    K = (0,)
    P = im.HeffPlan(ctx(), [_lay(a) for a in s["ALob"]], [_lay(a) for a in s["ARob"]], [_lay(a) for a in s["NLob"]],
                    [_lay(c) for c in s["Cob"]], s["H"], s["E0"], K, 1)
    tensors = ([_dev(a) for a in s["ALob"]], [_dev(a) for a in s["ARob"]], [_dev(a) for a in s["NLob"]], [_dev(c) for c in s["Cob"]])
    c = s["Cob"][-1]
    bad = im.BlockLayout(1, np.array([1], np.int32), np.array([[40]], np.int32), np.array([[40]], np.int32), np.array([1], np.int32),
                         np.array([1], np.int32))
    good = (_lay(c), _dev(c))
    with pytest.raises(im.ImpsError, match="outside the bond structure"):
        P.set_state(*tensors, L=(bad, torch.ones(1, dtype=torch.float64, device="cuda")), LM=good, RM=good)
