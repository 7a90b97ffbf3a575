"""Excitation observables on the GPU (imps_opeff_plan: OPeffConstants + ApplyOPeff through the excitation
engine) against the oracle (oracle/excitations.py, pinned in tests/test_excitation_observables.py).

States: the D = 4 Hubbard state of the oracle's pins (double occupancy, n_up n_down), the D = 16
Hubbard U(1)xU(1) state (double occupancy and n_up) and the D = 16 XXZ state (Sz, whose form is the
excitation's charge).  Each observable gets its ground-state expectation subtracted (the reference's
eobs); the engine's traces tr(OPL R), tr(L OPR) then vanish.  Four momenta per batched call.
Tolerances: O_eff X 1e-10 relative per momentum (relative to max(|O_eff X|, |X|)); MeasureExcitations'
per-site values 1e-10.
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
import excitation_state as es
from oracle import excitations as ex

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MOMENTA = [0.0, 0.3, 1.0, 0.5]
DOCC = np.diag([0.0, 0.0, 0.0, 1.0])
NUP = np.diag([0.0, 0.0, 1.0, 1.0])
SZ = np.diag([0.5, -0.5])
_CTX = []
_PLANS = {}


def ctx():
    if not _CTX:
        _CTX.append(im.Context(0))
    return _CTX[0]


def _lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def _dev(ob):
    return torch.from_numpy(np.ascontiguousarray(ob.data, np.float64)).cuda()


def _state(which):
    return {"hub4": lambda: es.hubbard_state(2, 1, 5, 4.0), "hub16": es.hubbard_state, "xxz": es.state}[which]()


def _eobs(s, O):
    _, _, (tl, _) = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], O)
    return float(np.real(tl)) / s["N"]


def plan(which, name, O, K):
    key = (which, name, K)
    if key not in _PLANS:
        s = _state(which)
        e0 = _eobs(s, O)
        print(f"opeff plan {which} {name} K={K} <O>={e0:.6f}", flush=True)
        P = im.HeffPlan(ctx(), [_lay(a) for a in s["ALob"]], [_lay(a) for a in s["ARob"]], [_lay(a) for a in s["NLob"]],
                        [_lay(c) for c in s["Cob"]], None, e0, K, len(MOMENTA), observable=O)
        P.set_state([_dev(a) for a in s["ALob"]], [_dev(a) for a in s["ARob"]], [_dev(a) for a in s["NLob"]], [_dev(c) for c in s["Cob"]])
        _PLANS[key] = (P, s, O - e0 * np.eye(O.shape[0]))
    return _PLANS[key]


def oracle_opeff(s, O, K, p, z):
    xd = ex.xdims(s["NL"], s["AR"], K)
    kfac = np.exp(1j * p * np.pi * s["N"])
    proj = abs(kfac - 1) < 1e-13 and all(k == 0 for k in K)  # the engine projects wherever e^{i p pi N} = 1
    LM, RM = (ex.lr_overlaps(s["AL"], s["AR"], K)[:2]) if proj else (None, None)
    OPLtot, OPRtot, _ = ex.opeff_constants(s["AL"], s["AR"], s["L"], s["R"], O)
    X = ex.from_vec_list(z, xd)
    Y = ex.apply_opeff(X, kfac, s["AL"], s["AR"], s["NL"], LM, RM, O, OPLtot, OPRtot)
    return ex.vectorize_list(Y), np.array([np.vdot(ex.vectorize(x), ex.vectorize(y)) for x, y in zip(X, Y)])


CASES = [("hub4", "docc", DOCC, (0, 0)), ("hub4", "docc", DOCC, (1, 1)), ("hub4", "docc", DOCC, (0, 2)),
         ("hub16", "docc", DOCC, (0, 0)), ("hub16", "nup", NUP, (1, 1)), ("xxz", "sz", SZ, (0,)), ("xxz", "sz", SZ, (2,))]


@pytest.mark.parametrize("which,name,O,K", CASES, ids=[f"{c[0]}-{c[1]}-{c[3]}" for c in CASES])
def test_opeff_apply_and_measure_match_oracle(which, name, O, K):
    P, s, O0 = plan(which, name, O, K)
    trl, trr, its = P.check()
    print(f"tr(OPL R) {trl:.2e} tr(L OPR) {trr:.2e} ({its} environment applications)", flush=True)
    assert abs(trl) < 1e-12 and abs(trr) < 1e-12
    rng = np.random.default_rng(31)
    z = rng.standard_normal((len(MOMENTA), P.mtot)) + 1j * rng.standard_normal((len(MOMENTA), P.mtot))
    x = torch.from_numpy(z).cuda()
    y = torch.empty_like(x)
    P.apply(MOMENTA, x, y)
    got = y.cpu().numpy()
    vals = P.measure(MOMENTA, x)
    for t, p in enumerate(MOMENTA):
        ref, rvals = oracle_opeff(s, O0, K, p, z[t])
        # relative to ||x|| where O_eff X nearly vanishes (a conserved charge at K = 0: O_eff = 0)
        err = np.linalg.norm(got[t] - ref) / max(np.linalg.norm(ref), np.linalg.norm(z[t]))
        verr = np.abs(vals[t] - rvals).max() / max(1.0, np.abs(rvals).max())
        print(f"{which} {name} K={K} p={p}: relerr {err:.2e}, per-site values {np.round(vals[t].real, 6)} ({verr:.1e})", flush=True)
        assert err < 1e-10
        assert verr < 1e-10


def test_conserved_charge_form_is_the_charge():
    P, s, O0 = plan("xxz", "sz", SZ, (2,))
    rng = np.random.default_rng(32)
    z = rng.standard_normal((len(MOMENTA), P.mtot)) + 1j * rng.standard_normal((len(MOMENTA), P.mtot))
    z /= np.linalg.norm(z, axis=1, keepdims=True)
    vals = P.measure(MOMENTA, torch.from_numpy(z).cuda())
    np.testing.assert_allclose(vals.sum(1), -1.0, atol=1e-10)  # K = 2 (2 Sz): Sz changes by -1 in this charge convention


def test_charge_changing_observable_is_rejected():
    s = _state("xxz")
    sx = np.array([[0.0, 0.5], [0.5, 0.0]])
    with pytest.raises(im.ImpsError):
        im.HeffPlan(ctx(), [_lay(a) for a in s["ALob"]], [_lay(a) for a in s["ARob"]], [_lay(a) for a in s["NLob"]],
                    [_lay(c) for c in s["Cob"]], None, 0.0, (0,), 1, observable=sx)
