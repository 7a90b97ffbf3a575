"""GPU VUMPS and excitations of the Z(q) and Bose-Hubbard chains (Models.cpp:325-507, 739-878),
pinned by closed forms and exact diagonalisation (tests/exact_diag.py, itself pinned against the
transverse-field Ising ring formula on the CPU).

* Quantum Ising, J = 1, h = 1.5 (sx = X/2, sz = Z/2: the transverse-field Ising chain
  -J' sum X X - g sum Z with J' = 1/4, g = 3/4, paramagnetic, gap 2 (g - J') = 1), D = 16 in two Z(2)
  sectors per bond: energy per site against the closed form -(1/pi) int_0^pi e(k)/2 dk, e(k) =
  2 sqrt(J'^2 + g^2 - 2 J' g cos k), 1e-10; <sz> = <Z>/2 and the bond value <sx sx> = <X X>/4
  against the Hellmann-Feynman derivatives -de/dg and -de/dJ' of the closed form, 1e-8; <sx> = 0
  (Z(2)-odd); the lowest two excitation energies of H_eff(p) in the Z(2)-odd sector (K = 1) against
  the two folded bands e(p pi), e(p pi + pi) of the two-site cell, 1e-8.
* Quantum Potts q = 2, h = 2 (-X X - h/2 (Z 1 + 1 Z): J' = 1, g = 2): energy against the closed form.
* Quantum Potts q = 3, h = 5 (Z(3), paramagnetic), D = 32: energy against exact diagonalisation of
  rings of 8, 10 and 12 sites (differences 5.9e-8, 8.3e-10: geometric with ratio about 70) with the
  geometric tail added, 1e-10.
* Bose-Hubbard d = 3, t = 1, U = 40, one boson per site (U(1), charges shifted by -1, 0, 1), eight
  sectors per bond: energy against exact diagonalisation of the 12-site ring at fixed particle
  number (ring sizes 10 and 12 differ by 1.5e-9, convergence factor 70 per two sites), 2e-10.
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
import exact_diag as ed

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

_CTX = []


@pytest.fixture(autouse=True)
def fixed_schedules(monkeypatch):
    monkeypatch.setenv("IMPS_AUTOTUNE", "0")


def ctx():
    if not _CTX:
        _CTX.append(im.Context(0))
    return _CTX[0]


def converge(model, QN, dims, params, seed, maxit=300, F=1e-11):
    from imps_cpp_tools_amd.vumps import VUMPS
    k, s = im.dimmap(model, 2, QN, dims)
    v = VUMPS(ctx(), model, 2, QN, k, s, params, seed=seed, krylov=30, env_krylov=100)
    for it in range(maxit):
        st = v.sweep()
        if max(st.F) < F:
            break
    print(f"model {hex(model)} D {v.bond_dims()} iterations {it + 1} E {st.energy:.15f} max F {max(st.F):.1e}", flush=True)
    assert max(st.F) < F
    return v, st


_QIM = {}


def qim_state():
    if not _QIM:
        _QIM["v"], _QIM["st"] = converge(im.QIM, [], [8, 8], [1.0, 1.5], seed=21)
    return _QIM["v"], _QIM["st"]


def test_quantum_ising_energy_and_observables():
    v, st = qim_state()
    Jp, g = 0.25, 0.75
    exact = ed.tfim_energy(Jp, g)
    e = v.energy()
    print(f"QIM E {e:.15f} exact {exact:.15f} diff {e - exact:.2e}", flush=True)
    assert abs(e - exact) < 1e-10
    # Hellmann-Feynman: <Z> = -de/dg, <X X> = -de/dJ' (central differences of the closed form)
    h = 1e-4
    dz = -(ed.tfim_energy(Jp, g + h) - ed.tfim_energy(Jp, g - h)) / (2 * h)
    dxx = -(ed.tfim_energy(Jp + h, g) - ed.tfim_energy(Jp - h, g)) / (2 * h)
    sxsx = np.kron(v.model_operators(im.QIM)["sx"], v.model_operators(im.QIM)["sx"])
    o = v.observables(["sz", "sx", "sp", "rz"])
    oxx = v.observables({"sxsx": sxsx})["sxsx"]
    print(f"<sz> {o['sz']} (2<sz> vs -de/dg {2 * o['sz'].mean() - dz:.2e}); <sx sx> {oxx} (vs {dxx / 4:.8f})", flush=True)
    np.testing.assert_allclose(2 * o["sz"], dz, atol=1e-8)
    np.testing.assert_allclose(o["rz"], 2 * o["sz"], atol=1e-14)
    np.testing.assert_allclose(4 * oxx, dxx, atol=1e-8)
    assert np.all(o["sx"] == 0.0) and np.all(o["sp"] == 0.0)
    eb = v.observables({"h": v.H})["h"]
    assert abs(eb.mean() - e) < 1e-10


def test_quantum_ising_excitations_are_the_fermion_bands():
    from imps_cpp_tools_amd.nullspace import left_nullspace
    v, _ = qim_state()
    E0 = v.energy()
    ps = [0.0, 0.25, 0.5, 0.8]
    nls = [left_nullspace(ctx(), v.AL[n], v.slay[n], seed=n) for n in range(2)]
    torch.cuda.synchronize()
    cl = [v.blay[(n + 1) % 2] for n in range(2)]
    P = im.HeffPlan(ctx(), v.slay, v.slay, [l for _, l in nls], cl, v.H, E0, (1,), len(ps))
    P.set_state(v.AL, v.AR, [a for a, _ in nls], v.C, inv_tol=1e-13)
    trl, trr, _ = P.check()
    assert abs(trl) < 1e-12 and abs(trr) < 1e-12
    ev, _, res, napply = P.eigs(ps, nev=2, tol=1e-11, maxit=2000)
    for t, p in enumerate(ps):
        ref = np.sort([ed.tfim_dispersion(0.25, 0.75, p * np.pi), ed.tfim_dispersion(0.25, 0.75, p * np.pi + np.pi)])
        print(f"p = {p}: {ev[t]} exact {ref} diff {np.abs(ev[t] - ref).max():.2e} residuals {res[t]}", flush=True)
        np.testing.assert_allclose(ev[t], ref, atol=1e-8)
    # K = 0 holds no single-particle states: its lowest energies lie above the two-particle threshold
    P0 = im.HeffPlan(ctx(), v.slay, v.slay, [l for _, l in nls], cl, v.H, E0, (0,), 1)
    P0.set_state(v.AL, v.AR, [a for a, _ in nls], v.C, inv_tol=1e-13)
    ev0, _, _, _ = P0.eigs([0.5], nev=1, tol=1e-9, maxit=2000)
    print(f"K = 0, p = 0.5: {ev0[0]} (two-particle threshold {2 * ed.tfim_dispersion(0.25, 0.75, 0.0):.6f})", flush=True)
    assert ev0[0, 0] > 2 * ed.tfim_dispersion(0.25, 0.75, 0.0) - 1e-6


def test_potts_two_states_is_ising():
    v, st = converge(im.QPOTTS(2), [], [6, 6], [2.0], seed=22)
    exact = ed.tfim_energy(1.0, 2.0)
    print(f"QPOTTS(2) E {v.energy():.15f} exact {exact:.15f}", flush=True)
    assert abs(v.energy() - exact) < 1e-10
    assert np.all(v.observables(["X"])["X"] == 0.0)


def test_potts_three_states_against_exact_diagonalisation():
    H = im.model_ham(im.QPOTTS(3), [5.0])
    e8, e10, e12 = (ed.ring_energy(H, 3, L) for L in (8, 10, 12))
    r = (e8 - e10) / (e10 - e12)  # geometric convergence in the ring size
    assert r > 20
    e_inf = e12 - (e10 - e12) / (r - 1)  # the remaining geometric tail (about 1e-11)
    v, st = converge(im.QPOTTS(3), [], [12, 10, 10], [5.0], seed=23)
    print(f"QPOTTS(3) E {v.energy():.15f} ED {e8:.15f} {e10:.15f} {e12:.15f} (ratio {r:.1f}) -> {e_inf:.15f} "
          f"diff {v.energy() - e_inf:.2e}", flush=True)
    assert abs(v.energy() - e_inf) < 1e-10
    z = v.observables(["Z"])["Z"]
    assert np.all(z > 0) and np.all(v.observables(["X"])["X"] == 0.0)


def test_bose_hubbard_against_exact_diagonalisation():
    H = im.model_ham(im.BHUB(3), [1.0, 40.0, 0.0])
    e12 = ed.ring_energy(H, 3, 12, fixed_total=12)
    # MakeDims shifts bond 1 by -1 against bond 0 (validpath [2, 0]: the reference product state
    # |2, 0>), while the Mott state carries the same charge c on every bond: eight sectors leave
    # c - 2 ... c + 2 on both bonds for c = 2 or 3
    v, st = converge(im.BHUB(3), [2], [1, 2, 3, 4, 4, 3, 2, 1], [1.0, 40.0, 0.0], seed=24)
    print(f"BHUB(3) bonds {[b[0].ravel().tolist() for b in v.bonds]} E {v.energy():.15f} ED(12) {e12:.15f} "
          f"diff {v.energy() - e12:.2e}", flush=True)
    assert abs(v.energy() - e12) < 2e-10
    n = v.observables(["n", "b"])
    np.testing.assert_allclose(n["n"].sum(), 2.0, atol=1e-10)  # two bosons per two-site cell
    assert np.all(n["b"] == 0.0)
