"""VUMPS on the D = 1024 Fermi-Hubbard chain converged to gradient norm F <= 1e-10, in FP64 and with
FP32 operators under FP64 refinement, from one common D = 1024 state.

Workload: t = 1, U = 4, half filling (QN = [2, 0] on the two-site cell), the reference's bond growth
(ExpandFromH, at most 256 states per bond and iteration, VUMPS_main.cpp:273-306) from a D = 64 bond
of 25 sectors to D = 1024 on both bonds, in FP64.  From that state (written and read back through
the state-file format) VUMPS iterates at fixed structure until max F <= 1e-10 (VUMPS_main.cpp:340).
At fixed structure the iteration has one fixed point, whose energy error at gradient norm F is
O(F^2): the FP64 and the mixed runs must agree to 1e-12.  The energy lies above the Lieb-Wu value
by the finite-D error (about 1e-7 at D = 1024).
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im

pytestmark = pytest.mark.gpu

MAX_ITERS = 1200
_E = {}


@pytest.fixture(autouse=True)
def fixed_schedules(monkeypatch):
    """Fixed panel schedules: every run follows the same iteration sequence."""
    monkeypatch.setenv("IMPS_AUTOTUNE", "0")


def _lieb_wu(U):
    """Ground-state energy per site of the half-filled chain H = -t sum c^+c + U sum n_up n_dn, t = 1:
    -4 int_0^inf J0(w) J1(w) / (w (1 + exp(w U / 2))) dw."""
    from scipy import integrate, special
    f = lambda w: special.j0(w) * special.j1(w) / (w * (1.0 + np.exp(min(w * U / 2.0, 700.0))))  # noqa: E731
    edges = np.linspace(0.0, 400.0, 4001)
    return -4.0 * sum(integrate.quad(f, a, b)[0] for a, b in zip(edges[:-1], edges[1:]))


def _vumps(prec, **kw):
    from imps_cpp_tools_amd import workloads as wl
    from imps_cpp_tools_amd.vumps import VUMPS
    cfg = wl.HUBBARD_1024
    k0, s0 = wl.fhub_profile(5, 8, 1, 64)
    return VUMPS(im.Context(0), cfg["model"], cfg["N"], cfg["QN"], k0, s0, cfg["params"], prec=prec, seed=1003, krylov=30,
                 env_krylov=300, **kw)


@pytest.fixture(scope="module")
def grown(tmp_path_factory):
    import os
    os.environ["IMPS_AUTOTUNE"] = "0"
    v = _vumps(im.FP64, mmax=1024, dm_max=256, expand_pause=1, expand_thresh=1.0)
    for _ in range(20):
        if min(v.bond_dims()) >= 1024:
            break
        v.sweep()
    assert v.bond_dims() == [1024, 1024]
    path = str(tmp_path_factory.mktemp("d1024") / "hub_d1024.imps")
    v.save(path)
    v.close()
    return path


def _converge(path, prec):
    v = _vumps(prec)
    v.load(path)
    it, st = 0, None
    while it < MAX_ITERS:
        st = v.sweep()
        it += 1
        if it % 100 == 0:
            print(f"iteration {it}: F {max(st.F):.3e} E {st.energy:.15f}", flush=True)
        if max(st.F) <= 1e-10:
            break
    v.close()
    print(f"converged in {it} iterations: F {max(st.F):.3e} E {st.energy:.15f}", flush=True)
    return it, st


def test_d1024_fp64_converges_to_1e10(grown):
    it, st = _converge(grown, im.FP64)
    assert max(st.F) <= 1e-10, f"F {max(st.F):.2e} after {it} iterations"
    exact = _lieb_wu(4.0) - 1.0
    assert exact < st.energy < exact + 3e-7
    _E["fp64"] = st.energy


def test_d1024_mixed_fp32_converges_to_fp64_energy(grown):
    it, st = _converge(grown, im.FP32)
    assert max(st.F) <= 1e-10, f"F {max(st.F):.2e} after {it} iterations"
    if "fp64" not in _E:
        pytest.skip("needs the FP64 run of this module")
    assert abs(st.energy - _E["fp64"]) <= 1e-12, f"mixed {st.energy:.15f} vs FP64 {_E['fp64']:.15f}"
