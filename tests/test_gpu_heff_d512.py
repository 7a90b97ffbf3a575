"""The C-ABI excitation engine at the size of the 32-momentum Hubbard workload (D = 512 on both bonds
of the two-site cell, 81 U(1)xU(1) sectors, K = (0, 0), 43084 complex parameters per momentum), on
the seeded random mixed-canonical state of that structure.  The dense oracle does not reach this
size, so the checks are size-independent properties and a second, independent implementation:

* HeffConstants' trace check tr(HL R) = tr(L HR) = 0 (Excitations_helpers.hpp:287-303), 1e-12;
* Hermiticity of H_eff(k) at all 32 momenta of one batched call: <w, H x> = conj(<x, H w>), 1e-10;
* H_eff X of the batched engine against the Python restatement of ApplyHeff over the library's
  block products (excitations.py: a different evaluation order and GMRES) at three momenta, 1e-9;
* the FP32-refined engine against the FP64 one: H_eff X 1e-9 and the lowest four energies of all
  32 momenta 1e-8 (both converged to 1e-10 residuals);
* the lowest energies of the batched eigensolver are the Rayleigh quotients of its unit Ritz
  vectors, which are orthonormal per momentum (1e-10).
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

PS = [j / 31 for j in range(32)]
_S = {}


@pytest.fixture(autouse=True)
def fixed_schedules(monkeypatch):
    monkeypatch.setenv("IMPS_AUTOTUNE", "0")


def setup():
    if _S:
        return _S
    import os
    os.environ["IMPS_AUTOTUNE"] = "0"
    from imps_cpp_tools_amd import workloads as wl
    from imps_cpp_tools_amd.nullspace import left_nullspace
    from imps_cpp_tools_amd.vumps import VUMPS
    ec = wl.HUBBARD_512
    ctx = im.Context(0)
    k5, s5 = wl.fhub_profile(ec["nside"], ec["peak"], 1, ec["total"])
    v = VUMPS(ctx, ec["model"], ec["N"], ec["QN"], k5, s5, ec["params"], prec=im.FP64, seed=1005, env_krylov=300)
    E0 = v.energy()
    nls = [left_nullspace(ctx, v.AL[n], v.slay[n], seed=n) for n in range(ec["N"])]
    torch.cuda.synchronize()
    cl = [v.blay[(n + 1) % ec["N"]] for n in range(ec["N"])]
    plans = {}
    for prec in (im.FP64, im.FP32):
        P = im.HeffPlan(ctx, v.slay, v.slay, [l for _, l in nls], cl, v.H, E0, (0, 0), len(PS), prec=prec)
        P.set_state(v.AL, v.AR, [a for a, _ in nls], v.C, inv_tol=1e-12)
        plans[prec] = P
    print(f"D {v.bond_dims()}, sectors {[len(b[1]) for b in v.bonds]}, mtot {plans[im.FP64].mtot}", flush=True)
    _S.update(ctx=ctx, v=v, nls=nls, E0=E0, cl=cl, plans=plans)
    return _S


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).cuda()


def test_constants_trace_check():
    s = setup()
    trl, trr, its = s["plans"][im.FP64].check()
    print(f"tr(HL R) {trl:.2e} tr(L HR) {trr:.2e} env applications {its}", flush=True)
    assert abs(trl) < 1e-12 and abs(trr) < 1e-12


def test_hermitian_at_all_momenta():
    s = setup()
    P = s["plans"][im.FP64]
    x, w = _rand((len(PS), P.mtot), 1), _rand((len(PS), P.mtot), 2)
    hx, hw = torch.empty_like(x), torch.empty_like(x)
    steps, rr = P.apply(PS, x, hx)
    P.apply(PS, w, hw)
    assert rr < 1e-12
    a = (w.conj() * hx).sum(1).cpu().numpy()
    b = (x.conj() * hw).sum(1).cpu().numpy()
    worst = float(np.max(np.abs(a - b.conj()) / np.abs(a)))
    print(f"hermiticity worst {worst:.2e}, GMRES steps {steps}", flush=True)
    assert worst < 1e-10


def test_batched_engine_equals_python_restatement():
    s = setup()
    from imps_cpp_tools_amd.excitations import ExcitationEngine
    v = s["v"]
    P = s["plans"][im.FP64]
    eng = ExcitationEngine(s["ctx"], v.AL, v.slay, v.AR, v.slay, [a for a, _ in s["nls"]], [l for _, l in s["nls"]], v.C, s["cl"], v.H,
                           4, s["E0"], (0, 0), inv_tol=1e-12)
    x = _rand((len(PS), P.mtot), 3)
    y = torch.empty_like(x)
    P.apply(PS, x, y)
    for t in (0, 5, 16):
        ref = eng.matvec(x[t], PS[t])
        err = float(torch.linalg.vector_norm(y[t] - ref) / torch.linalg.vector_norm(ref))
        print(f"p = {PS[t]:.4f}: C engine vs Python restatement {err:.2e}", flush=True)
        assert err < 1e-9


def test_fp32_refined_engine_equals_fp64():
    s = setup()
    P64, P32 = s["plans"][im.FP64], s["plans"][im.FP32]
    x = _rand((len(PS), P64.mtot), 4)
    y64, y32 = torch.empty_like(x), torch.empty_like(x)
    P64.apply(PS, x, y64)
    steps, rr = P32.apply(PS, x, y32)
    assert rr < 1e-12
    err = float((torch.linalg.vector_norm(y32 - y64, dim=1) / torch.linalg.vector_norm(y64, dim=1)).max())
    print(f"FP32-refined vs FP64 apply: {err:.2e} ({steps} GMRES steps)", flush=True)
    assert err < 1e-9
    e64, v64, r64, n64 = P64.eigs(PS, nev=4, tol=1e-10, krylov=20, vectors=True)
    e32, _, r32, n32 = P32.eigs(PS, nev=4, tol=1e-10, krylov=20)
    print(f"eigs: FP64 {n64} applies (res {r64.max():.1e}), FP32-refined {n32} (res {r32.max():.1e}), "
          f"energies differ by {np.abs(e32 - e64).max():.1e}", flush=True)
    assert np.abs(e32 - e64).max() < 1e-8
    # Ritz vectors: orthonormal per momentum, energies their Rayleigh quotients
    for t in (0, 13, 31):
        V = v64[t]
        G = (V.conj() @ V.T).cpu().numpy()
        assert np.abs(G - np.eye(4)).max() < 1e-10
    xs = v64[:, 0, :].contiguous()
    hy = torch.empty_like(xs)
    P64.apply(PS, xs, hy)
    rq = (xs.conj() * hy).sum(1).real.cpu().numpy()
    assert np.abs(rq - e64[:, 0]).max() < 1e-9
