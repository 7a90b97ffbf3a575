"""GPU Krylov drivers against the oracle: environment GMRES solves (InvertE_proj) in FP64 and with
low-precision inner solves inside FP64 refinement, and the Lanczos lowest eigenpair of H_AC / H_C
against a dense eigensolve of the explicitly assembled operator.

Tolerances: environments ||y - y_oracle|| / ||y_oracle|| < 1e-11, both solved to a relative residual
of 1e-13 (the random XXZ cell has a subdominant TM eigenvalue 0.997, so the projected system has a
condition number near 1e3 and 1e-14 is below what the reference's unrestarted MGS GMRES attains;
observed 2e-14 .. 5e-13); GradNorm 1e-12 relative (observed 1e-14);
eigenvalues |e - e_dense| / |e_dense| < 1e-11 in every precision (low-precision Lanczos is refined
with FP64 applies); eigenvector residual ||H x - e x|| < 1e-9 |e|.
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
from states import FHUB, XXZ, Cell, fhub_bond0, random_like, xxz_bond0
from oracle import pyoracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = im.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


CELLS = {}


def cell(name):
    if name not in CELLS:
        if name == "xxz":
            k, s = xxz_bond0(9, 10)
            CELLS[name] = Cell(XXZ, 2, [0], k, s, seed=201)
        else:
            k, s = fhub_bond0(4, 8)
            CELLS[name] = Cell(FHUB, 2, [2, 0], k, s, seed=202)
    return CELLS[name]


def fixed_point(c, As, direction):
    """Dominant fixed point of the cell TM on the block-diagonal subspace, from a dense eigensolve of
    the operator assembled column by column with the oracle; trace-normalised."""
    base = c.diag(0)
    n = base.nelem
    M = np.zeros((n, n))
    for i in range(n):
        e = c.diag(0, np.eye(n)[i])
        M[:, i] = orc.tm_chain(direction, As, e, c.diag(0)).data
    w, V = np.linalg.eig(M)
    k = np.argmax(np.abs(w))
    assert abs(w[k] - 1) < 1e-10
    v = np.real(V[:, k])
    x = c.diag(0, v)
    x.data /= sum(np.trace(m) for _, _, _, m in x.blocks())
    return x


def identity(c):
    x = c.diag(0)
    for _, _, _, m in x.blocks():
        m[...] = np.eye(m.shape[0])
    return x


@pytest.mark.parametrize("prec", [im.FP64, im.FP32, im.BF16X3])
@pytest.mark.parametrize("name", ["xxz", "fhub"])
@pytest.mark.parametrize("direction", [0, 1])
def test_env_solve_parity(ctx, prec, name, direction):
    c = cell(name)
    rng = np.random.default_rng(31)
    As = c.AL if direction == 0 else c.AR
    if direction == 0:
        L, R = identity(c), fixed_point(c, As, 1)
    else:
        L, R = fixed_point(c, As, 0), identity(c)
    x = random_like(c.diag(0), rng, sym=True)
    yref, flag = orc.invert_e_proj(direction, As, x, L, R, tol=1e-13)
    assert flag > 0
    p = im.EnvPlan(ctx, direction, [lay(a) for a in As], lay(x), prec)
    p.set_mps([dev(a.data) for a in As])
    y = torch.zeros(x.nelem, dtype=torch.float64, device="cuda")
    it, rr = p.solve(dev(x.data), dev(L.data), dev(R.data), y, tol=1e-13, krylov=30)
    assert rr < 1e-13
    err = relerr(y.cpu().numpy(), yref.data)
    print(name, direction, prec, "env relerr", err, "steps", it, flush=True)
    assert err < 1e-11, err


@pytest.mark.parametrize("name", ["xxz", "fhub"])
@pytest.mark.parametrize("direction", [0, 1])
def test_env_solve_long_krylov_space(ctx, name, direction):
    """GMRES with a Krylov space of up to 300 vectors (the size the VUMPS driver uses): the inner
    products over hundreds of basis vectors and their in-kernel partial sums, against the oracle."""
    c = cell(name)
    rng = np.random.default_rng(37)
    As = c.AL if direction == 0 else c.AR
    L, R = (identity(c), fixed_point(c, As, 1)) if direction == 0 else (fixed_point(c, As, 0), identity(c))
    x = random_like(c.diag(0), rng, sym=True)
    yref, flag = orc.invert_e_proj(direction, As, x, L, R, tol=1e-13)
    assert flag > 0
    p = im.EnvPlan(ctx, direction, [lay(a) for a in As], lay(x), im.FP64)
    p.set_mps([dev(a.data) for a in As])
    y = torch.zeros(x.nelem, dtype=torch.float64, device="cuda")
    it, rr = p.solve(dev(x.data), dev(L.data), dev(R.data), y, tol=1e-14, krylov=300)
    err = relerr(y.cpu().numpy(), yref.data)
    print(name, direction, "n", x.nelem, "steps", it, "env relerr", err, flush=True)
    assert rr < 1e-13
    assert err < 1e-11, err


def dense_operator(apply, n):
    M = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        M[:, i] = apply(e)
    return M


@pytest.mark.parametrize("prec", [im.FP64, im.FP32, im.BF16X3])
@pytest.mark.parametrize("name", ["xxz", "fhub"])
def test_eigsh_hac(ctx, prec, name):
    c = cell(name)
    H = orc.model_ham(c.model, [1.0, -1.0, 0.0] if c.model == XXZ else [1.0, 4.0, 0.0, 0.0, 0.0])
    rng = np.random.default_rng(32)
    HL = orc.reduced_op(H, c.d, c.AL[1], 0)
    HR = orc.reduced_op(H, c.d, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    x0 = random_like(c.struct[0], rng)
    n = x0.nelem

    def ap(v):
        t = random_like(c.struct[0], rng)
        t.data = v
        return orc.hac(EHL, EHR, HL, HR, t).data

    M = dense_operator(ap, n)
    assert np.abs(M - M.T).max() < 1e-12
    e_dense = np.linalg.eigvalsh(M)[0]
    args = (lay(x0), lay(EHL), lay(EHR), lay(HL), lay(HR))
    p = im.HACPlan(ctx, *args, prec)
    p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
    p64 = None
    if prec != im.FP64:
        p64 = im.HACPlan(ctx, *args, im.FP64)
        p64.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
    x = dev(x0.data)
    e, it, res = im.eigsh(p, x, tol=1e-12, krylov=30, plan_fp64=p64)
    assert abs(e - e_dense) < 1e-11 * abs(e_dense), (e, e_dense)
    v = x.cpu().numpy()
    assert np.linalg.norm(M @ v - e * v) < 1e-9 * abs(e)


def test_eigsh_hc(ctx):
    c = cell("fhub")
    H = orc.model_ham(FHUB, [1.0, 4.0, 0.0, 0.0, 0.0])
    rng = np.random.default_rng(33)
    AL, AR = c.AL[1], c.AR[0]
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(0), rng, sym=True)
    C0 = random_like(c.diag(0), rng)

    def ap(v):
        t = random_like(c.diag(0), rng)
        t.data = v
        return orc.hc(EHL, EHR, H, 4, AL, AR, t).data

    M = dense_operator(ap, C0.nelem)
    e_dense = np.linalg.eigvalsh(0.5 * (M + M.T))[0]
    HL = orc.reduced_op(H, 4, AL, 0)
    p = im.HCPlan(ctx, lay(C0), lay(EHL), lay(EHR), lay(HL), lay(AR), im.FP64)
    p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(AR.data))
    x = dev(C0.data)
    e, it, res = im.eigsh(p, x, tol=1e-12, krylov=30)
    assert abs(e - e_dense) < 1e-11 * abs(e_dense), (e, e_dense)


@pytest.mark.parametrize("name", ["xxz", "fhub"])
@pytest.mark.parametrize("direction", [0, 1])
def test_fixed_point_parity(ctx, name, direction):
    """TMDominantEig (trueLR fixed points) on the GPU (EnvPlan.fixed_point, from the identity) against
    the dense dominant eigenvector of the oracle's cell TM, both trace-normalised."""
    c = cell(name)
    As = c.AL if direction == 1 else c.AR  # R of the AL cell (right TM), L of the AR cell (left TM)
    ref = fixed_point(c, As, direction)
    p = im.EnvPlan(ctx, direction, [lay(a) for a in As], lay(c.diag(0)), im.FP64)
    p.set_mps([dev(a.data) for a in As])
    x0 = identity(c)
    x0.data /= sum(np.trace(m) for _, _, _, m in x0.blocks())
    x = dev(x0.data)
    it, delta = p.fixed_point(x, tol=1e-14, maxit=2000)
    assert delta < 1e-12
    assert relerr(x.cpu().numpy(), ref.data) < 1e-10


def test_gradnorm_matches_oracle():
    """GradNormRight at the seeded random mixed-canonical D = 64 Heisenberg start: the GPU VUMPS's F
    (H_AC on the GPU, null-space-free form) against the oracle's (explicit right null space)."""
    import json
    import os
    from imps_cpp_tools_amd.vumps import VUMPS
    from oracle_vumps import OracleVUMPS
    with open(os.path.join(os.path.dirname(__file__), "golden", "vumps_xxz_d64.json")) as f:
        g = json.load(f)
    args = (g["N"], g["QN"], np.array(g["keys0"], np.int32), np.array(g["sizes0"], np.int32), g["params"])
    v = VUMPS(im.Context(0), im.XXZ, *args, seed=g["seed"])
    o = OracleVUMPS(0, *args, seed=g["seed"])
    for n in range(g["N"]):
        left = (n - 1) % g["N"]
        hac, _ = v._hac_plans(n)
        AC, _ = v._prod(im.MAT_MPS, v.blay[n], v.C[left], v.slay[n], v.AR[n], ("CAR", n))
        F_gpu, _ = v.gradnorm(n, AC, hac)
        HL = orc.reduced_op(o.H, o.d, o.AL[left], 0)
        HR = orc.reduced_op(o.H, o.d, o.AR[(n + 1) % g["N"]], 1)
        F_ref = orc.gradnorm(1, o.AR[n], o.C[left], o.LE[n], o.RE[(n + 1) % g["N"]], HL, HR)
        print("gradnorm", n, F_gpu, F_ref, abs(F_gpu - F_ref) / F_ref, flush=True)
        assert abs(F_gpu - F_ref) <= 1e-12 * F_ref, (n, F_gpu, F_ref)


@pytest.mark.parametrize("prec", [im.FP64, im.BF16X3], ids=["fp64", "bf16x3"])
def test_eigsh_graph_cycles_match_eager(prec, monkeypatch):
    """Lanczos restart cycles replayed from a captured HIP graph (a context on its own stream)
    give bitwise the results of issuing the launches directly (IMPS_KRYLOV_GRAPH=0), and the dense
    eigenvalue; a small Krylov space forces many restarts, hence many replays."""
    c = cell("fhub")
    H = orc.model_ham(c.model, [1.0, 4.0, 0.0, 0.0, 0.0])
    rng = np.random.default_rng(33)
    HL = orc.reduced_op(H, c.d, c.AL[1], 0)
    HR = orc.reduced_op(H, c.d, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    x0 = random_like(c.struct[0], rng)

    def ap(v):
        t = random_like(c.struct[0], rng)
        t.data = v
        return orc.hac(EHL, EHR, HL, HR, t).data

    e_dense = np.linalg.eigvalsh(dense_operator(ap, x0.nelem))[0]
    sctx = im.Context(0, torch.cuda.Stream())
    args = (lay(x0), lay(EHL), lay(EHR), lay(HL), lay(HR))
    ops = (dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
    p = im.HACPlan(sctx, *args, prec)
    p.set_operator(*ops)
    p64 = None
    if prec != im.FP64:
        p64 = im.HACPlan(sctx, *args, im.FP64)
        p64.set_operator(*ops)
    got = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("IMPS_KRYLOV_GRAPH", flag)
        x = dev(x0.data)
        torch.cuda.synchronize()
        e, it, res = im.eigsh(p, x, tol=1e-12, krylov=12, plan_fp64=p64)
        torch.cuda.synchronize()
        got[flag] = (e, it, x.cpu().numpy())
    print("iterations", got["1"][1], flush=True)
    assert got["1"][1] == got["0"][1] and got["1"][1] > 3 * 12
    assert got["1"][0] == got["0"][0]
    np.testing.assert_array_equal(got["1"][2], got["0"][2])
    assert abs(got["1"][0] - e_dense) < 1e-11 * abs(e_dense)
