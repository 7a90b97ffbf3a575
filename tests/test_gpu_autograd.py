"""Adjoints of the transfer-matrix, H_AC and H_C applications (torch.autograd over the library's
C-ABI vector-Jacobian products imps_tm_vjp / imps_hac_vjp / imps_hc_vjp), on the GPU.

  * torch.autograd.gradcheck (central finite differences, eps = 1e-6, FP64) of the mixed and
    regular TM in both directions (inputs x and the MPS tensors) and of H_AC (inputs x, EHL, EHR,
    HL, HR) on small seeded structures;
  * the anisotropic Heisenberg workload (reference Delta = -1.5, D = 256, 17 U(1) sectors):
    directional central differences (h = 1e-6) against the vector-Jacobian products, FP64 forward,
    and the FP32 forward's VJP against the FP64 one;
  * at the converged D = 16 state, the autograd gradient of f(AC) = <AC, H_AC AC> / 2 equals the
    oracle's H_AC(AC), and the analytic VUMPS gradient H_AC(AC) - AL H_C(C) has no component
    outside AC (both terms are multiples of AC = AL C at the fixed point).
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
from states import XXZ, Cell, random_like, xxz_bond0
from oracle import pyoracle as orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def _t(a, grad=False):
    return torch.tensor(np.ascontiguousarray(a, np.float64), device="cuda", requires_grad=grad)


_CTX = {}


def ctx():
    if "c" not in _CTX:
        _CTX["c"] = im.Context(0)
    return _CTX["c"]


def small_cell():
    k, s = xxz_bond0(5, 3)
    return Cell(XXZ, 2, [0], k, s, seed=41)


@pytest.mark.parametrize("direction", [0, 1], ids=["left", "right"])
@pytest.mark.parametrize("mixed", [False, True], ids=["regular", "mixed"])
def test_tm_gradcheck(direction, mixed):
    from imps_cpp_tools_amd.autograd import TransferMatrix
    c = small_cell()
    A, B = c.AL[0], c.AR[0]
    xl, yl = (c.diag(0), c.diag(1)) if direction == 0 else (c.diag(1), c.diag(0))
    tm = TransferMatrix(ctx(), direction, _lay(A), _lay(B) if mixed else None, _lay(xl), _lay(yl))
    rng = np.random.default_rng(3)
    x = _t(rng.standard_normal(xl.nelem), True)
    a = _t(A.data, True)
    if mixed:
        b = _t(B.data, True)
        assert torch.autograd.gradcheck(lambda x, a, b: tm(x, a, b), (x, a, b), eps=1e-6, atol=1e-8, rtol=1e-6)
    else:
        assert torch.autograd.gradcheck(lambda x, a: tm(x, a), (x, a), eps=1e-6, atol=1e-8, rtol=1e-6)


def _hac_inputs(c, rng):
    H = orc.model_ham(XXZ, [1.0, -1.5, 0.0])
    HL = orc.reduced_op(H, c.d, c.AL[1], 0)
    HR = orc.reduced_op(H, c.d, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    return c.struct[0], EHL, EHR, HL, HR


def test_hac_gradcheck():
    from imps_cpp_tools_amd.autograd import TwoSiteHAC
    c = small_cell()
    rng = np.random.default_rng(4)
    AC, EHL, EHR, HL, HR = _hac_inputs(c, rng)
    h = TwoSiteHAC(ctx(), _lay(AC), _lay(EHL), _lay(EHR), _lay(HL), _lay(HR))
    # non-symmetric operator data too: the adjoint must not assume H_AC = H_AC^T
    ins = [_t(rng.standard_normal(o.nelem), True) for o in (AC, EHL, EHR, HL, HR)]
    assert torch.autograd.gradcheck(lambda *a: h(*a), tuple(ins), eps=1e-6, atol=1e-8, rtol=1e-6)


def _config4_cell():
    k, s = xxz_bond0(17, 28, 2, 256)  # D = 256 over 17 sectors
    return Cell(XXZ, 2, [0], k, s, seed=1004)


def test_config4_tm_directional_fd():
    from imps_cpp_tools_amd.autograd import TransferMatrix
    c = _config4_cell()
    A = c.AL[0]
    tm = TransferMatrix(ctx(), 0, _lay(A), None, _lay(c.diag(0)), _lay(c.diag(1)))
    tm32 = TransferMatrix(ctx(), 0, _lay(A), None, _lay(c.diag(0)), _lay(c.diag(1)), im.FP32)
    rng = np.random.default_rng(5)
    x = _t(rng.standard_normal(c.diag(0).nelem), True)
    a = _t(A.data, True)
    ybar = _t(rng.standard_normal(c.diag(1).nelem))
    y = tm(x, a)
    gx, ga = torch.autograd.grad(y, (x, a), ybar)
    dx = _t(rng.standard_normal(x.numel()))
    da = _t(rng.standard_normal(a.numel()))
    h = 1e-6
    with torch.no_grad():
        fd = (torch.dot(ybar, tm(x + h * dx, a + h * da)) - torch.dot(ybar, tm(x - h * dx, a - h * da))).item() / (2 * h)
    an = (torch.dot(gx, dx) + torch.dot(ga, da)).item()
    assert abs(fd - an) <= 1e-7 * abs(an)
    # FP32 forward: values agree to FP32 accuracy, and the central difference of the FP32 forward
    # (step large enough that FP32 rounding stays below the tolerance) matches the adjoint
    y32 = tm32(x, a)
    assert (torch.linalg.vector_norm(y32 - y) / torch.linalg.vector_norm(y)).item() < 1e-5
    h32 = 1e-2
    with torch.no_grad():
        fd32 = (torch.dot(ybar, tm32(x + h32 * dx, a + h32 * da)) - torch.dot(ybar, tm32(x - h32 * dx, a - h32 * da))).item() / (2 * h32)
    assert abs(fd32 - an) <= 2e-4 * abs(an)


def test_config4_hac_directional_fd():
    from imps_cpp_tools_amd.autograd import TwoSiteHAC
    c = _config4_cell()
    rng = np.random.default_rng(6)
    AC, EHL, EHR, HL, HR = _hac_inputs(c, rng)
    h = TwoSiteHAC(ctx(), _lay(AC), _lay(EHL), _lay(EHR), _lay(HL), _lay(HR))
    ins = [_t(o.data, True) for o in (AC, EHL, EHR, HL, HR)]
    ins[0] = _t(rng.standard_normal(AC.nelem), True)
    ybar = _t(rng.standard_normal(AC.nelem))
    grads = torch.autograd.grad(h(*ins), ins, ybar)
    dirs = [_t(rng.standard_normal(t.numel())) for t in ins]
    eps = 1e-6
    with torch.no_grad():
        fp = torch.dot(ybar, h(*[t + eps * d for t, d in zip(ins, dirs)])).item()
        fm = torch.dot(ybar, h(*[t - eps * d for t, d in zip(ins, dirs)])).item()
    fd = (fp - fm) / (2 * eps)
    an = sum(torch.dot(g, d).item() for g, d in zip(grads, dirs))
    assert abs(fd - an) <= 1e-7 * abs(an)


def test_vumps_gradient_at_convergence():
    """grad_AC <AC, H_AC AC>/2 by autograd equals H_AC(AC); at the fixed point it differs from
    AL H_C(C) only along AC."""
    import excitation_state as es
    from imps_cpp_tools_amd.autograd import TwoSiteHAC
    s = es.state()
    v = s["v"]
    H, d = s["H"], s["d"]
    HL = orc.reduced_op(H, d, v.AL[1], 0)
    HR = orc.reduced_op(H, d, v.AR[1], 1)
    EHL, EHR = v.LE[0], v.RE[1]
    AC = orc.mps_mat(v.AL[0], v.C[0])
    h = TwoSiteHAC(ctx(), _lay(AC), _lay(EHL), _lay(EHR), _lay(HL), _lay(HR))
    ac = _t(AC.data, True)
    ops = [_t(o.data) for o in (EHL, EHR, HL, HR)]
    f = 0.5 * torch.dot(ac, h(ac, *ops))
    (g,) = torch.autograd.grad(f, (ac,))
    hac = orc.hac(EHL, EHR, HL, HR, AC).data
    assert np.linalg.norm(g.cpu().numpy() - hac) <= 1e-12 * np.linalg.norm(hac)
    hc = orc.hc(v.LE[1], v.RE[1], H, d, v.AL[0], v.AR[1], v.C[0])
    G = hac - orc.mps_mat(v.AL[0], hc).data
    a = AC.data
    G_perp = G - (a @ G) / (a @ a) * a
    assert np.linalg.norm(G_perp) <= 1e-8 * np.linalg.norm(hac)


def _hc_inputs(c, rng):
    H = orc.model_ham(XXZ, [1.0, -1.5, 0.0])
    AL, AR = c.AL[1], c.AR[0]
    HL = orc.reduced_op(H, c.d, AL, 0)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(0), rng, sym=True)
    C = random_like(c.diag(0), rng)
    return C, EHL, EHR, HL, AR, H, AL


def test_hc_gradcheck():
    from imps_cpp_tools_amd.autograd import BondHC
    c = small_cell()
    rng = np.random.default_rng(8)
    C, EHL, EHR, HL, AR, _, _ = _hc_inputs(c, rng)
    h = BondHC(ctx(), _lay(C), _lay(EHL), _lay(EHR), _lay(HL), _lay(AR))
    ins = [_t(rng.standard_normal(o.nelem), True) for o in (C, EHL, EHR, HL, AR)]
    assert torch.autograd.gradcheck(lambda *a: h(*a), tuple(ins), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_hc_forward_matches_oracle_and_fd_at_config4():
    """H_C on the D = 256 anisotropic Heisenberg structure: the differentiable forward equals the
    oracle's reference chain, and its adjoint matches a directional central difference."""
    from imps_cpp_tools_amd.autograd import BondHC
    c = _config4_cell()
    rng = np.random.default_rng(10)
    C, EHL, EHR, HL, AR, H, AL = _hc_inputs(c, rng)
    h = BondHC(ctx(), _lay(C), _lay(EHL), _lay(EHR), _lay(HL), _lay(AR))
    ins = [_t(o.data, True) for o in (C, EHL, EHR, HL, AR)]
    y = h(*ins)
    ref = orc.hc(EHL, EHR, H, c.d, AL, AR, C).data
    assert np.linalg.norm(y.detach().cpu().numpy() - ref) <= 1e-12 * np.linalg.norm(ref)
    ybar = _t(rng.standard_normal(C.nelem))
    grads = torch.autograd.grad(y, ins, ybar)
    dirs = [_t(rng.standard_normal(t.numel())) for t in ins]
    eps = 1e-6
    with torch.no_grad():
        fp = torch.dot(ybar, h(*[t + eps * d for t, d in zip(ins, dirs)])).item()
        fm = torch.dot(ybar, h(*[t - eps * d for t, d in zip(ins, dirs)])).item()
    fd = (fp - fm) / (2 * eps)
    an = sum(torch.dot(g, d).item() for g, d in zip(grads, dirs))
    assert abs(fd - an) <= 1e-7 * abs(an)


def test_vjp_entries_match_autograd_of_dense_contractions():
    """The C-ABI adjoints called directly equal the gradients torch computes for the same
    contractions written densely (one sector structure embedded in dense matrices)."""
    from states import dense_mps
    c = small_cell()
    A = c.AL[0]
    xl, yl = c.diag(0), c.diag(1)
    p = im.TMPlan(ctx(), 0, _lay(A), None, _lay(xl), _lay(yl))
    rng = np.random.default_rng(12)
    xd = rng.standard_normal(xl.nelem)
    yb = rng.standard_normal(yl.nelem)
    xbar = torch.empty(xl.nelem, dtype=torch.float64, device="cuda")
    abar = torch.empty(A.nelem, dtype=torch.float64, device="cuda")
    p.vjp(_t(A.data), None, _t(xd), _t(yb), xbar, abar)
    # dense: y = sum_s A_s^T x A_s over the full bonds
    l0, D0 = c.bond_offsets(0)
    l1, D1 = c.bond_offsets(1)
    Ad = torch.tensor(dense_mps(A, l0, D0, l1, D1), requires_grad=True)
    X = torch.zeros(D0, D0, dtype=torch.float64)
    for _, qi, qo, m in orc.OB.like(xl, xd).blocks():
        a, _ = l0[tuple(int(v) for v in qi)]
        X[a:a + m.shape[0], a:a + m.shape[1]] = torch.tensor(m)
    X.requires_grad_(True)
    Y = sum(Ad[s].T @ X @ Ad[s] for s in range(c.d))
    Yb = torch.zeros(D1, D1, dtype=torch.float64)
    for _, qi, qo, m in orc.OB.like(yl, yb).blocks():
        a, _ = l1[tuple(int(v) for v in qi)]
        Yb[a:a + m.shape[0], a:a + m.shape[1]] = torch.tensor(m)
    gX, gA = torch.autograd.grad(Y, (X, Ad), Yb)
    from states import extract
    ref_x = extract(xl, gX.numpy(), l0, l0)
    ref_a = extract(A, gA.numpy(), l0, l1, per_entry=True)
    assert np.abs(xbar.cpu().numpy() - ref_x).max() < 1e-13
    assert np.abs(abar.cpu().numpy() - ref_a).max() < 1e-13


def _chain(c, direction, mixed, cells, every=None, prec=im.FP64):
    """TransferChain over `cells` unit cells of c (left: sites 0, 1, ... from bond 0; right: sites
    N-1, ..., 0 from bond 0), with the per-step MPS tensors."""
    from imps_cpp_tools_amd.autograd import TransferChain, TransferMatrix
    N = c.N
    order = list(range(N)) * cells if direction == 0 else list(range(N - 1, -1, -1)) * cells
    tms, As, Bs = [], [], []
    for n in order:
        A, B = c.AL[n], c.AR[n]
        if direction == 0:
            xl, yl = c.diag(n), c.diag((n + 1) % N)
        else:
            xl, yl = c.diag((n + 1) % N), c.diag(n)
        tms.append(TransferMatrix(ctx(), direction, _lay(A), _lay(B) if mixed else None, _lay(xl), _lay(yl), prec))
        As.append(A)
        Bs.append(B)
    return TransferChain(tms, every), As, Bs


@pytest.mark.parametrize("direction", [0, 1], ids=["left", "right"])
@pytest.mark.parametrize("mixed", [False, True], ids=["regular", "mixed"])
def test_chain_gradcheck(direction, mixed):
    """The checkpointed chain (recomputed environments) under central finite differences: two cells,
    a checkpoint every other step; the same tensor at two steps gets both contributions."""
    c = small_cell()
    ch, As, Bs = _chain(c, direction, mixed, 2, every=2)
    rng = np.random.default_rng(7)
    x = _t(rng.standard_normal(ch.tms[0].xl.nelem), True)
    a = [_t(A.data, True) for A in c.AL]
    b = [_t(B.data, True) for B in c.AR]
    order = [0, 1, 0, 1] if direction == 0 else [1, 0, 1, 0]
    if mixed:
        f = lambda x, a0, a1, b0, b1: ch(x, [[a0, a1][n] for n in order], [[b0, b1][n] for n in order])  # noqa: E731
        assert torch.autograd.gradcheck(f, (x, a[0], a[1], b[0], b[1]), eps=1e-6, atol=1e-8, rtol=1e-6)
    else:
        f = lambda x, a0, a1: ch(x, [[a0, a1][n] for n in order])  # noqa: E731
        assert torch.autograd.gradcheck(f, (x, a[0], a[1]), eps=1e-6, atol=1e-8, rtol=1e-6)
    assert ch.last_checkpoints == 2  # the input and the environment after two steps


def test_config4_chain_recompute_equals_stored():
    """D = 256: three cells of left TMs.  The checkpointed chain keeps 2 of the 6 environments and
    gives the gradients of the step-by-step graph that stores all of them (bitwise for x)."""
    from imps_cpp_tools_amd.autograd import TransferMatrix
    c = _config4_cell()
    ch, As, _ = _chain(c, 0, False, 3)
    rng = np.random.default_rng(8)
    x0 = rng.standard_normal(ch.tms[0].xl.nelem)
    w = _t(rng.standard_normal(ch.tms[-1].yl.nelem))
    x1, a1 = _t(x0, True), [_t(A.data, True) for A in c.AL]
    y = ch(x1, [a1[n] for n in [0, 1] * 3])
    g1 = torch.autograd.grad(torch.dot(w, y), [x1] + a1)
    assert ch.last_checkpoints == 2 and ch.every == 3
    x2, a2 = _t(x0, True), [_t(A.data, True) for A in c.AL]
    z = x2
    for i, n in enumerate([0, 1] * 3):
        z = ch.tms[i](z, a2[n])
    assert torch.equal(z, y)
    g2 = torch.autograd.grad(torch.dot(w, z), [x2] + a2)
    assert torch.equal(g1[0], g2[0])
    for p, q in zip(g1[1:], g2[1:]):  # three contributions per tensor, summed in autograd's order
        assert (torch.linalg.vector_norm(p - q) / torch.linalg.vector_norm(q)).item() < 1e-14
