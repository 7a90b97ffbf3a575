"""Adjoints of the transfer-matrix and H_AC applications (torch.autograd), on the GPU.

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
    # FP32 forward: same adjoint (FP64 backward), forward values agree to FP32 accuracy
    y32 = tm32(x, a)
    assert (torch.linalg.vector_norm(y32 - y) / torch.linalg.vector_norm(y)).item() < 1e-5
    gx32, ga32 = torch.autograd.grad(y32, (x, a), ybar)
    assert torch.allclose(gx32, gx) and torch.allclose(ga32, ga)


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
