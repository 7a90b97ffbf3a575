"""Edge cases of the operator applications against the oracle: sectors larger than the 64 x 64
split-copy tile, an MPS with an empty physical entry, all sectors of size 1, and batches that are
not a multiple of the four vectors of a workgroup.  Tolerances as in test_gpu_parity."""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
from states import XXZ, Cell, random_like
from oracle import pyoracle as orc
from oracle.pyoracle import OB

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {im.FP64: 1e-12, im.FP32: 1e-5, im.BF16X3: 2e-4}
PRECS = [im.FP64, im.FP32, im.BF16X3]


@pytest.fixture(scope="module")
def ctx():
    c = im.Context(0)
    yield c
    c.close()


def lay(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _cell(kind):
    if kind == "big":   # sectors of 40, 130, 40: blocks span several split tiles
        return Cell(XXZ, 2, [0], np.array([[-2], [0], [2]], np.int32), np.array([40, 130, 40], np.int32), seed=11)
    return Cell(XXZ, 2, [0], np.array([[2 * (i - 4)] for i in range(9)], np.int32), np.ones(9, np.int32), seed=12)


def _hac_check(ctx, c, nvec, rng):
    H = orc.model_ham(XXZ, [1.0, -1.5, 0.3])
    HL = orc.reduced_op(H, 2, c.AL[1], 0)
    HR = orc.reduced_op(H, 2, c.AR[1], 1)
    EHL = random_like(c.diag(0), rng, sym=True)
    EHR = random_like(c.diag(1), rng, sym=True)
    ACs = [random_like(c.struct[0], rng) for _ in range(nvec)]
    refs = [orc.hac(EHL, EHR, HL, HR, a).data for a in ACs]
    X = dev(np.concatenate([a.data for a in ACs]))
    for prec in PRECS:
        p = im.HACPlan(ctx, lay(c.struct[0]), lay(EHL), lay(EHR), lay(HL), lay(HR), prec)
        p.set_operator(dev(EHL.data), dev(EHR.data), dev(HL.data), dev(HR.data))
        y = torch.full((X.numel(),), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(X, y, nvec)
        got = y.cpu().numpy().reshape(nvec, -1)
        for v in range(nvec):
            assert relerr(got[v], refs[v]) < TOL[prec], (im.PREC_NAMES[prec], nvec, v)


@pytest.mark.parametrize("kind", ["big", "ones"])
@pytest.mark.parametrize("nvec", [1, 3, 37])
def test_hac_edge_structures(ctx, kind, nvec):
    _hac_check(ctx, _cell(kind), nvec, np.random.default_rng(21))


@pytest.mark.parametrize("kind", ["big", "ones"])
@pytest.mark.parametrize("direction", [0, 1])
def test_tm_edge_structures(ctx, kind, direction):
    c = _cell(kind)
    rng = np.random.default_rng(22)
    A = c.AL[0]
    xl, yl = (c.diag(0), c.diag(1)) if direction == 0 else (c.diag(1), c.diag(0))
    xs = [random_like(xl, rng) for _ in range(5)]
    refs = [orc.tm_mixed(direction, A, None, x, yl).data for x in xs]
    X = dev(np.concatenate([x.data for x in xs]))
    for prec in PRECS:
        p = im.TMPlan(ctx, direction, lay(A), None, lay(xl), lay(yl), prec)
        p.set_mps(dev(A.data))
        y = torch.full((yl.nelem * 5,), float("nan"), dtype=torch.float64, device="cuda")
        p.apply(X, y, 5)
        got = y.cpu().numpy().reshape(5, -1)
        for v in range(5):
            assert relerr(got[v], refs[v]) < TOL[prec], (im.PREC_NAMES[prec], v)


def test_tm_with_empty_physical_entry(ctx):
    """An MPS tensor whose second physical entry has no blocks: the TM sums over the first only."""
    c = _cell("big")
    A = c.AL[0]
    n0 = int(A.nblk[0])
    size0 = int((A.rows[:n0].astype(np.int64) * A.cols[:n0]).sum())
    A1 = OB(A.nsym, np.array([n0, 0], np.int32), A.qin[:n0], A.qout[:n0], A.rows[:n0], A.cols[:n0], A.data[:size0].copy())
    rng = np.random.default_rng(23)
    for direction in (0, 1):
        xl, yl = (c.diag(0), c.diag(1)) if direction == 0 else (c.diag(1), c.diag(0))
        x = random_like(xl, rng)
        ref = orc.tm_mixed(direction, A1, None, x, yl).data
        for prec in PRECS:
            p = im.TMPlan(ctx, direction, lay(A1), None, lay(xl), lay(yl), prec)
            p.set_mps(dev(A1.data))
            y = torch.full((yl.nelem,), float("nan"), dtype=torch.float64, device="cuda")
            p.apply(dev(x.data), y)
            assert relerr(y.cpu().numpy(), ref) < TOL[prec], (direction, im.PREC_NAMES[prec])
