"""Host-side block arithmetic of the excitation engine on CPU tensors: the cached gather-scale-sum
behind ApplyOperator and the reduced-operator grids must give exactly the structure of the chain
of sector-set sums it replaces (BlockMat::operator+= applied term by term) and its values.

The objects are synthetic: four list entries (a d = 2 two-site tensor) whose sector sets differ
and overlap, so that the sums insert, merge and keep sectors; one operator row is empty."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from imps_cpp_tools_amd import BlockLayout
from imps_cpp_tools_amd.excitations import BlockOps, DObj


def _object(rng, nsym=1):
    sets = [[0, 2], [2, 4, 6], [0, 6], [4]]  # per entry, the ingoing charges present
    shapes = {}
    nblk, qin, qout, rows, cols = [], [], [], [], []
    for ch in sets:
        nblk.append(len(ch))
        for q in sorted(ch):
            r, c = shapes.setdefault(q, (int(rng.integers(1, 4)), int(rng.integers(1, 4))))
            qin.append([q]), qout.append([q + 1]), rows.append(r), cols.append(c)
    lay = BlockLayout(nsym, nblk, qin, qout, rows, cols)
    data = torch.from_numpy(rng.standard_normal((2, lay.nelem)))
    return DObj(lay, data)


def _ops():
    ops = BlockOps(None, 1, d=2, H=None)
    ops.dev = "cpu"
    return ops


def _chain(ops, AA, H):
    """The former ApplyOperator: rows combined term by term with add, then concatenated."""
    D = H.shape[0]
    ents = [ops.entry(AA, e) for e in range(D)]
    outs = []
    for row in range(D):
        terms = [(H[row, col], ents[col]) for col in range(D) if H[row, col] != 0.0]
        if terms:
            outs.append(ops.combine(terms))
        else:
            outs.append(DObj(BlockLayout(1, [0], np.zeros((0, 1)), np.zeros((0, 1)), [], []), ops._zeros(0, True)))
    return ops.concat(outs)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_apply_operator_matches_chain_of_sums(seed):
    rng = np.random.default_rng(seed)
    AA = _object(rng)
    H = rng.standard_normal((4, 4))
    H[np.abs(H) < 0.4] = 0.0
    H[2, :] = 0.0  # an empty row
    ops = _ops()
    got = ops.apply_operator(AA, H)
    ref = _chain(ops, AA, H)
    assert got.lay == ref.lay
    np.testing.assert_allclose(got.data.numpy(), ref.data.numpy(), rtol=0, atol=1e-14)
    # cached map reused on new data of the same structure
    AA2 = DObj(AA.lay, torch.from_numpy(rng.standard_normal((2, AA.lay.nelem))))
    np.testing.assert_allclose(ops.apply_operator(AA2, H).data.numpy(), _chain(ops, AA2, H).data.numpy(), rtol=0, atol=1e-14)


def test_apply_operator_dense_sector():
    """One sector present in every entry: out[row] = sum_col H[row, col] AA[col] on that block."""
    rng = np.random.default_rng(5)
    lay = BlockLayout(1, [1, 1, 1, 1], [[3]] * 4, [[4]] * 4, [2] * 4, [3] * 4)
    AA = DObj(lay, torch.from_numpy(rng.standard_normal((2, lay.nelem))))
    H = rng.standard_normal((4, 4))
    got = _ops().apply_operator(AA, H)
    blocks = AA.data.numpy().reshape(2, 4, 6)
    want = np.einsum("rc,pcx->prx", H, blocks).reshape(2, -1)
    np.testing.assert_allclose(got.data.numpy(), want, rtol=0, atol=1e-14)


@pytest.mark.parametrize("cplx", [(False, False), (True, False), (False, True), (True, True)])
def test_sector_trace(cplx):
    """_tr(a, b) = sum over sectors of trace(A_q B_q'), B's block ingoing where A's goes out and
    back; sectors without a partner contribute nothing."""
    rng = np.random.default_rng(9)
    qa = [0, 2, 4]          # A: q -> q + 1, B: q + 1 -> q for q in 0, 2 (4 has no partner)
    shp = {0: (2, 3), 2: (3, 1), 4: (2, 2)}
    la = BlockLayout(1, [3], [[q] for q in qa], [[q + 1] for q in qa], [shp[q][0] for q in qa], [shp[q][1] for q in qa])
    lb = BlockLayout(1, [2], [[1], [3]], [[0], [2]], [shp[0][1], shp[2][1]], [shp[0][0], shp[2][0]])

    def obj(lay, c):
        n = lay.nelem
        return DObj(lay, torch.from_numpy(rng.standard_normal((2, n)) if c else rng.standard_normal(n)))
    a, b = obj(la, cplx[0]), obj(lb, cplx[1])
    ops = _ops()

    def blocks(x, lay):
        d = x.data.numpy()
        z = d[0] + 1j * d[1] if x.cplx else d.astype(complex)
        out, o = {}, 0
        for q, r, c in zip(lay.qin[:, 0], lay.rows, lay.cols):
            out[int(q)] = z[o:o + r * c].reshape(r, c, order="F")
            o += r * c
        return out
    A, B = blocks(a, la), blocks(b, lb)
    want = sum(np.trace(A[q] @ B[q + 1]) for q in (0, 2))
    got = ops._tr(a, b)
    assert abs(got - want) <= 1e-13 * max(1.0, abs(want))
