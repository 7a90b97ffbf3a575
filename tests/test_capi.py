"""CPU tests of libimps_hip.so: it loads, exports every entry point of include/imps_capi.h, its
bookkeeping equals the oracle's and the hand-derived answers exactly, and host-only plans
(created without a device context) report the algorithmic cost of the block contractions."""
import ctypes
import os
import re

import numpy as np
import pytest

import imps_cpp_tools_amd as im
from states import FHUB, XXZ, Cell, fhub_bond0, xxz_bond0
from oracle import pyoracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "imps_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(imps_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(im.LIB_PATH)
    names = header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_no_device_is_reported_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(im.ImpsError) as e:
        im.Context(0)
    assert e.value.status == 5  # IMPS_NO_DEVICE


@pytest.mark.parametrize("model,N,QN", [(XXZ, 2, [0]), (XXZ, 3, [1]), (XXZ, 4, [0]), (FHUB, 2, [2, 0]), (FHUB, 4, [4, 0])])
def test_bookkeeping_matches_oracle(model, N, QN):
    assert np.array_equal(im.i2k(model, N, QN), orc.i2k(model, N, QN))
    assert np.array_equal(im.validpath(model, N, QN), orc.validpath(model, N, QN))
    dims = [2, 4, 3] if model == XXZ else [2, 3]
    k0, s0 = im.dimmap(model, N, QN, dims)
    ob = orc.make_dims_model(model, dims, N, QN)
    assert np.array_equal(k0, ob[0][0]) and np.array_equal(s0, ob[0][1])
    bonds = im.make_dims(model, N, QN, k0, s0)
    for (ka, sa), (kb, sb) in zip(bonds, ob):
        assert np.array_equal(ka, kb) and np.array_equal(sa, sb)
    st = im.mps_structure(model, N, QN, bonds[0], bonds[1 % N])
    so = orc.mps_structure(model, N, QN, ob[0], ob[1 % N])
    assert np.array_equal(st.nblk, so.nblk) and np.array_equal(st.qin, so.qin) and np.array_equal(st.qout, so.qout)
    assert np.array_equal(st.rows, so.rows) and np.array_equal(st.cols, so.cols)


def test_hamiltonians_match_oracle():
    for model, p in [(XXZ, [1.0, -1.0, 0.3]), (XXZ, [0.7, 1.5, 0.0]), (FHUB, [1.0, 4.0, 0.2, 0.1, 0.05])]:
        assert np.array_equal(im.model_ham(model, p), orc.model_ham(model, p))


def test_known_answers_product():
    assert im.i2k(FHUB, 2, [2, 0]).tolist() == [[-1, 0], [0, -1], [0, 1], [1, 0]]
    assert im.validpath(FHUB, 2, [2, 0]).tolist() == [3, 0]
    k, s = im.dimmap(XXZ, 2, [0], [2, 4, 2])
    assert k.ravel().tolist() == [-2, 0, 2] and s.tolist() == [2, 4, 2]


def test_invalid_layouts_rejected():
    # blocks out of ascending order
    bad = im.BlockLayout(1, [2], [[2], [0]], [[2], [0]], [3, 3], [3, 3])
    ok = im.BlockLayout.diag(1, [[0], [2]], [3, 3])
    A = im.BlockLayout(1, [1, 1], [[0], [2]], [[2], [0]], [3, 3], [3, 3])
    with pytest.raises(im.ImpsError) as e:
        im.TMPlan(None, im.LEFT, A, None, bad, ok)
    assert e.value.status == 1
    # unknown precision
    with pytest.raises(im.ImpsError):
        im.TMPlan(None, im.LEFT, A, None, ok, ok, prec=7)
    # matched sectors with disagreeing sizes
    A2 = im.BlockLayout(1, [1, 1], [[0], [2]], [[2], [0]], [4, 3], [3, 3])
    with pytest.raises(im.ImpsError):
        im.TMPlan(None, im.LEFT, A2, None, ok, ok)


def _layout(ob):
    return im.BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def test_plan_costs_are_algorithmic_flops():
    k, s = xxz_bond0(9, 8)
    c = Cell(XXZ, 2, [0], k, s, seed=3)
    A = _layout(c.struct[0])
    x = _layout(c.diag(0))
    y = _layout(c.diag(1))
    p = im.TMPlan(None, im.LEFT, A, None, x, y)
    flops, bytes_ = p.cost()
    # sum over blocks q -> q' of 2 m_q m_q' (m_q + m_q')  (x A first, then A^T (x A))
    ref = sum(2.0 * r * cc * (r + cc) for r, cc in zip(A.rows, A.cols))
    assert flops == ref
    assert bytes_ == 8.0 * (A.nelem + x.nelem + y.nelem)
    launches, tiles = p.shape()
    assert launches == 2 and tiles == len(A.rows) + len(y.rows)


def test_redop_output_layout_matches_oracle():
    for model, (k, s), QN, Hp in [(XXZ, xxz_bond0(7, 6), [0], [1.0, -1.0, 0.2]), (FHUB, fhub_bond0(3, 4), [2, 0], [1.0, 4.0, 0.0, 0.0, 0.0])]:
        c = Cell(model, 2, QN, k, s, seed=4)
        H = orc.model_ham(model, Hp)
        for direction, A in [(0, c.AL[1]), (1, c.AR[1])]:
            ref = orc.reduced_op(H, c.d, A, direction)
            p = im.RedOpPlan(None, H, c.d, _layout(A), direction)
            lay = p.out_layout
            assert np.array_equal(lay.nblk, ref.nblk)
            assert np.array_equal(lay.qin, ref.qin) and np.array_equal(lay.qout, ref.qout)
            assert np.array_equal(lay.rows, ref.rows) and np.array_equal(lay.cols, ref.cols)
