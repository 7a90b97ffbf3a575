"""Host-side argument checks of the Python plan wrappers (no GPU needed).

The wrappers pass raw device addresses to the C ABI, whose kernels touch exactly the extents a plan
derives from its layouts; a buffer that is too short, or that lives on another GPU, must be refused
before any launch.  A stand-in object with the tensor attributes the wrappers read lets these checks
run without a device."""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
from oracle import pyoracle as orc
from states import XXZ, Cell


class FakeDev:
    """Minimal stand-in for a contiguous float64 CUDA tensor (never dereferenced)."""

    class _D:
        def __init__(self, i):
            self.index = i

    def __init__(self, n, device=0):
        import torch
        self.n = n
        self.dtype = torch.float64
        self.is_cuda = True
        self.device = self._D(device)

    def element_size(self):
        return 8

    def is_contiguous(self):
        return True

    def numel(self):
        return self.n

    def data_ptr(self):
        return 0x1000


def _lay(o):
    return im.BlockLayout(o.nsym, o.nblk, o.qin, o.qout, o.rows, o.cols)


def test_ptr_checks_size_and_device():
    t = FakeDev(10, device=1)
    assert im._ptr(t, 10, 1) == 0x1000
    with pytest.raises(ValueError, match="needs 11"):
        im._ptr(t, 11, 1)
    with pytest.raises(ValueError, match="cuda:0"):
        im._ptr(t, 10, 0)
    assert im._ptr(None, 5, 0) is None


def test_plan_wrappers_refuse_short_buffers():
    c = Cell(XXZ, 2, [0], np.array([[-2], [0], [2]], np.int32), np.array([2, 4, 2], np.int32), seed=3)
    x, y = c.diag(0), c.diag(1)
    p = im.TMPlan(None, im.LEFT, _lay(c.AL[0]), None, _lay(x), _lay(y), im.FP64)
    with pytest.raises(ValueError, match="tm x"):
        p.apply(FakeDev(x.nelem - 1), FakeDev(y.nelem))
    with pytest.raises(ValueError, match="tm y"):
        p.apply(FakeDev(2 * x.nelem), FakeDev(2 * y.nelem - 1), nvec=2)
    with pytest.raises(ValueError, match="tm A"):
        p.set_mps(FakeDev(c.AL[0].nelem - 1))
    H = orc.model_ham(XXZ, [1.0, 1.0, 0.0])
    HL = orc.reduced_op(H, 2, c.AL[1], 0)
    HR = orc.reduced_op(H, 2, c.AR[1], 1)
    AC = c.struct[0]
    h = im.HACPlan(None, _lay(AC), _lay(c.diag(0)), _lay(c.diag(1)), _lay(HL), _lay(HR), im.FP64)
    with pytest.raises(ValueError, match="hac x"):
        h.apply(FakeDev(3 * AC.nelem - 1), FakeDev(3 * AC.nelem), nvec=3)
    with pytest.raises(ValueError, match="hac HR"):
        h.set_operator(FakeDev(c.diag(0).nelem), FakeDev(c.diag(1).nelem), FakeDev(HL.nelem), FakeDev(HR.nelem - 1))
    q = im.ProdPlan(None, im.MPS_MAT, _lay(c.AL[0]), _lay(y))
    with pytest.raises(ValueError, match="prod y"):
        q.apply_n(FakeDev(2 * c.AL[0].nelem), FakeDev(y.nelem), FakeDev(2 * q.out_layout.nelem - 1), 2)
