"""MI355X block-sparse iMPS operator path.

Python view of the C ABI in include/imps_capi.h (libimps_hip.so, built in-tree by
``imps_cpp_tools_amd.build``).  The operator applications run as grouped HIP kernels on gfx950;
there is no CPU fallback: loading fails loudly when the library is missing, and every apply
needs a GPU context.

Layouts follow the reference's Vectorize order (BlockObj.hpp:764-780, MPSBlockMat.hpp:446-465).
Device buffers are torch float64 tensors on ``cuda:N``; torch is used only for device memory.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libimps_hip.so")

FP64, FP32, BF16X3 = 0, 1, 2
LEFT, RIGHT = 0, 1
PREC_NAMES = {FP64: "fp64", FP32: "fp32", BF16X3: "bf16x3"}
XXZ, FHUB, QIM = 0, 1, 2


def QPOTTS(q: int) -> int:
    """Model id of the q-state quantum Potts chain (Z(q) charges)."""
    return 3 | (int(q) << 8)


def BHUB(d: int) -> int:
    """Model id of the Bose-Hubbard chain with local dimension d (U(1) charge)."""
    return 4 | (int(d) << 8)


STATUS = {0: "OK", 1: "INVALID_ARG", 2: "NOT_CONVERGED", 3: "NUMERICAL", 4: "HIP_ERROR", 5: "NO_DEVICE"}


class ImpsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class _CLayout(ctypes.Structure):
    _fields_ = [("nsym", ctypes.c_int), ("nlist", ctypes.c_int), ("nblk", ctypes.POINTER(ctypes.c_int)),
                ("qin", ctypes.POINTER(ctypes.c_int)), ("qout", ctypes.POINTER(ctypes.c_int)),
                ("rows", ctypes.POINTER(ctypes.c_int)), ("cols", ctypes.POINTER(ctypes.c_int)),
                ("periods", ctypes.POINTER(ctypes.c_int))]


_lib = None


def lib():
    """Load libimps_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m imps_cpp_tools_amd.build`")
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as the one this
        # library is linked against; whichever is loaded first serves the whole process.  Load
        # torch's first so that torch (which owns the device buffers and streams) and this library
        # share one runtime; torch cannot start on a runtime other than its own.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        P, I, D, V = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
        LP = ctypes.POINTER(_CLayout)
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        sig = {
            "imps_last_error": (ctypes.c_char_p, []),
            "imps_version": (ctypes.c_char_p, []),
            "imps_create": (I, [ctypes.POINTER(P), I]),
            "imps_destroy": (I, [P]),
            "imps_set_stream": (I, [P, P]),
            "imps_synchronize": (I, [P]),
            "imps_plan_destroy": (I, [P]),
            "imps_plan_cost": (I, [P, dp, dp]),
            "imps_plan_timing": (I, [P, I, dp, ip]),
            "imps_plan_schedule": (I, [P, I, ip, ip]),
            "imps_plan_shape": (I, [P, ip, ip]),
            "imps_plan_out_layout": (I, [P, ip, ip, ctypes.POINTER(ctypes.c_long), ip, ip, ip, ip, ip]),
            "imps_tm_plan": (I, [P, ctypes.POINTER(P), I, LP, LP, LP, LP, I]),
            "imps_tm_set_mps": (I, [P, V, V]),
            "imps_tm_apply": (I, [P, V, V, I]),
            "imps_hac_plan": (I, [P, ctypes.POINTER(P), LP, LP, LP, LP, LP, I]),
            "imps_hac_set_operator": (I, [P, V, V, V, V]),
            "imps_hac_apply": (I, [P, V, V, I]),
            "imps_hac_apply_static": (I, [P, V, V, I]),
            "imps_hc_plan": (I, [P, ctypes.POINTER(P), LP, LP, LP, LP, LP, I]),
            "imps_hc_set_operator": (I, [P, V, V, V, V]),
            "imps_hc_apply": (I, [P, V, V, I]),
            "imps_redop_plan": (I, [P, ctypes.POINTER(P), dp, I, LP, I, I]),
            "imps_redop_apply": (I, [P, V, V]),
            "imps_env_plan": (I, [P, ctypes.POINTER(P), I, I, LP, LP, I]),
            "imps_env_set_mps": (I, [P, ctypes.POINTER(V)]),
            "imps_env_solve": (I, [P, V, V, V, V, V, D, I, I, ip, dp]),
            "imps_eigsh": (I, [P, P, V, D, I, I, dp, ip, dp]),
            "imps_env_tm_apply": (I, [P, V, V]),
            "imps_env_fixed_point": (I, [P, V, D, I, ip, dp]),
            "imps_prod_plan": (I, [P, ctypes.POINTER(P), I, LP, LP]),
            "imps_prod_apply": (I, [P, V, V, V]),
            "imps_prod_apply_n": (I, [P, V, V, V, I]),
            "imps_qr_plan": (I, [P, ctypes.POINTER(P), LP, I]),
            "imps_qr_apply": (I, [P, V, V, V]),
            "imps_tm_vjp": (I, [P, V, V, V, V, V, V, V]),
            "imps_hac_vjp": (I, [P, V, V, V, V, V, V, V, V, V, V, V]),
            "imps_hc_vjp": (I, [P, V, V, V, V, V, V, V, V, V, V, V]),
            "imps_heff_plan": (I, [P, ctypes.POINTER(P), I, I, LP, LP, LP, LP, dp, D, ip, I, I]),
            "imps_opeff_plan": (I, [P, ctypes.POINTER(P), I, I, LP, LP, LP, LP, dp, D, ip, I, I]),
            "imps_heff_set_state": (I, [P, ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V), D]),
            "imps_heff_set_state_lr": (I, [P, ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V), LP, V, LP, V,
                                           LP, V, D]),
            "imps_heff_check": (I, [P, dp, dp, ip]),
            "imps_heff_apply": (I, [P, dp, V, V, ip, dp]),
            "imps_heff_eigs": (I, [P, dp, I, D, I, I, dp, V, dp, ip]),
            "imps_model_nsym": (I, [I]),
            "imps_model_d": (I, [I]),
            "imps_model_periods": (I, [I, ip]),
            "imps_model_i2k": (I, [I, I, ip, ip]),
            "imps_model_validpath": (I, [I, I, ip, ip]),
            "imps_model_dimmap": (I, [I, I, ip, ip, I, ip, ip]),
            "imps_model_make_dims": (I, [I, I, ip, I, ip, ip, ip, ip]),
            "imps_model_ham": (I, [I, dp, dp]),
            "imps_mps_structure": (I, [I, I, ip, I, ip, ip, I, ip, ip, ip, ip, ip, ip, ip]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int):
    if status != 0:
        raise ImpsError(status, lib().imps_last_error().decode())


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ----------------------------------------------------------------------------------------------
@dataclass
class BlockLayout:
    """Structure of a block-sparse object: nlist entries (1 = BlockMat/BlockDiagMat, d = MPS,
    d*d = RedOp), each a list of blocks in ascending ingoing charge.  `periods`: the symmetry group
    (per charge 0 for U(1), N for Z(N)); None means U(1) for every charge."""
    nsym: int
    nblk: np.ndarray
    qin: np.ndarray
    qout: np.ndarray
    rows: np.ndarray
    cols: np.ndarray
    periods: Optional[np.ndarray] = None

    def __post_init__(self):
        per = np.zeros(self.nsym, np.int32) if self.periods is None else np.ascontiguousarray(self.periods, dtype=np.int32).reshape(-1)
        if per.shape != (self.nsym,):
            raise ValueError(f"periods: {per.size} entries for {self.nsym} charges")
        self.periods = np.where(per == 1, 0, per).astype(np.int32)  # period 1 is U(1), as in symobj::init
        self.nblk = np.ascontiguousarray(self.nblk, dtype=np.int32).reshape(-1)
        nb = int(self.nblk.sum())
        self.qin = np.ascontiguousarray(self.qin, dtype=np.int32).reshape(nb, self.nsym)
        self.qout = np.ascontiguousarray(self.qout, dtype=np.int32).reshape(nb, self.nsym)
        self.rows = np.ascontiguousarray(self.rows, dtype=np.int32).reshape(nb)
        self.cols = np.ascontiguousarray(self.cols, dtype=np.int32).reshape(nb)

    @property
    def nlist(self) -> int:
        return len(self.nblk)

    def _sized(self):
        """(block sizes, element count, block offsets), computed once per rows / cols arrays
        (read-only arrays: layouts are shared)."""
        src = (id(self.rows), id(self.cols))
        c = self.__dict__.get("_size_cache")
        if c is None or c[0] != src:
            sz = self.rows.astype(np.int64) * self.cols.astype(np.int64)
            off = np.concatenate([[0], np.cumsum(sz)[:-1]]).astype(np.int64)
            sz.setflags(write=False)
            off.setflags(write=False)
            c = (src, sz, int(sz.sum()), off)
            self._size_cache = c
        return c

    @property
    def sizes(self) -> np.ndarray:
        return self._sized()[1]

    @property
    def nelem(self) -> int:
        return self._sized()[2]

    @property
    def offsets(self) -> np.ndarray:
        return self._sized()[3]

    def entry(self, e: int) -> "BlockLayout":
        s = int(self.nblk[:e].sum())
        n = int(self.nblk[e])
        return BlockLayout(self.nsym, [n], self.qin[s:s + n], self.qout[s:s + n], self.rows[s:s + n], self.cols[s:s + n], self.periods)

    def entry_range(self, e: int):
        """Element range [a, b) of list entry e in the flat buffer."""
        s = int(self.nblk[:e].sum())
        n = int(self.nblk[e])
        a = int(self.sizes[:s].sum())
        return a, a + int(self.sizes[s:s + n].sum())

    def c(self) -> _CLayout:
        self._keep = (self.nblk, self.qin, self.qout, self.rows, self.cols, self.periods)
        return _CLayout(self.nsym, self.nlist, _iptr(self.nblk), _iptr(self.qin), _iptr(self.qout), _iptr(self.rows), _iptr(self.cols),
                        _iptr(self.periods))

    def __eq__(self, o):
        return (self.nsym == o.nsym and np.array_equal(self.periods, o.periods) and np.array_equal(self.nblk, o.nblk)
                and np.array_equal(self.qin, o.qin) and np.array_equal(self.qout, o.qout) and np.array_equal(self.rows, o.rows)
                and np.array_equal(self.cols, o.cols))

    @staticmethod
    def concat(parts: Sequence["BlockLayout"]) -> "BlockLayout":
        return BlockLayout(parts[0].nsym, np.concatenate([p.nblk for p in parts]), np.concatenate([p.qin for p in parts]),
                           np.concatenate([p.qout for p in parts]), np.concatenate([p.rows for p in parts]),
                           np.concatenate([p.cols for p in parts]), parts[0].periods)

    @staticmethod
    def diag(nsym: int, keys: np.ndarray, sizes: np.ndarray, periods=None) -> "BlockLayout":
        keys = np.asarray(keys, dtype=np.int32).reshape(-1, nsym)
        return BlockLayout(nsym, [len(keys)], keys, keys, sizes, sizes, periods)


# ----------------------------------------------------------------------------------------------
# bookkeeping (host only)
def _qn(QN) -> np.ndarray:
    """The quantum numbers as int32 (the Z(q) models take none: one placeholder entry)."""
    q = np.asarray(list(QN), np.int32).reshape(-1)
    return q if q.size else np.zeros(1, np.int32)


def model_nsym(model: int) -> int:
    return lib().imps_model_nsym(model)


def model_d(model: int) -> int:
    return lib().imps_model_d(model)


def model_periods(model: int) -> np.ndarray:
    """The model's symmetry group: per charge 0 (U(1)) or N (Z(N))."""
    out = np.zeros(4, np.int32)
    n = lib().imps_model_periods(model, _iptr(out))
    if n < 0:
        raise ImpsError(1, "imps_model_periods: unknown model")
    return out[:n].copy()


def i2k(model: int, N: int, QN: Sequence[int]) -> np.ndarray:
    ns, d = model_nsym(model), model_d(model)
    out = np.zeros(d * ns, np.int32)
    q = _qn(QN)
    r = lib().imps_model_i2k(model, N, _iptr(q), _iptr(out))
    if r < 0:
        raise ImpsError(1, "imps_model_i2k failed")
    return out.reshape(d, ns)


def validpath(model: int, N: int, QN: Sequence[int]) -> np.ndarray:
    out = np.zeros(N, np.int32)
    q = _qn(QN)
    if lib().imps_model_validpath(model, N, _iptr(q), _iptr(out)) < 0:
        raise ImpsError(1, "imps_model_validpath failed")
    return out


def dimmap(model: int, N: int, QN: Sequence[int], dims: Sequence[int]):
    ns = model_nsym(model)
    nmax = len(dims) ** (2 if model == FHUB else 1)
    keys = np.zeros(nmax * ns, np.int32)
    sizes = np.zeros(nmax, np.int32)
    q = _qn(QN)
    dd = np.asarray(dims, np.int32)
    n = lib().imps_model_dimmap(model, N, _iptr(q), _iptr(dd), len(dd), _iptr(keys), _iptr(sizes))
    if n < 0:
        raise ImpsError(1, "imps_model_dimmap failed")
    return keys[:n * ns].reshape(n, ns), sizes[:n]


def make_dims(model: int, N: int, QN: Sequence[int], keys0: np.ndarray, sizes0: np.ndarray):
    """MakeDims from a bond-0 dim map; returns a list of N (keys, sizes) bonds."""
    ns = model_nsym(model)
    k0 = np.ascontiguousarray(keys0, np.int32).reshape(-1)
    s0 = np.ascontiguousarray(sizes0, np.int32)
    nsec = len(s0)
    keys = np.zeros(N * nsec * ns, np.int32)
    sizes = np.zeros(N * nsec, np.int32)
    q = _qn(QN)
    if lib().imps_model_make_dims(model, N, _iptr(q), nsec, _iptr(k0), _iptr(s0), _iptr(keys), _iptr(sizes)) < 0:
        raise ImpsError(1, "imps_model_make_dims failed")
    keys = keys.reshape(N, nsec, ns)
    sizes = sizes.reshape(N, nsec)
    return [(keys[k], sizes[k]) for k in range(N)]


def model_ham(model: int, params: Sequence[float]) -> np.ndarray:
    d = model_d(model)
    p = np.asarray(params, np.float64)
    H = np.zeros(d * d * d * d, np.float64)
    if lib().imps_model_ham(model, _dptr(p), _dptr(H)) < 0:
        raise ImpsError(1, "imps_model_ham failed")
    return H.reshape(d * d, d * d, order="F")


def mps_structure(model: int, N: int, QN: Sequence[int], left, right) -> BlockLayout:
    ns, d = model_nsym(model), model_d(model)
    kl, sl = (np.ascontiguousarray(a, np.int32) for a in left)
    kr, sr = (np.ascontiguousarray(a, np.int32) for a in right)
    cap = d * len(sl)
    nblk = np.zeros(d, np.int32)
    qin = np.zeros(cap * ns, np.int32)
    qout = np.zeros(cap * ns, np.int32)
    rows = np.zeros(cap, np.int32)
    cols = np.zeros(cap, np.int32)
    q = _qn(QN)
    n = lib().imps_mps_structure(model, N, _iptr(q), len(sl), _iptr(kl.reshape(-1)), _iptr(sl), len(sr), _iptr(kr.reshape(-1)),
                                 _iptr(sr), _iptr(nblk), _iptr(qin), _iptr(qout), _iptr(rows), _iptr(cols))
    if n < 0:
        raise ImpsError(1, "imps_mps_structure failed")
    return BlockLayout(ns, nblk, qin[:n * ns], qout[:n * ns], rows[:n], cols[:n], model_periods(model))


# ----------------------------------------------------------------------------------------------
# device context and plans
def _ptr(t, n: Optional[int] = None, device: Optional[int] = None, what: str = "buffer") -> Optional[int]:
    """Device address of a contiguous float64 CUDA tensor.  With n, the tensor must hold at least n
    elements; with device, it must live on that GPU: the library reads and writes exactly the
    extents its plans derive, so a short or foreign buffer is rejected here, before any launch."""
    if t is None:
        return None
    if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 8 and t.is_contiguous()):
        raise ValueError(f"{what}: device buffers must be contiguous float64 CUDA tensors")
    if n is not None and t.numel() < n:
        raise ValueError(f"{what}: {t.numel()} elements, the plan needs {n}")
    if device is not None and t.device.index != device:
        raise ValueError(f"{what}: tensor on cuda:{t.device.index}, the plan's context is on cuda:{device}")
    return t.data_ptr()


class Context:
    """A device context (one GPU, one default-priority stream)."""

    def __init__(self, device: int = 0, stream=None):
        """`stream`: a torch.cuda.Stream (or raw hipStream_t int); default: torch's current stream,
        so that library launches are ordered with torch's own copies and kernels."""
        h = ctypes.c_void_p()
        _check(lib().imps_create(ctypes.byref(h), device))
        self.h = h
        self.device = device
        if stream is None:
            import torch
            stream = torch.cuda.current_stream(device)
        self.set_stream(stream)

    def set_stream(self, stream):
        raw = stream if isinstance(stream, int) else stream.cuda_stream
        _check(lib().imps_set_stream(self.h, ctypes.c_void_p(raw)))

    def synchronize(self):
        _check(lib().imps_synchronize(self.h))

    def close(self):
        if self.h:
            lib().imps_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    kind = ""

    def __init__(self, handle, ctx: Optional["Context"] = None):
        self.h = handle
        self.device = ctx.device if ctx is not None else None

    def _p(self, t, n: int, what: str):
        return _ptr(t, n, self.device, f"{self.kind} {what}")

    def _out_layout(self, nsym: int, periods=None) -> BlockLayout:
        nl, nb, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        z = ctypes.POINTER(ctypes.c_int)()
        _check(lib().imps_plan_out_layout(self.h, ctypes.byref(nl), ctypes.byref(nb), ctypes.byref(ne), z, z, z, z, z))
        n = nb.value
        nblk = np.zeros(nl.value, np.int32)
        qin = np.zeros(max(n, 1) * nsym, np.int32)
        qout = np.zeros(max(n, 1) * nsym, np.int32)
        rows = np.zeros(max(n, 1), np.int32)
        cols = np.zeros(max(n, 1), np.int32)
        _check(lib().imps_plan_out_layout(self.h, None, None, None, _iptr(nblk), _iptr(qin), _iptr(qout), _iptr(rows), _iptr(cols)))
        return BlockLayout(nsym, nblk, qin[:n * nsym], qout[:n * nsym], rows[:n], cols[:n], periods)

    def cost(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().imps_plan_cost(self.h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def timing(self, enable: bool):
        """GPU ms of the GEMM kernels of the applies timed since the previous call and their count;
        then switches the per-apply event timing on or off (imps_plan_timing)."""
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(lib().imps_plan_timing(self.h, int(enable), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def schedule(self, nvec: int = 1):
        """(ksplit, shared) of the panel kernels for applies of nvec vectors (imps_plan_schedule)."""
        k, sh = ctypes.c_int(), ctypes.c_int()
        _check(lib().imps_plan_schedule(self.h, nvec, ctypes.byref(k), ctypes.byref(sh)))
        return k.value, bool(sh.value)

    def shape(self):
        a, t = ctypes.c_int(), ctypes.c_int()
        _check(lib().imps_plan_shape(self.h, ctypes.byref(a), ctypes.byref(t)))
        return a.value, t.value

    def close(self):
        if self.h:
            lib().imps_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ctx_handle(ctx: Optional[Context]):
    return ctx.h if ctx is not None else None


class TMPlan(Plan):
    """Mixed / regular transfer-matrix apply (TMBlockFunctions.hpp:352-390, :494-533)."""
    kind = "tm"

    def __init__(self, ctx: Optional[Context], direction: int, A: BlockLayout, B: Optional[BlockLayout], x: BlockLayout,
                 y: BlockLayout, prec: int = FP64):
        h = ctypes.c_void_p()
        cA, cx, cy = A.c(), x.c(), y.c()
        cB = B.c() if B is not None else None
        _check(lib().imps_tm_plan(_ctx_handle(ctx), ctypes.byref(h), direction, ctypes.byref(cA),
                                  ctypes.byref(cB) if cB is not None else None, ctypes.byref(cx), ctypes.byref(cy), prec))
        super().__init__(h, ctx)
        self.mixed = B is not None
        self.xlen, self.ylen = x.nelem, y.nelem
        self.alen, self.blen = A.nelem, (B.nelem if B is not None else 0)

    def set_mps(self, A, B=None):
        _check(lib().imps_tm_set_mps(self.h, self._p(A, self.alen, "A"), self._p(B, self.blen, "B")))

    def apply(self, x, y, nvec: int = 1):
        _check(lib().imps_tm_apply(self.h, self._p(x, nvec * self.xlen, "x"), self._p(y, nvec * self.ylen, "y"), nvec))

    def vjp(self, A, B, x, ybar, xbar=None, Abar=None, Bbar=None):
        """Adjoint of the apply (imps_tm_vjp, FP64): writes the given outputs."""
        _check(lib().imps_tm_vjp(self.h, self._p(A, self.alen, "A"), self._p(B, self.blen, "B") if self.mixed else None,
                                 self._p(x, self.xlen, "x"), self._p(ybar, self.ylen, "ybar"), self._p(xbar, self.xlen, "xbar"),
                                 self._p(Abar, self.alen, "Abar"), self._p(Bbar, self.blen, "Bbar") if self.mixed else None))


class HACPlan(Plan):
    """Two-site H_AC apply (VUMPS_helpers.hpp:38-88)."""
    kind = "hac"

    def __init__(self, ctx: Optional[Context], AC: BlockLayout, EHL: BlockLayout, EHR: BlockLayout, HL: BlockLayout,
                 HR: BlockLayout, prec: int = FP64):
        h = ctypes.c_void_p()
        c = [l.c() for l in (AC, EHL, EHR, HL, HR)]
        _check(lib().imps_hac_plan(_ctx_handle(ctx), ctypes.byref(h), *[ctypes.byref(x) for x in c], prec))
        super().__init__(h, ctx)
        self.n = AC.nelem
        self.oplen = (EHL.nelem, EHR.nelem, HL.nelem, HR.nelem)

    def set_operator(self, EHL, EHR, HL, HR):
        a, b, c, d = self.oplen
        _check(lib().imps_hac_set_operator(self.h, self._p(EHL, a, "EHL"), self._p(EHR, b, "EHR"), self._p(HL, c, "HL"),
                                           self._p(HR, d, "HR")))

    def apply(self, x, y, nvec: int = 1):
        _check(lib().imps_hac_apply(self.h, self._p(x, nvec * self.n, "x"), self._p(y, nvec * self.n, "y"), nvec))

    def apply_static(self, x, y, nvec: int = 1):
        """The apply without allocation (imps_hac_apply_static: split copies in the fixed scratch)."""
        _check(lib().imps_hac_apply_static(self.h, self._p(x, nvec * self.n, "x"), self._p(y, nvec * self.n, "y"), nvec))

    def vjp(self, ops, x, ybar, xbar=None, opbars=(None, None, None, None)):
        """Adjoint of the apply (imps_hac_vjp, FP64): ops = (EHL, EHR, HL, HR), opbars their gradients."""
        names = ("EHL", "EHR", "HL", "HR")
        o = [self._p(t, n, w) for t, n, w in zip(ops, self.oplen, names)]
        g = [self._p(t, n, w + "bar") for t, n, w in zip(opbars, self.oplen, names)]
        _check(lib().imps_hac_vjp(self.h, *o, self._p(x, self.n, "x"), self._p(ybar, self.n, "ybar"), self._p(xbar, self.n, "xbar"), *g))


class HCPlan(Plan):
    """Bond H_C apply (VUMPS_helpers.hpp:91-133) in reduced form with HL = ReducedOp(H, AL, l)."""
    kind = "hc"

    def __init__(self, ctx: Optional[Context], C: BlockLayout, EHL: BlockLayout, EHR: BlockLayout, HL: BlockLayout,
                 AR: BlockLayout, prec: int = FP64):
        h = ctypes.c_void_p()
        c = [l.c() for l in (C, EHL, EHR, HL, AR)]
        _check(lib().imps_hc_plan(_ctx_handle(ctx), ctypes.byref(h), *[ctypes.byref(x) for x in c], prec))
        super().__init__(h, ctx)
        self.n = C.nelem
        self.oplen = (EHL.nelem, EHR.nelem, HL.nelem, AR.nelem)

    def set_operator(self, EHL, EHR, HL, AR):
        a, b, c, d = self.oplen
        _check(lib().imps_hc_set_operator(self.h, self._p(EHL, a, "EHL"), self._p(EHR, b, "EHR"), self._p(HL, c, "HL"),
                                          self._p(AR, d, "AR")))

    def apply(self, x, y, nvec: int = 1):
        _check(lib().imps_hc_apply(self.h, self._p(x, nvec * self.n, "x"), self._p(y, nvec * self.n, "y"), nvec))

    def vjp(self, ops, x, ybar, xbar=None, opbars=(None, None, None, None)):
        """Adjoint of the apply (imps_hc_vjp, FP64): ops = (EHL, EHR, HL, AR), opbars their gradients."""
        names = ("EHL", "EHR", "HL", "AR")
        o = [self._p(t, n, w) for t, n, w in zip(ops, self.oplen, names)]
        g = [self._p(t, n, w + "bar") for t, n, w in zip(opbars, self.oplen, names)]
        _check(lib().imps_hc_vjp(self.h, *o, self._p(x, self.n, "x"), self._p(ybar, self.n, "ybar"), self._p(xbar, self.n, "xbar"), *g))


class RedOpPlan(Plan):
    """ReducedOp(H, A, dir) (MPSBlockUtilities.hpp:39-74); output layout derived from A."""
    kind = "redop"

    def __init__(self, ctx: Optional[Context], H: np.ndarray, d: int, A: BlockLayout, direction: int, prec: int = FP64):
        h = ctypes.c_void_p()
        Hf = np.asfortranarray(H, dtype=np.float64).reshape(-1, order="F").copy()
        cA = A.c()
        _check(lib().imps_redop_plan(_ctx_handle(ctx), ctypes.byref(h), _dptr(Hf), d, ctypes.byref(cA), direction, prec))
        super().__init__(h, ctx)
        self.out_layout = self._out_layout(A.nsym, A.periods)
        self.alen = A.nelem

    def apply(self, A, y):
        _check(lib().imps_redop_apply(self.h, self._p(A, self.alen, "A"), self._p(y, self.out_layout.nelem, "y")))


class EnvPlan(Plan):
    """Environment (geometric-sum) solver over a unit cell (InvertE_proj, InvEBlockFunctions.hpp:205-266)."""
    kind = "env"

    def __init__(self, ctx: Context, direction: int, A: Sequence[BlockLayout], x: BlockLayout, prec: int = FP64):
        h = ctypes.c_void_p()
        arr = (_CLayout * len(A))(*[a.c() for a in A])
        self._A = list(A)
        cx = x.c()
        _check(lib().imps_env_plan(_ctx_handle(ctx), ctypes.byref(h), direction, len(A), arr, ctypes.byref(cx), prec))
        super().__init__(h, ctx)
        self.n = x.nelem

    def set_mps(self, tensors):
        if len(tensors) != len(self._A):
            raise ValueError(f"env: {len(tensors)} tensors for a cell of {len(self._A)}")
        arr = (ctypes.c_void_p * len(tensors))(*[self._p(t, a.nelem, f"A[{k}]") for k, (t, a) in enumerate(zip(tensors, self._A))])
        _check(lib().imps_env_set_mps(self.h, arr))

    def tm_apply(self, x, y):
        _check(lib().imps_env_tm_apply(self.h, self._p(x, self.n, "x"), self._p(y, self.n, "y")))

    def fixed_point(self, x, tol: float = 1e-14, maxit: int = 0, check: bool = True):
        it, dl = ctypes.c_int(), ctypes.c_double()
        st = lib().imps_env_fixed_point(self.h, self._p(x, self.n, "x"), tol, maxit, ctypes.byref(it), ctypes.byref(dl))
        if check:
            _check(st)
        return it.value, dl.value

    def solve(self, x, L, R, y, y0=None, tol: float = 1e-14, maxit: int = 0, krylov: int = 40, check: bool = True):
        it, rr = ctypes.c_int(), ctypes.c_double()
        n = self.n
        st = lib().imps_env_solve(self.h, self._p(x, n, "x"), self._p(L, n, "L"), self._p(R, n, "R"), self._p(y0, n, "y0"),
                                  self._p(y, n, "y"), tol, maxit, krylov, ctypes.byref(it), ctypes.byref(rr))
        if check:
            _check(st)
        return it.value, rr.value


def eigsh(plan: Plan, x, tol: float = 1e-12, maxit: int = 0, krylov: int = 20, plan_fp64: Optional[Plan] = None, check: bool = True):
    """Lowest eigenpair of an H_AC / H_C plan (eigs_rs "SA", nev = 1); x: start vector in, eigenvector out."""
    ev, it, rr = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
    st = lib().imps_eigsh(plan.h, plan_fp64.h if plan_fp64 is not None else None, plan._p(x, plan.n, "x"), tol, maxit, krylov,
                          ctypes.byref(ev), ctypes.byref(it), ctypes.byref(rr))
    if check:
        _check(st)
    return ev.value, it.value, rr.value


MAT_MPS, MPS_MAT, MPS_MATT, MATT_MPS, TM_LEFT_UNITY, TM_RIGHT_UNITY, MPS_MPS, REDOP_MPS, MPS_REDOP = range(9)


class ProdPlan(Plan):
    """Two-operand block product (FP64); output layout in .out_layout."""
    kind = "prod"

    def __init__(self, ctx: Optional[Context], kind: int, X1: BlockLayout, X2: BlockLayout):
        h = ctypes.c_void_p()
        c1, c2 = X1.c(), X2.c()
        _check(lib().imps_prod_plan(_ctx_handle(ctx), ctypes.byref(h), kind, ctypes.byref(c1), ctypes.byref(c2)))
        super().__init__(h, ctx)
        self.out_layout = self._out_layout(X1.nsym, X1.periods)
        self.len1, self.len2 = X1.nelem, X2.nelem

    def apply(self, x1, x2, y):
        _check(lib().imps_prod_apply(self.h, self._p(x1, self.len1, "x1"), self._p(x2, self.len2, "x2"),
                                     self._p(y, self.out_layout.nelem, "y")))

    def apply_n(self, x1, x2, y, nvec: int):
        """nvec first operands back to back in x1 (outputs back to back in y), one x2."""
        nvec = int(nvec)
        _check(lib().imps_prod_apply_n(self.h, self._p(x1, nvec * self.len1, "x1"), self._p(x2, self.len2, "x2"),
                                       self._p(y, nvec * self.out_layout.nelem, "y"), nvec))


class QRPlan(Plan):
    """Per-sector QR of an MPS tensor (qr(R, A, dir) with qr_pos), FP64 on the GPU; R layout in .out_layout."""
    kind = "qr"

    def __init__(self, ctx: Optional[Context], A: BlockLayout, direction: int):
        h = ctypes.c_void_p()
        cA = A.c()
        _check(lib().imps_qr_plan(_ctx_handle(ctx), ctypes.byref(h), ctypes.byref(cA), direction))
        super().__init__(h, ctx)
        self.out_layout = self._out_layout(A.nsym, A.periods)
        self.alen = A.nelem

    def apply(self, A, Q, R):
        _check(lib().imps_qr_apply(self.h, self._p(A, self.alen, "A"), self._p(Q, self.alen, "Q"),
                                   self._p(R, self.out_layout.nelem, "R")))


class HeffPlan(Plan):
    """Excitation effective Hamiltonian H_eff(k) for nmom momenta at once (ApplyHeff with its
    constants, Excitations_helpers.hpp:20-347): imps_heff_plan / set_state / apply / eigs.

    AL, AR, NL: per-site MPS layouts; C: per-site bond layouts (C[n] right of site n); H: dense
    d^2 x d^2 two-site Hamiltonian; E0: ground-state energy per site (subtracted from H); K: charge;
    prec: FP64, or FP32 / BF16X3 transfer-matrix chains in the geometric sums' Arnoldi steps under
    FP64 restarts.  Excitation vectors are complex128 CUDA tensors of shape (nmom, mtot) in XDims order.
    With `observable` (a d x d one-site operator; H is then ignored and E0 is the observable's
    ground-state expectation) the plan is the effective operator of that observable
    (imps_opeff_plan: OPeffConstants / ApplyOPeff); `.apply` applies it and `.measure` gives
    MeasureExcitations' per-site values."""
    kind = "heff"

    def __init__(self, ctx: Context, AL: Sequence[BlockLayout], AR: Sequence[BlockLayout], NL: Sequence[BlockLayout],
                 C: Sequence[BlockLayout], H: Optional[np.ndarray], E0: float, K, nmom: int = 1, prec: int = FP64,
                 observable: Optional[np.ndarray] = None):
        h = ctypes.c_void_p()
        N = len(AL)
        arr = [(_CLayout * N)(*[x.c() for x in L]) for L in (AL, AR, NL, C)]
        self._keep_layouts = (list(AL), list(AR), list(NL), list(C))
        op = H if observable is None else observable
        Hf = np.asfortranarray(op, dtype=np.float64).reshape(-1, order="F").copy()
        Kq = np.ascontiguousarray(np.atleast_1d(K), np.int32)
        entry = lib().imps_heff_plan if observable is None else lib().imps_opeff_plan
        _check(entry(ctx.h, ctypes.byref(h), N, AL[0].nlist, *arr, _dptr(Hf), float(E0), _iptr(Kq), int(nmom), int(prec)))
        super().__init__(h, ctx)
        self.N, self.nmom = N, int(nmom)
        self.xdims = self._out_layout(AL[0].nsym, AL[0].periods)
        self.mtot = self.xdims.nelem

    def set_state(self, AL, AR, NL, C, inv_tol: float = 1e-14, L=None, LM=None, RM=None):
        """Copies the state and computes the constants.  With AL C = C AR the bond objects are taken in
        closed form; for a transformed or shifted AR pass L, LM and RM as (layout, device tensor)
        pairs (imps_heff_set_state_lr; `lr_overlaps` and `apply_transformation` in
        imps_cpp_tools_amd.transform make them)."""
        lays = self._keep_layouts

        def arr(ts, ls, what):
            if len(ts) != self.N:
                raise ValueError(f"heff {what}: {len(ts)} tensors for a cell of {self.N}")
            return (ctypes.c_void_p * self.N)(*[self._p(t, l.nelem, f"{what}[{k}]") for k, (t, l) in enumerate(zip(ts, ls))])
        a = (arr(AL, lays[0], "AL"), arr(AR, lays[1], "AR"), arr(NL, lays[2], "NL"), arr(C, lays[3], "C"))
        given = [x is not None for x in (L, LM, RM)]
        if not any(given):
            _check(lib().imps_heff_set_state(self.h, *a, float(inv_tol)))
            return
        if not all(given):
            raise ValueError("heff: L, LM and RM are given together")
        cl = [x[0].c() for x in (L, LM, RM)]
        bond = []
        for c, (lay, t), what in zip(cl, (L, LM, RM), ("L", "LM", "RM")):
            if lay.nlist != 1:
                raise ValueError(f"heff {what}: a bond object has nlist = 1")
            bond += [ctypes.byref(c), self._p(t, lay.nelem, what)]
        _check(lib().imps_heff_set_state_lr(self.h, *a, *bond, float(inv_tol)))

    def check(self):
        """(tr(HL R), tr(L HR), environment-solver applications) of HeffConstants' check."""
        a, b, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(lib().imps_heff_check(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(it)))
        return a.value, b.value, it.value

    def _cp(self, t, n, what):
        import torch
        if t.dtype != torch.complex128:
            raise ValueError(f"heff {what}: complex128 tensors are needed")
        return _ptr(torch.view_as_real(t), 2 * n, self.device, f"heff {what}")

    def _moms(self, p):
        pv = np.ascontiguousarray(np.atleast_1d(p), np.float64)
        if pv.size != self.nmom:
            raise ValueError(f"heff: {pv.size} momenta for a plan of {self.nmom}")
        return pv

    def apply(self, p, x, y):
        """y[t] = H_eff(p[t]) x[t]; returns (batched GMRES steps, worst GMRES relative residual)."""
        pv = self._moms(p)
        n = self.nmom * self.mtot
        st, rr = ctypes.c_int(), ctypes.c_double()
        _check(lib().imps_heff_apply(self.h, _dptr(pv), self._cp(x, n, "x"), self._cp(y, n, "y"), ctypes.byref(st), ctypes.byref(rr)))
        return st.value, rr.value

    def measure(self, p, x):
        """MeasureExcitations' values (Excitations_helpers.hpp:819-820) of an observable plan:
        <x[t] restricted to site n | (O_eff x[t]) restricted to site n> as an (nmom, N) array."""
        import torch
        y = torch.empty_like(x)
        self.apply(p, x, y)
        out = np.zeros((self.nmom, self.N), complex)
        for n in range(self.N):
            a, b = self.xdims.entry_range(n)
            out[:, n] = (x[:, a:b].conj() * y[:, a:b]).sum(1).cpu().numpy()
        return out

    def eigs(self, p, nev: int = 4, tol: float = 1e-10, maxit: int = 0, krylov: int = 0, vectors: bool = False, check: bool = True):
        """Lowest nev eigenpairs per momentum: (energies (nmom, nev), vectors (nmom, nev, mtot) or
        None, residual norms (nmom, nev), batched applications)."""
        import torch
        pv = self._moms(p)
        ev = np.zeros((self.nmom, nev))
        rs = np.zeros((self.nmom, nev))
        vec = torch.empty((self.nmom, nev, self.mtot), dtype=torch.complex128, device=f"cuda:{self.device}") if vectors else None
        na = ctypes.c_int()
        st = lib().imps_heff_eigs(self.h, _dptr(pv), int(nev), float(tol), int(maxit), int(krylov), _dptr(ev),
                                  self._cp(vec, self.nmom * nev * self.mtot, "vectors") if vectors else None, _dptr(rs), ctypes.byref(na))
        if check:
            _check(st)
        return ev, vec, rs, na.value
