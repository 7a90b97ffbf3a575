"""ctypes binding of liboracle.so (the FP64 CPU oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
the cpu_baseline leg of bench.py.  Build with ``make -C oracle``.

Objects are lists of BlockMats with a flat FP64 data array in the reference's Vectorize order.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        ip = ctypes.POINTER(ctypes.c_int)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(ctypes.c_void_p)
        sig = {
            "orc_last_error": (ctypes.c_char_p, []),
            "orc_new": (P, [I, I, ip, ip, ip, ip, ip, dp]),
            "orc_free": (None, [P]),
            "orc_nlist": (I, [P]),
            "orc_nsym": (I, [P]),
            "orc_counts": (ctypes.c_long, [P, ip]),
            "orc_export": (None, [P, ip, ip, ip, ip, dp]),
            "orc_tm_mixed": (P, [I, P, P, P, P]),
            "orc_tm_ret": (P, [I, P, P]),
            "orc_tm_chain": (P, [I, pp, I, P, P]),
            "orc_tm_unity": (P, [I, P, P]),
            "orc_mps_mul": (P, [P, P]),
            "orc_mat_mps": (P, [P, P]),
            "orc_mps_mat": (P, [P, P]),
            "orc_bmul": (P, [P, P]),
            "orc_btrans": (P, [P]),
            "orc_apply_operator": (P, [P, dp, I]),
            "orc_reduced_op": (P, [dp, I, P, I]),
            "orc_op_tm_diag": (P, [I, dp, I, P]),
            "orc_get_h": (P, [I, pp, I, dp, I]),
            "orc_hac": (P, [P, P, P, P, P]),
            "orc_hac_plan": (P, [P, P, P, P, P]),
            "orc_hac_plan_free": (None, [P]),
            "orc_hac_plan_nx": (ctypes.c_long, [P]),
            "orc_hac_plan_apply": (None, [P, dp, dp, I]),
            "orc_max_threads": (I, []),
            "orc_hc": (P, [P, P, dp, I, P, P, P]),
            "orc_invert_e_proj": (P, [I, pp, I, P, P, P, D, I, P, ip]),
            "orc_mps_qr": (P, [P, I, pp]),
            "orc_nullspace": (P, [P, I]),
            "orc_gradnorm_right": (D, [P, P, P, P, P, P]),
            "orc_gradnorm_left": (D, [P, P, P, P, P, P]),
            "orc_i2k": (I, [I, I, ip, ip]),
            "orc_validpath": (I, [I, I, ip, ip]),
            "orc_make_dims_model": (I, [I, ip, I, I, ip, ip, ip]),
            "orc_make_dims": (I, [I, I, ip, I, I, ip, ip, ip, ip]),
            "orc_mps_structure": (P, [I, I, ip, I, I, ip, ip, I, ip, ip]),
            "orc_model_ham": (None, [I, dp, dp]),
            "orc_set_periods": (None, [I, ip]),
            "orc_model_d": (I, [I]),
            "orc_model_nsym": (I, [I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class OB:
    """Block-sparse object: structure arrays + flat data; `periods` is the symmetry group (per
    charge 0 = U(1), N = Z(N); None = U(1)), made the oracle's current group by every operation on
    the object."""
    nsym: int
    nblk: np.ndarray
    qin: np.ndarray
    qout: np.ndarray
    rows: np.ndarray
    cols: np.ndarray
    data: np.ndarray
    periods: object = None

    @staticmethod
    def like(layout, data=None) -> "OB":
        n = int((np.asarray(layout.rows, np.int64) * np.asarray(layout.cols, np.int64)).sum())
        dat = np.zeros(n) if data is None else np.ascontiguousarray(data, np.float64).reshape(-1)
        assert dat.size == n
        return OB(layout.nsym, np.asarray(layout.nblk, np.int32), np.asarray(layout.qin, np.int32).reshape(-1, layout.nsym),
                  np.asarray(layout.qout, np.int32).reshape(-1, layout.nsym), np.asarray(layout.rows, np.int32),
                  np.asarray(layout.cols, np.int32), dat, getattr(layout, "periods", None))

    @property
    def nelem(self):
        return self.data.size

    def handle(self):
        set_group(self.nsym, self.periods)
        nb = np.ascontiguousarray(self.nblk, np.int32)
        qi = np.ascontiguousarray(self.qin, np.int32).reshape(-1)
        qo = np.ascontiguousarray(self.qout, np.int32).reshape(-1)
        r = np.ascontiguousarray(self.rows, np.int32)
        c = np.ascontiguousarray(self.cols, np.int32)
        d = np.ascontiguousarray(self.data, np.float64)
        return _H(lib().orc_new(self.nsym, len(nb), _ip(nb), _ip(qi), _ip(qo), _ip(r), _ip(c), _dp(d)))

    def blocks(self):
        """Yield (entry, qin tuple, qout tuple, matrix view) in flat order."""
        pos = 0
        b = 0
        for e, n in enumerate(self.nblk):
            for _ in range(int(n)):
                r, c = int(self.rows[b]), int(self.cols[b])
                yield e, tuple(self.qin[b]), tuple(self.qout[b]), self.data[pos:pos + r * c].reshape(c, r).T
                pos += r * c
                b += 1


class _H:
    def __init__(self, h):
        if not h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        self.h = h

    def __del__(self):
        try:
            lib().orc_free(self.h)
        except Exception:
            pass

    def get(self) -> OB:
        L = lib()
        nl = L.orc_nlist(self.h)
        ns = L.orc_nsym(self.h)
        nblk = np.zeros(nl, np.int32)
        n = L.orc_counts(self.h, _ip(nblk))
        nb = int(nblk.sum())
        qin = np.zeros(max(nb, 1) * ns, np.int32)
        qout = np.zeros(max(nb, 1) * ns, np.int32)
        rows = np.zeros(max(nb, 1), np.int32)
        cols = np.zeros(max(nb, 1), np.int32)
        data = np.zeros(max(n, 1))
        L.orc_export(self.h, _ip(qin), _ip(qout), _ip(rows), _ip(cols), _dp(data))
        return OB(ns, nblk, qin[:nb * ns].reshape(nb, ns), qout[:nb * ns].reshape(nb, ns), rows[:nb], cols[:nb], data[:n], _group[1])


_alive = []
_group = [1, None]  # (nsym, periods) of the oracle's current group


def set_group(nsym: int, periods=None):
    """Make (nsym, periods) the oracle's current symmetry group (None: U(1) for every charge)."""
    per = np.zeros(nsym, np.int32) if periods is None else np.ascontiguousarray(periods, np.int32).reshape(-1)
    lib().orc_set_periods(nsym, _ip(per))
    _group[0], _group[1] = nsym, (None if periods is None else per.copy())


def model_periods(model):
    """The model's group as the oracle restates it (0 = U(1), N = Z(N))."""
    t, d = model & 0xFF, model_d(model)
    return np.array({1: [0, 0], 2: [2], 3: [d]}.get(t, [0]), np.int32)


def model_d(model):
    return lib().orc_model_d(model)


def model_nsym(model):
    return lib().orc_model_nsym(model)


def _qn(QN):
    q = np.asarray(list(QN), np.int32).reshape(-1)
    return q if q.size else np.zeros(1, np.int32)


def _h(x: OB):
    """Handle of x kept alive until the current oracle call returns (see _call)."""
    h = x.handle()
    _alive.append(h)
    return h


def _op(f):
    """Release the argument handles once an oracle operation has returned."""
    import functools

    @functools.wraps(f)
    def g(*a, **k):
        try:
            return f(*a, **k)
        finally:
            _alive.clear()
    return g


def _pp(hs):
    arr = (ctypes.c_void_p * len(hs))(*[h.h for h in hs])
    return arr


# ---- operations (each returns a new OB) ----
@_op
def tm_mixed(direction, A: OB, B, x: OB, y_layout: OB) -> OB:
    hA, hx, hy = _h(A), _h(x), _h(y_layout)
    hB = _h(B) if B is not None else None
    return _H(lib().orc_tm_mixed(direction, hA.h, hB.h if hB else None, hx.h, hy.h)).get()


@_op
def tm_ret(direction, A: OB, x: OB) -> OB:
    return _H(lib().orc_tm_ret(direction, _h(A).h, _h(x).h)).get()


@_op
def tm_chain(direction, As, x: OB, y_layout: OB) -> OB:
    hs = [_h(a) for a in As]
    return _H(lib().orc_tm_chain(direction, _pp(hs), len(hs), _h(x).h, _h(y_layout).h)).get()


@_op
def tm_unity(kind, A: OB, B=None) -> OB:
    hA = _h(A)
    hB = _h(B) if B is not None else None
    return _H(lib().orc_tm_unity(kind, hA.h, hB.h if hB else None)).get()


@_op
def mps_mul(l: OB, r: OB) -> OB:
    return _H(lib().orc_mps_mul(_h(l).h, _h(r).h)).get()


@_op
def mat_mps(m: OB, a: OB) -> OB:
    return _H(lib().orc_mat_mps(_h(m).h, _h(a).h)).get()


@_op
def mps_mat(a: OB, m: OB) -> OB:
    return _H(lib().orc_mps_mat(_h(a).h, _h(m).h)).get()


@_op
def bmul(l: OB, r: OB) -> OB:
    return _H(lib().orc_bmul(_h(l).h, _h(r).h)).get()


@_op
def btrans(a: OB) -> OB:
    return _H(lib().orc_btrans(_h(a).h)).get()


@_op
def apply_operator(AA: OB, H: np.ndarray) -> OB:
    Hf = np.asfortranarray(H, np.float64).reshape(-1, order="F").copy()
    return _H(lib().orc_apply_operator(_h(AA).h, _dp(Hf), H.shape[0])).get()


@_op
def reduced_op(H: np.ndarray, d: int, A: OB, direction: int) -> OB:
    Hf = np.asfortranarray(H, np.float64).reshape(-1, order="F").copy()
    return _H(lib().orc_reduced_op(_dp(Hf), d, _h(A).h, direction)).get()


@_op
def op_tm_diag(direction, H, d, AA: OB) -> OB:
    Hf = np.asfortranarray(H, np.float64).reshape(-1, order="F").copy()
    return _H(lib().orc_op_tm_diag(direction, _dp(Hf), d, _h(AA).h)).get()


@_op
def get_h(direction, As, H, d) -> OB:
    Hf = np.asfortranarray(H, np.float64).reshape(-1, order="F").copy()
    hs = [_h(a) for a in As]
    return _H(lib().orc_get_h(direction, _pp(hs), len(hs), _dp(Hf), d)).get()


@_op
def hac(EHL: OB, EHR: OB, HL: OB, HR: OB, x: OB) -> OB:
    return _H(lib().orc_hac(_h(EHL).h, _h(EHR).h, _h(HL).h, _h(HR).h, _h(x).h)).get()


@_op
def hc(EHL: OB, EHR: OB, H, d, AL: OB, AR: OB, x: OB) -> OB:
    Hf = np.asfortranarray(H, np.float64).reshape(-1, order="F").copy()
    return _H(lib().orc_hc(_h(EHL).h, _h(EHR).h, _dp(Hf), d, _h(AL).h, _h(AR).h, _h(x).h)).get()


@_op
def invert_e_proj(direction, As, x: OB, L: OB, R: OB, tol=1e-14, maxit=0, y0=None):
    hs = [_h(a) for a in As]
    flag = ctypes.c_int(0)
    hy0 = _h(y0) if y0 is not None else None
    out = _H(lib().orc_invert_e_proj(direction, _pp(hs), len(hs), _h(x).h, _h(L).h, _h(R).h, tol, maxit,
                                     hy0.h if hy0 else None, ctypes.byref(flag))).get()
    return out, flag.value


@_op
def mps_qr(A: OB, direction: int):
    R = ctypes.c_void_p()
    Q = _H(lib().orc_mps_qr(_h(A).h, direction, ctypes.byref(R))).get()
    return Q, _H(R.value).get()


@_op
def nullspace(A: OB, direction: int) -> OB:
    return _H(lib().orc_nullspace(_h(A).h, direction)).get()


@_op
def gradnorm(direction, A: OB, C: OB, EHL, EHR, HL, HR) -> float:
    f = lib().orc_gradnorm_right if direction == 1 else lib().orc_gradnorm_left
    return f(_h(A).h, _h(C).h, _h(EHL).h, _h(EHR).h, _h(HL).h, _h(HR).h)


# ---- bookkeeping ----
@_op
def i2k(model, N, QN):
    d, ns = model_d(model), model_nsym(model)
    out = np.zeros(d * ns, np.int32)
    q = _qn(QN)
    assert lib().orc_i2k(model, N, _ip(q), _ip(out)) == d
    return out.reshape(d, ns)


@_op
def validpath(model, N, QN):
    out = np.zeros(N, np.int32)
    q = _qn(QN)
    assert lib().orc_validpath(model, N, _ip(q), _ip(out)) == N
    return out


@_op
def make_dims_model(model, dims, N, QN):
    ns = model_nsym(model)
    nsec = len(dims) ** 2 if model == 1 else len(dims)
    keys = np.zeros(N * nsec * ns, np.int32)
    sizes = np.zeros(N * nsec, np.int32)
    dd = np.asarray(dims, np.int32)
    q = _qn(QN)
    n = lib().orc_make_dims_model(model, _ip(dd), len(dd), N, _ip(q), _ip(keys), _ip(sizes))
    assert n > 0, lib().orc_last_error()
    return [(keys.reshape(N, n, ns)[k], sizes.reshape(N, n)[k]) for k in range(N)]


@_op
def make_dims(model, N, QN, keys0, sizes0):
    ns = model_nsym(model)
    k0 = np.ascontiguousarray(keys0, np.int32).reshape(-1)
    s0 = np.ascontiguousarray(sizes0, np.int32)
    n0 = len(s0)
    keys = np.zeros(N * n0 * ns, np.int32)
    sizes = np.zeros(N * n0, np.int32)
    q = _qn(QN)
    n = lib().orc_make_dims(model, N, _ip(q), ns, n0, _ip(k0), _ip(s0), _ip(keys), _ip(sizes))
    assert n == n0, lib().orc_last_error()
    return [(keys.reshape(N, n, ns)[k], sizes.reshape(N, n)[k]) for k in range(N)]


@_op
def mps_structure(model, N, QN, left, right) -> OB:
    ns = model_nsym(model)
    kl, sl = (np.ascontiguousarray(a, np.int32) for a in left)
    kr, sr = (np.ascontiguousarray(a, np.int32) for a in right)
    q = _qn(QN)
    set_group(ns, model_periods(model))
    return _H(lib().orc_mps_structure(model, N, _ip(q), ns, len(sl), _ip(kl.reshape(-1)), _ip(sl), len(sr),
                                      _ip(kr.reshape(-1)), _ip(sr))).get()


@_op
def model_ham(model, params):
    d = model_d(model)
    p = np.asarray(params, np.float64)
    H = np.zeros(d ** 4)
    lib().orc_model_ham(model, _dp(p), _dp(H))
    return H.reshape(d * d, d * d, order="F")


class HacPlanCPU:
    """H_AC on flat Vectorize-order arrays, compiled once, OpenMP-threaded (orc_hac_plan): the CPU
    baseline of the H_AC workload (the products of apply_hac, VUMPS_helpers.hpp:38-88)."""

    def __init__(self, EHL, EHR, HL, HR, AC):
        hs = [_h(o) for o in (EHL, EHR, HL, HR, AC)]
        self.h = lib().orc_hac_plan(*[x.h for x in hs])
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())
        self.n = int(lib().orc_hac_plan_nx(self.h))

    def apply(self, x: np.ndarray, y: np.ndarray, threads: int = 0):
        lib().orc_hac_plan_apply(self.h, _dp(x), _dp(y), int(threads))

    def __del__(self):
        try:
            lib().orc_hac_plan_free(self.h)
        except Exception:
            pass


def max_threads() -> int:
    return int(lib().orc_max_threads())
