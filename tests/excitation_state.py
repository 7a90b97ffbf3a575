"""A small converged ground state for excitation tests (built with the oracle VUMPS on the CPU).

Anisotropic Heisenberg chain in the reference convention Delta = -1.5 (antiferromagnetic, gapped),
J = 1, hz = 0, two-site cell, QN = [0], bond 0: 7 sectors of sizes [1, 2, 3, 4, 3, 2, 1] (D = 16),
seed 7, iterated to gradient norm F < 1e-12.  Also returns the null spaces NL (oracle Nullspace),
L = C^T C and R = C C^T on bond 0 and the oracle's excitation constants."""
from __future__ import annotations

import functools

import numpy as np

from oracle import excitations as ex
from oracle import pyoracle as orc


@functools.lru_cache(maxsize=None)
def state(delta: float = -1.5, sizes: tuple = (1, 2, 3, 4, 3, 2, 1), seed: int = 7):
    """The default is the D = 16 state above; other (delta, sizes, seed) give further XXZ states
    built the same way (bond-0 charges 2 (i - len(sizes) // 2))."""
    from oracle_vumps import OracleVUMPS
    ns = len(sizes)
    keys0 = np.array([[2 * (i - ns // 2)] for i in range(ns)], np.int32)
    sizes0 = np.array(sizes, np.int32)
    v = OracleVUMPS(0, 2, [0], keys0, sizes0, [1.0, delta, 0.0], seed=seed)
    e, F = 0.0, [1.0]
    for _ in range(400):
        e, F = v.sweep()
        if max(F) < 1e-12:
            break
    return _finish(v, e, F)


@functools.lru_cache(maxsize=None)
def hubbard_state(nside: int = 4, peak: int = 2, seed: int = 5, U: float = 4.0):
    """A small Fermi-Hubbard U(1)xU(1) state (t = 1, U = 4, half filling QN = [2, 0], two-site cell),
    bond 0 from the (N, M) checkerboard profile (nside = 4, peak 2: D = 16 over 16 sectors), iterated
    by the oracle VUMPS to F < 1e-11 at that fixed structure (a stationary point)."""
    from oracle_vumps import OracleVUMPS
    from imps_cpp_tools_amd.workloads import fhub_profile
    keys0, sizes0 = fhub_profile(nside, peak, 1)
    v = OracleVUMPS(1, 2, [2, 0], keys0, sizes0, [1.0, U, 0.0, 0.0, 0.0], seed=seed)
    e, F = 0.0, [1.0]
    for _ in range(400):
        e, F = v.sweep()
        if max(F) < 1e-11:
            break
    return _finish(v, e, F)


def _finish(v, e, F):
    NLob = [orc.nullspace(a, 0) for a in v.AL]
    H = np.array(v.H)
    AL = [ex.from_ob(a) for a in v.AL]
    AR = [ex.from_ob(a) for a in v.AR]
    NL = [ex.from_ob(a) for a in NLob]
    C = [ex.from_ob(c)[0] for c in v.C]
    H0 = H - e * np.eye(H.shape[0])
    L = ex.bm_mul(ex.bm_t(C[-1]), C[-1])
    R = ex.bm_mul(C[-1], ex.bm_t(C[-1]))
    HLtot, HRtot = ex.heff_constants(AL, AR, L, R, H0)
    return dict(v=v, E0=e, F=max(F), H=H, H0=H0, ALob=v.AL, ARob=v.AR, Cob=v.C, NLob=NLob, AL=AL, AR=AR, NL=NL, C=C, L=L, R=R,
                HLtot=HLtot, HRtot=HRtot, N=2, d=v.d)


def oracle_heff(K, p, z, s=None, proj_at_unit_phase=False):
    """Oracle ApplyHeff on the flat complex vector z (XDims order) at momentum p (units of pi per site).
    Projected sums at K = 0 and p = 0 (the reference's rule, Excitations_main.cpp:264-268), or with
    proj_at_unit_phase wherever e^{i p pi N} = 1 (the C-ABI engine's rule)."""
    s = state() if s is None else s
    xd = ex.xdims(s["NL"], s["AR"], K)
    unit = abs(np.exp(1j * p * np.pi * s["N"]) - 1) < 1e-13 if proj_at_unit_phase else abs(p) < 1e-13
    proj = unit and all(k == 0 for k in K)
    LM, RM = (None, None)
    if proj:
        LM, RM, _ = ex.lr_overlaps(s["AL"], s["AR"], K)
    kfac = np.exp(1j * p * np.pi * s["N"])
    X = ex.from_vec_list(z, xd)
    return ex.vectorize_list(ex.apply_heff(X, kfac, s["AL"], s["AR"], s["NL"], LM, RM, s["H0"], s["HLtot"], s["HRtot"]))


def oracle_heff_matrix(K, p, s=None, proj_at_unit_phase=False):
    s = state() if s is None else s
    xd = ex.xdims(s["NL"], s["AR"], K)
    n = sum(r * c for dv in xd for *_, r, c in dv)
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        M[:, j] = oracle_heff(K, p, e, s, proj_at_unit_phase)
    return M


MODELS = {"xxz": (0, "XXZ", [0]), "hub": (1, "FHUB", [2, 0])}


@functools.lru_cache(maxsize=None)
def transformed_state(which: str, op: str, K: tuple, shift: int = 0):
    """The oracle's right half of the excitation program (Excitations_main.cpp:142-177) for the state
    of `which` ("xxz": state(), "hub": hubbard_state()): ApplyTransformation(AR, L, op) when op is
    given, the relative shift of AR and L by `shift` sites (site n takes site n - shift), LRoverlaps
    in K and AR[0] times the sign of the overlap; HLtot / HRtot of HeffConstants with the new L."""
    s = state() if which == "xxz" else hubbard_state()
    model, mname, QN = MODELS[which]
    N = s["N"]
    b = (N - 1 - shift) % N
    L = ex.bm_mul(ex.bm_t(s["C"][b]), s["C"][b])
    AR = list(s["AR"])
    if op:
        I2K = [tuple(int(v) for v in k) for k in np.asarray(orc.i2k(model, N, QN)).reshape(s["d"], -1)]
        wh, signs = ex.transformation_masks(mname, op, N)
        AR, L = ex.apply_transformation(AR, L, I2K, wh, signs)
    AR = [AR[(n - shift) % N] for n in range(N)]
    _, _, ol = ex.lr_overlaps(s["AL"], AR, K)
    if ol < 0:
        AR[0] = [ex.bm_scale(bm, -1.0) for bm in AR[0]]
    HLtot, HRtot = ex.heff_constants(s["AL"], AR, L, s["R"], s["H0"])
    out = dict(s)
    out.update(AR=AR, L=L, HLtot=HLtot, HRtot=HRtot, OL=ol, model=model, QN=QN)
    return out
