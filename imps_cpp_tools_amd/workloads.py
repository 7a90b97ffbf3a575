"""Seeded synthetic workloads at the measured shapes.

Sector profiles:
  * XXZ, one U(1): charges 2(i - n/2), Gaussian-in-charge sizes (transfer-matrix microbenchmark:
    17 sectors, sizes 2..64, D = 512).
  * Fermi-Hubbard U(1)xU(1) in (N, M) representation: keys (i+j-n+1, i-j) for i, j < n (the
    checkerboard of GetDimMap, Models.cpp:680-697), sizes Gaussian in the charge radius, clipped
    to [1, 60] and adjusted to sum exactly to D (D = 1024 over 121 sectors for the H_AC workload).
Operator data for timing runs are seeded normal entries at isometry scale; the reduced operators
HL/HR are computed on the GPU by the library's ReducedOp.  Everything is deterministic in the seed.
"""
from __future__ import annotations

import numpy as np

from . import FHUB, XXZ, BlockLayout, make_dims, mps_structure


def _fix_total(sizes: np.ndarray, total: int, lo: int, hi: int) -> np.ndarray:
    sizes = sizes.copy()
    order = np.argsort(-sizes, kind="stable")
    k = 0
    while sizes.sum() != total:
        i = order[k % len(order)]
        step = 1 if sizes.sum() < total else -1
        if lo <= sizes[i] + step <= hi:
            sizes[i] += step
        k += 1
        if k > 100 * len(sizes):
            raise ValueError("cannot reach the requested total")
    return sizes


def _gauss_sizes(r2: np.ndarray, peak: int, lo: int, total: int | None) -> np.ndarray:
    """max(lo, round(peak * exp(-r2 / (2 s2)))) with s2 chosen by bisection so that the sizes sum
    to `total` as closely as possible; the remainder is then spread one by one."""
    def sizes(s2):
        return np.maximum(lo, np.round(peak * np.exp(-r2 / (2 * s2)))).astype(np.int64)
    if total is None:
        return sizes(max(r2.max() / 8.0, 1.0))
    a, b = 1e-3, 1e6
    for _ in range(200):
        m = np.sqrt(a * b)
        if sizes(m).sum() > total:
            b = m
        else:
            a = m
    return _fix_total(sizes(a), total, lo, peak)


def xxz_profile(nsec: int, peak: int, lo: int = 1, total: int | None = None):
    """Bond-0 (keys, sizes) of an XXZ chain: charges 2(i - nsec//2), Gaussian sizes."""
    x = np.arange(nsec) - (nsec - 1) / 2
    s = _gauss_sizes(x.astype(float) ** 2, peak, lo, total)
    keys = np.array([[2 * (i - nsec // 2)] for i in range(nsec)], np.int32)
    return keys, s.astype(np.int32)


def fhub_profile(nside: int, peak: int, lo: int = 1, total: int | None = None):
    """Bond-0 (keys, sizes) of a Hubbard chain, keys in ascending (map) order."""
    keys, r2 = [], []
    for i in range(nside):
        for j in range(nside):
            n, m = i + j - nside + 1, i - j
            keys.append((n, m))
            r2.append(n * n + m * m)
    sizes = _gauss_sizes(np.array(r2, float), peak, lo, total)
    order = sorted(range(len(keys)), key=lambda k: keys[k])
    return np.array([keys[k] for k in order], np.int32), sizes[order].astype(np.int32)


HUBBARD_1024 = dict(model=FHUB, N=2, QN=[2, 0], nside=11, peak=60, total=1024, params=[1.0, 4.0, 0.0, 0.0, 0.0])
TM_512 = dict(model=XXZ, N=2, QN=[0], nsec=17, peak=64, lo=2, total=512)
# quasiparticle excitations on the Hubbard chain, D = 512 over 81 (N, M) sectors
HUBBARD_512 = dict(model=FHUB, N=2, QN=[2, 0], nside=9, peak=32, total=512, params=[1.0, 4.0, 0.0, 0.0, 0.0])


def hubbard_bonds(cfg=HUBBARD_1024):
    k0, s0 = fhub_profile(cfg["nside"], cfg["peak"], 1, cfg["total"])
    return make_dims(cfg["model"], cfg["N"], cfg["QN"], k0, s0)


def diag_layout(nsym: int, bond) -> BlockLayout:
    k, s = bond
    return BlockLayout.diag(nsym, k, s)


def hac_workload_layouts(cfg=HUBBARD_1024):
    """Layouts of the H_AC problem on site 0 of a two-site cell: AC: bond0 -> bond1, left
    neighbour AL[1]: bond1 -> bond0, right neighbour AR[1]: bond1 -> bond0."""
    bonds = hubbard_bonds(cfg)
    ns = 2
    AC = mps_structure(cfg["model"], cfg["N"], cfg["QN"], bonds[0], bonds[1])
    A10 = mps_structure(cfg["model"], cfg["N"], cfg["QN"], bonds[1], bonds[0])
    return dict(bonds=bonds, AC=AC, AL=A10, AR=A10, EHL=diag_layout(ns, bonds[0]), EHR=diag_layout(ns, bonds[1]))


def random_data(layout: BlockLayout, rng, scale_by_rows: bool = False, symmetric: bool = False) -> np.ndarray:
    out = rng.standard_normal(layout.nelem)
    offs = layout.offsets
    for b, (r, c) in enumerate(zip(layout.rows, layout.cols)):
        blk = out[offs[b]:offs[b] + r * c].reshape(c, r)
        if symmetric:
            blk[...] = 0.5 * (blk + blk.T)
        if scale_by_rows:
            blk /= np.sqrt(max(r, 1) * 4)
    return out
