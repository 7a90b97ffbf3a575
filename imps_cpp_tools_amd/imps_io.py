"""IMPS state files: the reference's saveIMPS / loadIMPS container format
(MPSBlockUtilities.hpp:1212-1327), for states exchanged with the reference's VUMPS and
Excitations programs.

A file holds, in order, the Schmidt values, the bond matrices, AL and AR of an N-site unit cell:

  BlockLamArray      "BLOCKLAMARRAY\\n" u32 N, N x BlockLam            (BlockObj.hpp:2692-2708)
  BlockLam           "BLOCKLAM\\n" u32 nsec, nsec x (key, column)       (BlockObj.hpp:150-184)
  BlockDiagMatArray  "RBLOCKDIAGMATARRAY\\n" u32 N, N x BlockDiagMat   (BlockObj.hpp:2877-2893)
  BlockDiagMat       "RBLOCKDIAGMAT\\n" u32 nsec, nsec x (key, matrix)  (BlockObj.hpp:1968-2000)
  MPSBlockMatArray   "RBLOCKMPSARRAY\\n" u32 N, N x MPSBlockMat        (MPSBlockMat.hpp:1326-1342)
  MPSBlockMat        "RBLOCKMPS\\n" u32 d, u32 nsites, d x BlockMat    (MPSBlockMat.hpp:1085-1103)
  BlockMat           "RBLOCKMAT\\n" u32 nblk, nblk x (key in, key out, matrix)  (BlockObj.hpp:905-945)
  key                u32 nsym, nsym x i32                              (KeyTypes.hpp:99-118)

Tags are text lines (the reader takes a whitespace-delimited word and skips one character);
counts are native 32-bit unsigned integers (little-endian here); sectors follow the map order
(ascending charge).  Matrices use Armadillo's arma_binary layout, a third-party format not held
in the reference (Armadillo's version is not pinned there): the line "ARMA_MAT_BIN_FN008", the line
"<rows> <cols>", then the column-major doubles; a Schmidt column is a rows x 1 matrix.  The
symmetry group is not stored (the reader's model supplies it, KeyTypes.hpp:101-103).  Real states
only (the "R" tags; VUMPS saves real states).

The excitation program's output files use the same pieces: saveALRN (Excitations_helpers.hpp:948-989)
is C, L, R as BlockDiagMatArrays and AL, AR, NL as MPSBlockMatArrays, one after the other; its
excitation vectors (Excitations_main.cpp:285-292, the generic save of BlockObj.hpp:3196-3222) are
"CBLOCKMATARRAYVEC\n" u32 nev, nev x ("CBLOCKMATARRAY\n" u32 N, N x ("CBLOCKMAT\n" u32 nblk,
nblk x (key in, key out, complex matrix))) with complex matrices "ARMA_MAT_BIN_FC016", "<rows> <cols>",
then the column-major (real, imaginary) pairs.

Cell convention (as randMPS_LR builds it and VUMPS_main.cpp:367-372 restores it before saving):
AL[n] and AR[n] map bond n to bond n + 1, C[n] lives on bond n + 1 (AL[n] C[n] = C[n-1] AR[n]) and
lam[n] are C[n]'s singular values per sector.
"""
from __future__ import annotations

import struct
from typing import BinaryIO, Dict, List, Tuple

import numpy as np

from . import BlockLayout

Key = Tuple[int, ...]
ARMA_REAL = b"ARMA_MAT_BIN_FN008"
ARMA_CPLX = b"ARMA_MAT_BIN_FC016"


# ---------------------------------------------------------------------------------------------
# primitives
def _w_tag(f: BinaryIO, tag: str):
    f.write(tag.encode("ascii") + b"\n")


def _w_u32(f: BinaryIO, v: int):
    f.write(struct.pack("<I", int(v)))


def _w_key(f: BinaryIO, key: Key):
    _w_u32(f, len(key))
    f.write(struct.pack("<%di" % len(key), *[int(q) for q in key]))


def _w_mat(f: BinaryIO, m: np.ndarray):
    m = np.asarray(m, np.float64)
    if m.ndim == 1:
        m = m.reshape(-1, 1)
    f.write(ARMA_REAL + b"\n" + ("%d %d\n" % m.shape).encode("ascii"))
    f.write(np.asfortranarray(m).astype("<f8").tobytes(order="F"))


def _w_cmat(f: BinaryIO, m: np.ndarray):
    m = np.asarray(m, np.complex128)
    f.write(ARMA_CPLX + b"\n" + ("%d %d\n" % m.shape).encode("ascii"))
    f.write(np.asfortranarray(m).astype("<c16").tobytes(order="F"))


def _r_word(f: BinaryIO) -> bytes:
    """operator>> on a string: skip whitespace, read to the next whitespace (not consumed)."""
    c = f.read(1)
    while c and c.isspace():
        c = f.read(1)
    out = b""
    while c and not c.isspace():
        out += c
        c = f.read(1)
    if c:
        f.seek(-1, 1)
    return out


def _r_tag(f: BinaryIO, want: str):
    got = _r_word(f)
    f.read(1)  # the newline after the tag (file.get() in the reference's readers)
    if got != want.encode("ascii"):
        raise ValueError(f"IMPS file: header {got!r}, expected {want!r}")


def _r_u32(f: BinaryIO) -> int:
    b = f.read(4)
    if len(b) != 4:
        raise ValueError("IMPS file: truncated")
    return struct.unpack("<I", b)[0]


def _r_key(f: BinaryIO) -> Key:
    n = _r_u32(f)
    b = f.read(4 * n)
    if len(b) != 4 * n:
        raise ValueError("IMPS file: truncated key")
    return tuple(struct.unpack("<%di" % n, b))


def _r_mat(f: BinaryIO) -> np.ndarray:
    head = _r_word(f)
    if head != ARMA_REAL:
        raise ValueError(f"IMPS file: matrix header {head!r} (only real FP64 matrices are supported)")
    f.read(1)
    rows, cols = int(_r_word(f)), int(_r_word(f))
    f.read(1)
    b = f.read(8 * rows * cols)
    if len(b) != 8 * rows * cols:
        raise ValueError("IMPS file: truncated matrix")
    return np.frombuffer(b, "<f8").astype(np.float64).reshape(rows, cols, order="F")


# ---------------------------------------------------------------------------------------------
# block containers on flat buffers (this package's layouts)
def _diag_items(lay: BlockLayout, data) -> List[Tuple[Key, np.ndarray]]:
    out = []
    for q, o, r, c in zip(lay.qin, lay.offsets, lay.rows, lay.cols):
        out.append((tuple(int(v) for v in q), np.asarray(data[o:o + int(r) * int(c)]).reshape(int(r), int(c), order="F")))
    return out


def _mps_items(lay: BlockLayout, data) -> List[List[Tuple[Key, Key, np.ndarray]]]:
    out, b, offs = [], 0, lay.offsets
    for e in range(lay.nlist):
        ent = []
        for j in range(b, b + int(lay.nblk[e])):
            r, c = int(lay.rows[j]), int(lay.cols[j])
            ent.append((tuple(int(v) for v in lay.qin[j]), tuple(int(v) for v in lay.qout[j]),
                        np.asarray(data[offs[j]:offs[j] + r * c]).reshape(r, c, order="F")))
        out.append(ent)
        b += int(lay.nblk[e])
    return out


def diag_from_items(nsym: int, items: List[Tuple[Key, np.ndarray]], periods=None) -> Tuple[BlockLayout, np.ndarray]:
    items = sorted(items, key=lambda t: t[0])
    lay = BlockLayout(nsym, [len(items)], np.asarray([k for k, _ in items], np.int32).reshape(-1, nsym),
                      np.asarray([k for k, _ in items], np.int32).reshape(-1, nsym), [m.shape[0] for _, m in items],
                      [m.shape[1] for _, m in items], periods)
    data = np.concatenate([m.reshape(-1, order="F") for _, m in items]) if items else np.zeros(0)
    return lay, data


def mps_from_items(nsym: int, items: List[List[Tuple[Key, Key, np.ndarray]]], periods=None) -> Tuple[BlockLayout, np.ndarray]:
    nblk, qin, qout, rows, cols, parts = [], [], [], [], [], []
    for ent in items:
        ent = sorted(ent, key=lambda t: t[0])
        nblk.append(len(ent))
        for ki, ko, m in ent:
            qin.append(ki), qout.append(ko), rows.append(m.shape[0]), cols.append(m.shape[1])
            parts.append(m.reshape(-1, order="F"))
    nb = sum(nblk)
    lay = BlockLayout(nsym, nblk, np.asarray(qin, np.int32).reshape(nb, nsym), np.asarray(qout, np.int32).reshape(nb, nsym), rows, cols, periods)
    return lay, (np.concatenate(parts) if parts else np.zeros(0))


# ---------------------------------------------------------------------------------------------
def _w_diag_array(f: BinaryIO, arr: List[Tuple[BlockLayout, np.ndarray]]):
    _w_tag(f, "RBLOCKDIAGMATARRAY")
    _w_u32(f, len(arr))
    for lay, data in arr:
        items = _diag_items(lay, data)
        _w_tag(f, "RBLOCKDIAGMAT")
        _w_u32(f, len(items))
        for k, m in items:
            _w_key(f, k)
            _w_mat(f, m)


def _w_mps_array(f: BinaryIO, arr: List[Tuple[BlockLayout, np.ndarray]]):
    _w_tag(f, "RBLOCKMPSARRAY")
    _w_u32(f, len(arr))
    for lay, data in arr:
        _w_tag(f, "RBLOCKMPS")
        _w_u32(f, lay.nlist)
        _w_u32(f, 1)  # one-site tensors
        for ent in _mps_items(lay, data):
            _w_tag(f, "RBLOCKMAT")
            _w_u32(f, len(ent))
            for ki, ko, m in ent:
                _w_key(f, ki)
                _w_key(f, ko)
                _w_mat(f, m)


def write_imps(path: str, lam: List[Dict[Key, np.ndarray]], C: List[Tuple[BlockLayout, np.ndarray]],
               AL: List[Tuple[BlockLayout, np.ndarray]], AR: List[Tuple[BlockLayout, np.ndarray]]):
    """saveIMPS(Lamvec, Cvec, ALvec, ARvec, file): lam[n] = {charge: Schmidt values}, C[n] a
    block-diagonal (layout, data), AL[n] / AR[n] one-site MPS tensors (layout with d entries, data)."""
    with open(path, "wb") as f:
        _w_tag(f, "BLOCKLAMARRAY")
        _w_u32(f, len(lam))
        for site in lam:
            _w_tag(f, "BLOCKLAM")
            _w_u32(f, len(site))
            for k in sorted(site):
                _w_key(f, k)
                _w_mat(f, np.asarray(site[k], np.float64).reshape(-1, 1))
        _w_diag_array(f, C)
        _w_mps_array(f, AL)
        _w_mps_array(f, AR)


def write_alrn(path: str, C, L, R, AL, AR, NL):
    """saveALRN (Excitations_helpers.hpp:948-989): C, L, R (per bond, (layout, data)) and AL, AR, NL
    (per site), as the excitation program saves the state it excites (with its shifts and
    transformations)."""
    with open(path, "wb") as f:
        for arr in (C, L, R):
            _w_diag_array(f, arr)
        for arr in (AL, AR, NL):
            _w_mps_array(f, arr)


def write_excitations(path: str, X: np.ndarray, xdims: BlockLayout):
    """The excitation vectors of one momentum (Excitations_main.cpp:285-292): X (nev, mtot) complex in
    XDims order; xdims the layout with one entry per site (X[n]'s blocks)."""
    X = np.atleast_2d(np.asarray(X, np.complex128))
    if X.shape[1] != xdims.nelem:
        raise ValueError("excitation vectors do not match XDims")
    with open(path, "wb") as f:
        _w_tag(f, "CBLOCKMATARRAYVEC")
        _w_u32(f, X.shape[0])
        for x in X:
            _w_tag(f, "CBLOCKMATARRAY")
            _w_u32(f, xdims.nlist)
            for ent in _mps_items(xdims, x):
                _w_tag(f, "CBLOCKMAT")
                _w_u32(f, len(ent))
                for ki, ko, m in ent:
                    _w_key(f, ki)
                    _w_key(f, ko)
                    _w_cmat(f, m)


def read_imps(path: str, nsym: int, periods=None):
    """loadIMPS: returns (lam, C, AL, AR) in write_imps' form (layouts rebuilt from the stored
    sectors, nsym charges per key; `periods`: the reader's symmetry group, which the file does not
    hold)."""
    with open(path, "rb") as f:
        _r_tag(f, "BLOCKLAMARRAY")
        lam = []
        for _ in range(_r_u32(f)):
            _r_tag(f, "BLOCKLAM")
            site = {}
            for _ in range(_r_u32(f)):
                k = _r_key(f)
                site[k] = _r_mat(f)[:, 0].copy()
            lam.append(site)
        _r_tag(f, "RBLOCKDIAGMATARRAY")
        C = []
        for _ in range(_r_u32(f)):
            _r_tag(f, "RBLOCKDIAGMAT")
            items = []
            for _ in range(_r_u32(f)):
                k = _r_key(f)
                items.append((k, _r_mat(f)))
            C.append(diag_from_items(nsym, items, periods))
        arrs = []
        for _ in range(2):
            _r_tag(f, "RBLOCKMPSARRAY")
            arr = []
            for _ in range(_r_u32(f)):
                _r_tag(f, "RBLOCKMPS")
                d, nsites = _r_u32(f), _r_u32(f)
                if nsites != 1:
                    raise ValueError("IMPS file: only one-site tensors are supported")
                items = []
                for _ in range(d):
                    _r_tag(f, "RBLOCKMAT")
                    ent = []
                    for _ in range(_r_u32(f)):
                        ki = _r_key(f)
                        ko = _r_key(f)
                        ent.append((ki, ko, _r_mat(f)))
                    items.append(ent)
                arr.append(mps_from_items(nsym, items, periods))
            arrs.append(arr)
    if any(len(k) != nsym for site in lam for k in site):
        raise ValueError("IMPS file: charge keys do not have nsym entries")
    return lam, C, arrs[0], arrs[1]
