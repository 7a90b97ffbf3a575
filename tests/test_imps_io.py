"""IMPS state files (saveIMPS/loadIMPS, MPSBlockUtilities.hpp:1212-1327).

The byte-level fixture below is assembled by hand from the reference's writers (tags as text
lines, native u32 counts, keys as u32 length + i32 charges, matrices in Armadillo's arma_binary
layout), independently of imps_io's writer, for a one-site cell with one U(1) sector: the writer
must produce exactly these bytes and the reader must return the objects.  Byte-exact agreement
with files written by the reference itself is not pinned (no such file exists in the reference).
"""
import os
import struct
import tempfile

import numpy as np
import pytest

from states import FHUB, XXZ, Cell, fhub_bond0, random_like, xxz_bond0
from imps_cpp_tools_amd import BlockLayout
from imps_cpp_tools_amd import imps_io


def _lay(ob):
    return BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def _u32(v):
    return struct.pack("<I", v)


def _key(*q):
    return _u32(len(q)) + struct.pack("<%di" % len(q), *q)


def _mat(m):
    m = np.asarray(m, np.float64)
    return b"ARMA_MAT_BIN_FN008\n" + ("%d %d\n" % m.shape).encode() + m.tobytes(order="F")


def _fixture():
    lam = np.array([0.8, 0.6])
    C = np.array([[0.8, 0.0], [0.1, 0.6]])
    A0 = np.array([[1.0, 2.0], [3.0, 4.0]])   # s = 0: charge (0) -> (1), 2 x 2
    A1 = np.array([[5.0, 6.0], [7.0, 8.0]])   # s = 1: charge (2) -> (1)
    B0 = np.array([[0.5], [0.25]])            # AR, s = 0: (1) -> (2), 2 x 1
    data = (b"BLOCKLAMARRAY\n" + _u32(1) + b"BLOCKLAM\n" + _u32(1) + _key(1) + _mat(lam.reshape(2, 1))
            + b"RBLOCKDIAGMATARRAY\n" + _u32(1) + b"RBLOCKDIAGMAT\n" + _u32(1) + _key(1) + _mat(C)
            + b"RBLOCKMPSARRAY\n" + _u32(1) + b"RBLOCKMPS\n" + _u32(2) + _u32(1)
            + b"RBLOCKMAT\n" + _u32(1) + _key(0) + _key(1) + _mat(A0)
            + b"RBLOCKMAT\n" + _u32(1) + _key(2) + _key(1) + _mat(A1)
            + b"RBLOCKMPSARRAY\n" + _u32(1) + b"RBLOCKMPS\n" + _u32(2) + _u32(1)
            + b"RBLOCKMAT\n" + _u32(1) + _key(1) + _key(2) + _mat(B0)
            + b"RBLOCKMAT\n" + _u32(0))
    objs = dict(
        lam=[{(1,): lam}],
        C=[imps_io.diag_from_items(1, [((1,), C)])],
        AL=[imps_io.mps_from_items(1, [[((0,), (1,), A0)], [((2,), (1,), A1)]])],
        AR=[imps_io.mps_from_items(1, [[((1,), (2,), B0)], []])],
    )
    return data, objs


def _same(a, b):
    (la, da), (lb, db) = a, b
    assert la == lb
    np.testing.assert_array_equal(da, db)


def test_writer_matches_hand_assembled_bytes():
    data, o = _fixture()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cell.bin")
        imps_io.write_imps(path, o["lam"], o["C"], o["AL"], o["AR"])
        with open(path, "rb") as f:
            assert f.read() == data


def test_reader_parses_hand_assembled_bytes():
    data, o = _fixture()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cell.bin")
        with open(path, "wb") as f:
            f.write(data)
        lam, C, AL, AR = imps_io.read_imps(path, 1)
    assert list(lam[0]) == [(1,)]
    np.testing.assert_array_equal(lam[0][(1,)], o["lam"][0][(1,)])
    _same(C[0], o["C"][0])
    _same(AL[0], o["AL"][0])
    _same(AR[0], o["AR"][0])


@pytest.mark.parametrize("model", [XXZ, FHUB], ids=["xxz", "fhub"])
def test_round_trip_of_a_cell(model):
    if model == XXZ:
        (k, s), QN = xxz_bond0(7, 6), [0]
    else:
        (k, s), QN = fhub_bond0(3, 3), [2, 0]
    c = Cell(model, 2, QN, k, s, seed=21)
    rng = np.random.default_rng(22)
    Cs = [random_like(c.diag(n + 1), rng) for n in range(2)]
    lam = [{tuple(int(v) for v in q): np.sort(rng.random(int(r)))[::-1] for q, r in zip(C.qin, C.rows)} for C in Cs]
    args = (lam, [(_lay(C), C.data) for C in Cs], [(_lay(A), A.data) for A in c.AL], [(_lay(A), A.data) for A in c.AR])
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cell.bin")
        imps_io.write_imps(path, *args)
        got = imps_io.read_imps(path, c.nsym)
    for site_got, site_ref in zip(got[0], lam):
        assert list(site_got) == sorted(site_ref)
        for q in site_ref:
            np.testing.assert_array_equal(site_got[q], site_ref[q])
    for part in (1, 2, 3):
        for a, b in zip(got[part], args[part]):
            _same(a, b)


def test_reader_rejects_bad_files():
    data, _ = _fixture()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cell.bin")
        for bad in (data[:-5], b"BLOCKLAMS\n" + data[14:], data.replace(b"ARMA_MAT_BIN_FN008", b"ARMA_MAT_BIN_FC016", 1)):
            with open(path, "wb") as f:
                f.write(bad)
            with pytest.raises(ValueError):
                imps_io.read_imps(path, 1)


def _cmat(m):
    m = np.asarray(m, np.complex128)
    # arma_binary complex: header line, dims line, column-major (re, im) pairs
    pairs = np.stack([m.real.T.reshape(-1), m.imag.T.reshape(-1)], axis=1)
    return b"ARMA_MAT_BIN_FC016\n" + ("%d %d\n" % m.shape).encode() + pairs.astype("<f8").tobytes()


def test_excitation_vectors_bytes():
    """Two vectors over a two-site XDims (site 0: one 1 x 2 block (0) -> (2); site 1: none), as the
    generic vector save of BlockObj.hpp:3196-3222 writes std::vector<BlockMatArray<IKey, Complex>>."""
    xd = BlockLayout(1, np.array([1, 0], np.int32), np.array([[0]], np.int32), np.array([[2]], np.int32), np.array([1], np.int32),
                     np.array([2], np.int32))
    X = np.array([[1 + 2j, 3 - 4j], [0.5j, -1.0]])
    want = b"CBLOCKMATARRAYVEC\n" + _u32(2)
    for x in X:
        want += (b"CBLOCKMATARRAY\n" + _u32(2) + b"CBLOCKMAT\n" + _u32(1) + _key(0) + _key(2) + _cmat(x.reshape(1, 2))
                 + b"CBLOCKMAT\n" + _u32(0))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "x.bin")
        imps_io.write_excitations(p, X, xd)
        with open(p, "rb") as f:
            assert f.read() == want
        with pytest.raises(ValueError):
            imps_io.write_excitations(p, X[:, :1], xd)


def test_alrn_sections_in_order():
    """saveALRN: C, L, R as BlockDiagMatArrays, then AL, AR, NL as MPSBlockMatArrays."""
    data, objs = _fixture()
    C, AL, AR = objs["C"], objs["AL"], objs["AR"]
    L = [imps_io.diag_from_items(1, [((1,), np.array([[2.0, 0.0], [0.0, 3.0]]))])]
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "alrn.bin")
        imps_io.write_alrn(p, C, L, C, AL, AR, AL)
        with open(p, "rb") as f:
            got = f.read()
    diag = lambda m: b"RBLOCKDIAGMATARRAY\n" + _u32(1) + b"RBLOCKDIAGMAT\n" + _u32(1) + _key(1) + _mat(m)  # noqa: E731
    i = data.index(b"RBLOCKMPSARRAY")
    mps_al = data[i:data.index(b"RBLOCKMPSARRAY", i + 1)]
    mps_ar = data[data.index(b"RBLOCKMPSARRAY", i + 1):]
    Cm = np.array([[0.8, 0.0], [0.1, 0.6]])
    assert got == diag(Cm) + diag(np.diag([2.0, 3.0])) + diag(Cm) + mps_al + mps_ar + mps_al
