"""Bond growth (ExpandFromH, MPSBlockUtilities.hpp:407-608) on the host, against a dense
computation that ignores the symmetry: the singular values of (1 - P_L) Phi (1 - P_R), with Phi
the two-site H (AL C AR) regrouped as (s_l, left bond) x (s_r, right bond) and P_L, P_R the
projectors on the ranges of the dense AL and AR, are exactly the sector singular values the
expansion ranks (existing charges through the null spaces, new charges through Phi itself).  The
kept states must be the largest of them, leave the state unchanged and keep AL / AR isometric."""
import numpy as np
import pytest

from states import FHUB, XXZ, Cell, dense_mps, fhub_bond0, random_like, xxz_bond0
from oracle import pyoracle as orc
from imps_cpp_tools_amd import BlockLayout
from imps_cpp_tools_amd import expand as ex

CASES = {
    "xxz": (XXZ, xxz_bond0(5, 4), [0], [1.0, -1.0, 0.3]),
    "fhub": (FHUB, fhub_bond0(3, 3), [2, 0], [1.0, 4.0, 0.0, 0.0, 0.0]),
}


def _lay(ob):
    return BlockLayout(ob.nsym, ob.nblk, ob.qin, ob.qout, ob.rows, ob.cols)


def _offsets(keys, sizes):
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    return {tuple(int(v) for v in k): (int(o), int(s)) for k, o, s in zip(np.asarray(keys).reshape(len(sizes), -1), off, sizes)}, int(np.sum(sizes))


def _dense_diag(keys, sizes, blocks):
    off, D = _offsets(keys, sizes)
    out = np.zeros((D, D))
    for q, m in blocks.items():
        a, _ = off[q]
        out[a:a + m.shape[0], a:a + m.shape[1]] = m
    return out


def _setup(name, seed):
    model, (k, s), QN, Hp = CASES[name]
    c = Cell(model, 2, QN, k, s, seed=seed)
    rng = np.random.default_rng(seed + 100)
    C = random_like(c.diag(1), rng)  # bond 1 between AL[0] and AR[1]
    H = orc.model_ham(model, Hp)
    dq = orc.i2k(model, 2, QN)
    return model, QN, c, C, H, dq


def _dense_state(c, bond_l, bond_c, bond_r, AL, AR, Cd):
    """Dense (d, Dl, Dc) AL, (d, Dc, Dr) AR -> two-site psi[s_l, s_r] = AL C AR."""
    offl, Dl = _offsets(*bond_l)
    offc, Dc = _offsets(*bond_c)
    offr, Dr = _offsets(*bond_r)
    ALd = dense_mps(AL, offl, Dl, offc, Dc)
    ARd = dense_mps(AR, offc, Dc, offr, Dr)
    d = ALd.shape[0]
    psi = np.einsum("aij,jk,bkl->abil", ALd, Cd, ARd)
    return ALd, ARd, psi.reshape(d * d, Dl, Dr), (Dl, Dc, Dr)


def _projected_phi(H, ALd, ARd, psi):
    d, Dl, Dc = ALd.shape
    Dr = ARd.shape[2]
    phi = np.einsum("st,tij->sij", H, psi).reshape(d, d, Dl, Dr)
    Phi = phi.transpose(0, 2, 1, 3).reshape(d * Dl, d * Dr)  # rows (s_l, i), columns (s_r, j)
    L = ALd.reshape(d * Dl, Dc)
    R = ARd.transpose(1, 0, 2).reshape(Dc, d * Dr)
    PL = np.eye(d * Dl) - L @ L.T
    PR = np.eye(d * Dr) - R.T @ R
    return PL @ Phi @ PR, Phi


def _as_ob(lay: BlockLayout, data):
    from oracle.pyoracle import OB
    return OB(lay.nsym, lay.nblk, lay.qin, lay.qout, lay.rows, lay.cols, np.asarray(data, np.float64))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("dm_max", [4, 1000])
def test_expand_from_h_matches_dense(name, dm_max):
    model, QN, c, C, H, dq = _setup(name, seed=7)
    AL, AR = c.AL[0], c.AR[1]
    ALb = ex.mps_blocks(_lay(AL), AL.data)
    ARb = ex.mps_blocks(_lay(AR), AR.data)
    Cb = ex.diag_blocks(_lay(C), C.data)
    dm, new_cols, new_rows, lam = ex.expand_from_h(H, dq, ALb, Cb, ARb, dm_max)

    b0, b1 = c.bonds[0], c.bonds[1]
    ALd, ARd, psi, _ = _dense_state(c, b0, b1, b0, AL, AR, _dense_diag(*b1, Cb))
    M, Phi = _projected_phi(H, ALd, ARd, psi)
    sv = np.linalg.svd(M, compute_uv=False)
    scale = max(sv[0], 1e-300)
    ours = np.sort(np.concatenate([v for v in lam.values()]))[::-1]
    nz = int(np.sum(sv > 1e-10 * scale))
    assert nz > 0
    # every nonzero singular value of the dense projected residual is one of the sectors'
    np.testing.assert_allclose(ours[:nz], sv[:nz], rtol=1e-10, atol=1e-12 * scale)
    assert np.all(ours[nz:] <= 1e-9 * scale)
    total = sum(dm.values())
    assert total == min(dm_max, len(ours))
    # the new bond: old sectors grown, new charges inserted in order
    keys, sizes = ex.grow_bond(b1[0], b1[1], dm)
    old = {tuple(int(v) for v in k): int(s) for k, s in zip(b1[0].reshape(len(b1[1]), -1), b1[1])}
    for q, s in zip(map(tuple, keys.reshape(len(sizes), -1)), sizes):
        assert s == old.get(q, 0) + dm.get(q, 0)
    assert [tuple(q) for q in keys.reshape(len(sizes), -1)] == sorted(tuple(q) for q in keys.reshape(len(sizes), -1))
    nb1 = (keys, sizes)
    layAL = _lay(orc.mps_structure(model, 2, QN, b0, nb1))
    layAR = _lay(orc.mps_structure(model, 2, QN, nb1, b0))
    ALn = _as_ob(layAL, ex.flatten_mps(layAL, ex.append_columns(ALb, new_cols)))
    ARn = _as_ob(layAR, ex.flatten_mps(layAR, ex.append_rows(ARb, new_rows)))
    Cn = ex.flatten_diag(BlockLayout.diag(AL.nsym, keys, sizes), Cb)
    Cnb = ex.diag_blocks(BlockLayout.diag(AL.nsym, keys, sizes), Cn)
    ALd2, ARd2, psi2, (Dl, Dc2, Dr) = _dense_state(c, b0, nb1, b0, ALn, ARn, _dense_diag(keys, sizes, Cnb))
    d = ALd2.shape[0]
    # isometries, and the state is unchanged (C padded with zeros)
    L2 = ALd2.reshape(d * Dl, Dc2)
    R2 = ARd2.transpose(1, 0, 2).reshape(Dc2, d * Dr)
    np.testing.assert_allclose(L2.T @ L2, np.eye(Dc2), atol=1e-12)
    np.testing.assert_allclose(R2 @ R2.T, np.eye(Dc2), atol=1e-12)
    np.testing.assert_allclose(psi2, psi, atol=1e-13)
    # the kept states are the largest singular directions: what remains of the projected
    # residual after the growth is the rest of the spectrum
    M2, _ = _projected_phi(H, ALd2, ARd2, psi)
    sv2 = np.linalg.svd(M2, compute_uv=False)
    rest = sv[total:]
    np.testing.assert_allclose(sv2[:len(rest)], rest, rtol=1e-9, atol=1e-11 * scale)


def test_expansion_adds_new_charges():
    """On the Heisenberg chain some center charges of H (AL C AR) lie outside the bond: with room
    for every state they enter as new sectors (GetDenseSplitMat's isnew branch)."""
    model, QN, c, C, H, dq = _setup("xxz", seed=11)
    ALb = ex.mps_blocks(_lay(c.AL[0]), c.AL[0].data)
    ARb = ex.mps_blocks(_lay(c.AR[1]), c.AR[1].data)
    Cb = ex.diag_blocks(_lay(C), C.data)
    dm, _, _, lam = ex.expand_from_h(H, dq, ALb, Cb, ARb, 1000)
    assert any(q not in Cb for q in dm)


def test_schmidt_values_are_block_singular_values():
    _, _, c, C, _, _ = _setup("fhub", seed=5)
    lam = ex.schmidt_values(_lay(C), C.data)
    dense = _dense_diag(*c.bonds[1], ex.diag_blocks(_lay(C), C.data))
    np.testing.assert_allclose(np.sort(np.concatenate(list(lam.values()))), np.sort(np.linalg.svd(dense, compute_uv=False)), atol=1e-12)


# ---- Truncate's selection (MPSBlockUtilities.hpp:846-906) ----------------------------------------
def _brute_keep(lam, mmax, thresh):
    allv = sorted(((v, q) for q, vs in lam.items() for v in vs if v > thresh), key=lambda t: -t[0])[:mmax]
    keep = {q: 0 for q in lam}
    for _, q in allv:
        keep[q] += 1
    return keep


@pytest.mark.parametrize("seed", range(6))
def test_truncate_sizes_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    lam = {(2 * i,): np.sort(rng.random(int(rng.integers(1, 9))) ** 3)[::-1] for i in range(-3, 4)}
    for mmax in (1, 5, 17, 1000):
        for thresh in (1e-10, 0.05, 0.3):
            keep, changed = ex.truncate_sizes(lam, mmax, thresh)
            assert keep == _brute_keep(lam, mmax, thresh)
            n_all = sum(len(v) for v in lam.values())
            assert changed == (sum(keep.values()) < n_all)


def test_truncate_sizes_rules():
    lam = {(0,): np.array([0.8, 0.3, 1e-13]), (2,): np.array([0.5, 1e-11]), (4,): np.array([1e-12])}
    keep, changed = ex.truncate_sizes(lam, 10)  # only the values below lamthresh 1e-10 go
    assert changed and keep == {(0,): 2, (2,): 1, (4,): 0}
    keep, changed = ex.truncate_sizes(lam, 2)   # then the mmax largest: 0.8 and 0.5
    assert keep == {(0,): 1, (2,): 1, (4,): 0}
    keep, changed = ex.truncate_sizes({(0,): np.array([0.9, 0.1])}, 5)
    assert not changed and keep == {(0,): 2}
    # a tie at the cut keeps the lower charge first (the reference's sort leaves the order open)
    keep, _ = ex.truncate_sizes({(0,): np.array([0.5]), (2,): np.array([0.5])}, 1)
    assert keep == {(0,): 1, (2,): 0}


def test_schmidt_basis_rotation_keeps_the_product():
    rng = np.random.default_rng(3)
    C = {(0,): rng.standard_normal((4, 4)), (2,): rng.standard_normal((2, 2))}
    U, S, V = ex.schmidt_basis(C)
    for q in C:
        np.testing.assert_allclose(U[q] @ np.diag(S[q]) @ V[q].T, C[q], atol=1e-13)
        assert np.all(np.diff(S[q]) <= 0)
    A = [{(-1,): ((0,), rng.standard_normal((3, 4)))}, {(1,): ((2,), rng.standard_normal((5, 2)))}]
    B = [{(0,): ((1,), rng.standard_normal((4, 6)))}, {(2,): ((3,), rng.standard_normal((2, 3)))}]
    Ar, Br = ex.rotate_right_index(A, U), ex.rotate_left_index(B, V)
    # A C B = (A U) S (V^T B) blockwise
    for s in range(2):
        for qi, (Q, a) in A[s].items():
            for t in range(2):
                if Q in B[t]:
                    ref = a @ C[Q] @ B[t][Q][1]
                    got = Ar[s][qi][1] @ np.diag(S[Q]) @ Br[t][Q][1]
                    np.testing.assert_allclose(got, ref, atol=1e-12)
    cut = ex.cut_right_index(Ar, {(0,): 2, (2,): 0})
    assert cut[0][(-1,)][1].shape == (3, 2) and not cut[1]
