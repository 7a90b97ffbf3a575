"""Z(q) charges and the quantum Ising, quantum Potts and Bose-Hubbard chains on the CPU.

Bookkeeping (Models.cpp:325-507, 739-878; symobj.hpp:80-161): the library's answers equal the
oracle's exactly and the answers derived by hand from the reference's source text:
  * QIM: group Z(2), I2K = [{0}, {1}] (state s carries charge s), validpath [0, 0], GetDimMap
    exactly two sectors {0}, {1} (more dims are cut to two, fewer are an error);
  * QPOTTS(q): group Z(q), I2K = [{0}, ..., {q-1}]; an MPS block q_l -> q_l + s wraps modulo q;
  * BHUB(d): group U(1); N = 2, n = 2: dQ = (N/gcd, -n/gcd) = (1, -1), I2K = [-1, 0, 1, 2] for d = 4,
    validpath [2, 0], bond 1 = bond 0 - 1; N = 2, n = 1: dQ = (2, -1), I2K = [-1, 1, 3] for d = 3,
    GetDimMap sectors at 0, 2, 4, ...
Hamiltonians against independent constructions: QIM from Pauli matrices (sx = X/2, sz = Z/2);
QPOTTS(q) as the clock model -sum_a (S x S^dag)^a - h/2 (T x 1 + 1 x T) with the cyclic shift S
and T = sum_{a=1}^{q-1} tau^a, tau = diag(w^j) (complex, w = e^{2 pi i / q}); QPOTTS(2) = QIM(J = 4,
h = 2h'); BHUB(2) (hard-core bosons) = XXZ(J = 2t, Delta = 0, hz = -mu) - mu/2; BHUB at t = 0 has the
spectrum of its on-site energies.  The exact-diagonalisation helper of the GPU energy pins equals
the TFIM ring formula.  Layout checks: a Z(q) charge outside [0, q) and layouts of different groups
are rejected.  A pin of the oracle's pins: with U(1) arithmetic on a Z(3) cell the dense transfer-
matrix pin fails.
"""
import numpy as np
import pytest

import imps_cpp_tools_amd as im
import exact_diag as ed
from oracle import pyoracle as orc
from states import Cell, dense_mat, dense_mps, extract, random_like

MODELS = [(im.QIM, [], [3, 2]), (im.QPOTTS(3), [], [2, 3, 1]), (im.QPOTTS(4), [], [1, 2, 3, 4]), (im.BHUB(4), [2], [1, 2, 3]),
          (im.BHUB(3), [1], [2, 2]), (im.BHUB(5), [3], [1, 2, 2, 1])]


@pytest.mark.parametrize("model,QN,dims", MODELS, ids=[hex(m[0]) for m in MODELS])
@pytest.mark.parametrize("N", [2, 3])
def test_bookkeeping_equals_oracle(model, QN, dims, N):
    assert np.array_equal(im.i2k(model, N, QN), orc.i2k(model, N, QN))
    assert np.array_equal(im.validpath(model, N, QN), orc.validpath(model, N, QN))
    k, s = im.dimmap(model, N, QN, dims)
    ob = orc.make_dims_model(model, dims, N, QN)
    bonds = im.make_dims(model, N, QN, k, s)
    for (ka, sa), (kb, sb) in zip(bonds, ob):
        assert np.array_equal(ka, kb) and np.array_equal(sa, sb)
    for n in range(N):
        a = im.mps_structure(model, N, QN, bonds[n], bonds[(n + 1) % N])
        b = orc.mps_structure(model, N, QN, ob[n], ob[(n + 1) % N])
        for f in ("nblk", "qin", "qout", "rows", "cols"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert np.array_equal(a.periods, im.model_periods(model))


def test_known_answers():
    assert im.model_periods(im.QIM).tolist() == [2] and im.model_d(im.QIM) == 2
    assert im.model_periods(im.QPOTTS(5)).tolist() == [5] and im.model_d(im.QPOTTS(5)) == 5
    assert im.model_periods(im.BHUB(6)).tolist() == [0] and im.model_d(im.BHUB(6)) == 6
    assert im.model_periods(im.FHUB).tolist() == [0, 0] and im.model_periods(im.XXZ).tolist() == [0]
    assert im.model_d(im.QPOTTS(1)) == -1 and im.model_d(7) == -1
    # QIM
    assert im.i2k(im.QIM, 2, []).tolist() == [[0], [1]]
    assert im.validpath(im.QIM, 2, []).tolist() == [0, 0]
    k, s = im.dimmap(im.QIM, 2, [], [3, 2, 7])  # a third sector is dropped
    assert k.ravel().tolist() == [0, 1] and s.tolist() == [3, 2]
    with pytest.raises(im.ImpsError):
        im.dimmap(im.QIM, 2, [], [3])
    # QPOTTS(3): blocks q -> q + s (mod 3)
    b = im.make_dims(im.QPOTTS(3), 2, [], [[0], [1], [2]], [2, 3, 1])
    A = im.mps_structure(im.QPOTTS(3), 2, [], b[0], b[1])
    assert A.nblk.tolist() == [3, 3, 3]
    assert A.qout.ravel().tolist() == [0, 1, 2, 1, 2, 0, 2, 0, 1]
    assert A.rows.tolist() == [2, 3, 1] * 3 and A.cols.tolist() == [2, 3, 1, 3, 1, 2, 1, 2, 3]
    # BHUB(4), N = 2, n = 2
    assert im.i2k(im.BHUB(4), 2, [2]).ravel().tolist() == [-1, 0, 1, 2]
    assert im.validpath(im.BHUB(4), 2, [2]).tolist() == [2, 0]
    b = im.make_dims(im.BHUB(4), 2, [2], [[0], [1], [2]], [1, 2, 3])
    assert b[1][0].ravel().tolist() == [-1, 0, 1] and b[1][1].tolist() == [1, 2, 3]
    A = im.mps_structure(im.BHUB(4), 2, [2], b[0], b[1])
    assert A.nblk.tolist() == [3, 2, 1, 0]
    # BHUB(3), N = 2, n = 1
    assert im.i2k(im.BHUB(3), 2, [1]).ravel().tolist() == [-1, 1, 3]
    k, s = im.dimmap(im.BHUB(3), 2, [1], [2, 5, 4])
    assert k.ravel().tolist() == [0, 2, 4] and s.tolist() == [2, 5, 4]
    with pytest.raises(im.ImpsError):
        im.i2k(im.BHUB(3), 2, [6])  # 0 < n < N d


def test_hamiltonians():
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Z = np.diag([1.0, -1.0])
    I2 = np.eye(2)
    J, h = 0.7, 1.3
    ref = -(J / 4) * np.kron(X, X) - (h / 4) * (np.kron(Z, I2) + np.kron(I2, Z))
    assert np.abs(im.model_ham(im.QIM, [J, h]) - ref).max() < 1e-15
    for q, hp in ((2, 0.9), (3, 1.7), (5, 0.4)):
        S = np.roll(np.eye(q), -1, axis=0)  # S|j> = |j - 1 mod q>
        w = np.exp(2j * np.pi / q)
        tau = np.diag(w ** np.arange(q))
        T = sum(np.linalg.matrix_power(tau, a) for a in range(1, q))
        ref = -sum(np.linalg.matrix_power(np.kron(S, S.conj().T), a) for a in range(1, q)) \
            - hp / 2 * (np.kron(T, np.eye(q)) + np.kron(np.eye(q), T))
        assert np.abs(ref.imag).max() < 1e-12
        Hq = im.model_ham(im.QPOTTS(q), [hp])
        assert np.abs(Hq - ref.real).max() < 1e-12, q
        assert np.array_equal(Hq, orc.model_ham(im.QPOTTS(q), [hp]))
    assert np.abs(im.model_ham(im.QPOTTS(2), [0.9]) - im.model_ham(im.QIM, [4.0, 1.8])).max() < 1e-15
    t, U, mu = 0.8, 3.0, 0.35
    ref = im.model_ham(im.XXZ, [2 * t, 0.0, -mu]) - mu / 2 * np.eye(4)
    # hard-core boson |0> (empty) <-> spin up, |1> <-> spin down: the same matrix
    assert np.abs(im.model_ham(im.BHUB(2), [t, U, mu]) - ref).max() < 1e-15
    for d in (3, 4, 6):
        e = np.array([0.5 * U * n * (n - 1) - mu * n for n in range(d)])
        Hb = im.model_ham(im.BHUB(d), [0.0, U, mu])
        assert np.abs(np.sort(np.linalg.eigvalsh(Hb)) - np.sort(0.5 * np.add.outer(e, e).ravel())).max() < 1e-12
        assert np.abs(im.model_ham(im.BHUB(d), [t, U, mu]) - orc.model_ham(im.BHUB(d), [t, U, mu])).max() < 1e-14
    assert np.array_equal(im.model_ham(im.QIM, [J, h]), orc.model_ham(im.QIM, [J, h]))


def test_exact_diag_helper_equals_tfim_ring():
    H = im.model_ham(im.QIM, [1.0, 1.5])  # J' = 1/4, g = 3/4
    for L in (6, 9, 12):
        assert abs(ed.ring_energy(H, 2, L) - ed.tfim_energy(0.25, 0.75, L)) < 1e-13
    # U(1)-restricted: two hard-core bosons on a ring of 6 = two free fermions of XX hopping
    Hb = im.model_ham(im.BHUB(2), [1.0, 0.0, 0.0])
    k = 2 * np.pi * np.arange(6) / 6 + np.pi / 6  # N = 2 (even): antiperiodic fermions
    assert abs(ed.ring_energy(Hb, 2, 6, fixed_total=2) * 6 - np.sort(-2 * np.cos(k))[:2].sum()) < 1e-12


def _lay(nsym, keys, sizes, periods):
    return im.BlockLayout.diag(nsym, keys, sizes, periods)


def test_zq_layouts_checked():
    b = im.make_dims(im.QIM, 2, [], [[0], [1]], [3, 2])
    A = im.mps_structure(im.QIM, 2, [], b[0], b[1])
    x = _lay(1, [[0], [1]], [3, 2], [2])
    p = im.TMPlan(None, im.LEFT, A, None, x, x)  # host-only plan
    assert p.cost()[0] > 0
    with pytest.raises(im.ImpsError):  # charge 2 is no Z(2) charge
        im.TMPlan(None, im.LEFT, A, None, _lay(1, [[0], [2]], [3, 2], [2]), x)
    with pytest.raises(im.ImpsError):  # a U(1) vector for a Z(2) tensor
        im.TMPlan(None, im.LEFT, A, None, _lay(1, [[0], [1]], [3, 2], None), x)
    with pytest.raises(ValueError):
        im.BlockLayout.diag(1, [[0]], [3], [2, 2])


def test_oracle_pins_see_the_wraparound():
    """The dense right-TM pin of the oracle (y = sum_s A_s x A_s^T, whose sector lookup subtracts the
    slice's charge shift) fails when the Z(3) charges are subtracted without the modulus."""
    model = im.QPOTTS(3)
    bd = orc.make_dims_model(model, [3, 4, 2], 2, [])
    c = Cell(model, 2, [], bd[0][0], bd[0][1], seed=17)
    rng = np.random.default_rng(5)
    lo, Dl = c.bond_offsets(0)
    ro, Dr = c.bond_offsets(1)
    A = random_like(c.struct[0], rng)
    x = random_like(c.diag(1), rng)
    Ad = dense_mps(A, lo, Dl, ro, Dr)
    want = extract(c.diag(0), np.einsum("sab,bc,sdc->ad", Ad, dense_mat(x, ro, Dr, ro, Dr), Ad), lo, lo)
    assert np.abs(orc.tm_mixed(1, A, None, x, c.diag(0)).data - want).max() < 1e-12

    class U1(type(A)):
        def handle(self):
            h = super().handle()
            orc.set_group(self.nsym, None)  # the planted mistake: U(1) differences for a Z(3) charge
            return h
    bad = [U1(o.nsym, o.nblk, o.qin, o.qout, o.rows, o.cols, o.data, o.periods) for o in (A, x, c.diag(0))]
    got = orc.tm_mixed(1, bad[0], None, bad[1], bad[2]).data
    assert np.abs(got - want).max() > 1e-3
