"""Exact diagonalisation of periodic chains of a two-site Hamiltonian (dense d^2 x d^2, pair index
s_l d + s_r, the models' convention), for energy pins of the models without closed forms.  Test
infrastructure: plain numpy / scipy, no sector logic of the library or the oracle."""
import numpy as np


def ring_energy(H2, d, L, fixed_total=None):
    """Lowest eigenvalue per site of sum_i H2(i, i+1) on a ring of L sites; with fixed_total the
    states are restricted to sum_i s_i = fixed_total (a U(1) charge whose local value is s)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    n = d ** L
    idx = np.arange(n, dtype=np.int64)
    w = [d ** (L - 1 - i) for i in range(L)]
    digits = [(idx // w[i]) % d for i in range(L)]
    rows, cols, vals = [], [], []
    for a, b in zip(*np.nonzero(np.abs(H2) > 0)):
        for i in range(L):
            j = (i + 1) % L
            src = idx[digits[i] * d + digits[j] == b]
            dst = src + (a // d - b // d) * w[i] + (a % d - b % d) * w[j]
            rows.append(dst)
            cols.append(src)
            vals.append(np.full(src.size, H2[a, b]))
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    if fixed_total is not None:
        keep = np.nonzero(sum(digits) == fixed_total)[0]
        M = M[keep][:, keep]
    return float(sla.eigsh(M, k=1, which="SA", tol=1e-14)[0][0]) / L


def tfim_energy(Jp, g, L=None):
    """Ground-state energy per site of -Jp sum sx_i sx_{i+1} - g sum sz_i (Pauli matrices): on a
    ring of L sites (even fermion parity, antiperiodic modes k = pi (2n + 1) / L) -sum_k e(k) / (2L)
    with e(k) = 2 sqrt(Jp^2 + g^2 - 2 Jp g cos k); without L the infinite chain,
    -(1/pi) int_0^pi e(k) / 2 dk, as the same midpoint sum at L = 8192 (exponentially accurate for
    the smooth periodic integrand of a gapped chain)."""
    e = lambda k: 2.0 * np.sqrt(Jp * Jp + g * g - 2.0 * Jp * g * np.cos(k))  # noqa: E731
    L = 8192 if L is None else L
    k = np.pi * (2 * np.arange(L) + 1) / L
    return -float(np.sum(e(k))) / (2.0 * L)


def tfim_dispersion(Jp, g, k):
    return 2.0 * np.sqrt(Jp * Jp + g * g - 2.0 * Jp * g * np.cos(k))
