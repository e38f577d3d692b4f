"""The second pass of the two-pass ground state runs only the shortest converged prefix of the first.

For a k-step Lanczos tridiagonal the residual of its lowest Ritz pair is b_k |y_{k-1}| (the standard
estimate; general/lanczos_diagonalization.py:81-121 leaves this to ARPACK).  Checked here on the CPU on the
recurrence of a random Hermitian matrix: the chosen prefix meets the bound, the one before it does not, and the
returned vector is that prefix's lowest eigenvector.
"""

import numpy as np
from scipy.linalg import eigh_tridiagonal


def _lanczos(A, v, n):
    a, b = [], [0.0]
    q_prev = np.zeros_like(v)
    q = v / np.linalg.norm(v)
    for _ in range(n):
        w = A @ q
        ak = np.vdot(q, w).real
        w = w - ak * q - b[-1] * q_prev
        a.append(ak)
        bk = np.linalg.norm(w)
        b.append(bk)
        q_prev, q = q, w / bk
    return np.array(a), np.array(b[:-1]), b[-1]


def test_prefix_choice():
    from exactdiag_amd.forms import _shortest_converged_prefix

    rng = np.random.default_rng(3)
    n = 400
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = (M + M.conj().T) / 2
    A[0, 0] -= 40.0  # an isolated lowest eigenvalue: quick convergence
    a, b, _ = _lanczos(A, rng.standard_normal(n) + 0j, 80)
    tol = 1e-8
    y = _shortest_converged_prefix(a, b, tol)
    k = len(y)
    assert 1 < k < 80

    def est(kk):
        th, Y = eigh_tridiagonal(a[:kk], b[1:kk], select="i", select_range=(0, 0))
        return abs(b[kk] * Y[-1, 0]), th[0], Y[:, 0]

    e_k, th, yk = est(k)
    assert e_k <= 0.25 * tol * max(1.0, abs(th))
    assert est(k - 1)[0] > 0.25 * tol * max(1.0, abs(est(k - 1)[1]))
    assert np.allclose(np.abs(y), np.abs(yk))
    # the Ritz value of that prefix is the lowest eigenvalue to the bound
    assert abs(th - np.linalg.eigvalsh(A)[0]) <= 1e-6


def test_prefix_without_convergence_uses_everything():
    from exactdiag_amd.forms import _shortest_converged_prefix

    rng = np.random.default_rng(4)
    n = 300
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2
    a, b, _ = _lanczos(A, rng.standard_normal(n) + 0j, 10)
    assert len(_shortest_converged_prefix(a, b, 1e-14)) == 10
