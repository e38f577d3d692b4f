"""Brute-force second-quantised check of the oracle on tiny ladders.

Independent of the reference's algorithms: Fock states are ordered creation
strings (highest site leftmost), group elements act as site permutations
with the sign of the reordering, symmetric states come from the projector
sum_g exp(-i theta_g) U(g), and the Hamiltonian and operators are applied as
explicit fermion operators.  A dropped term, a wrong sign or a transposed
convention in the oracle fails one of these comparisons.
"""

import numpy as np
import pytest

from fock import Fock, _fock_operator, ann_sign
from oracle import oracle
from oracle.oracle import OracleSector


SECTORS = [(3, 1, 2, 0, 0), (3, 1, 2, 1, 1), (3, 2, 2, 1, 0), (3, 2, 1, 2, 1), (4, 2, 3, 1, 1), (4, 2, 3, 2, 0),
           (4, 1, 3, 0, 1), (4, 3, 2, 3, 1), (3, 0, 3, 1, 0), (4, 0, 4, 2, 1), (3, 2, 2, 0, 1)]


@pytest.mark.parametrize("sector", SECTORS)
def test_basis_and_hamiltonian(sector):
    L, h, nd, kx, ky = sector
    fk = Fock(L, h, nd)
    reps, B = fk.symmetric_basis(kx, ky)
    o = OracleSector(*sector)
    sparse, counts = o.basis()
    assert [o.rank(np.array(r, np.int16)) for r in reps] == list(sparse)
    for t in range(4):
        w = np.zeros(4)
        w[t] = 1.0
        Hs = B.conj().T @ fk.hamiltonian(w) @ B
        M = np.conj(o.toarray(w))  # the matrix applied by dot
        assert np.max(np.abs(M - Hs)) < 1e-12, (sector, t)
        y = o.hv(w, np.arange(o.dim) + 1j)
        assert np.max(np.abs(y - Hs @ (np.arange(o.dim) + 1j))) < 1e-11


OPS = [((3, 2, 2, 0, 0), oracle.OP_SZQ, (1, 1), (3, 2, 2, 1, 1)),
       ((4, 2, 3, 1, 0), oracle.OP_SZQ, (2, 0), (4, 2, 3, -1, 0)),
       ((3, 1, 2, 0, 0), oracle.OP_APLUS_DOWN, (1, 0), (3, 0, 3, 1, 0)),
       ((4, 2, 3, 1, 1), oracle.OP_APLUS_UP, (2, 1), (4, 1, 3, -1, 0)),
       ((3, 1, 2, 1, 0), oracle.OP_AMINUS_DOWN, (-1, 1), (3, 2, 1, 0, 1)),
       ((4, 1, 3, 0, 1), oracle.OP_AMINUS_UP, (1, 0), (4, 2, 3, 1, 1))]


# N(q) and the two currents (matrix_setup.pyx:248-279).  The rung current is called with qy + 1, as
# setup_excitation_operator does (matrix_setup.pyx:251), and its final sector has ky + qy + 1.
OPS += [((3, 1, 2, 0, 0), oracle.OP_NQ, (1, 0), (3, 1, 2, 1, 0)),
        ((4, 2, 3, 1, 1), oracle.OP_NQ, (2, 1), (4, 2, 3, -1, 0)),
        ((3, 1, 2, 0, 0), oracle.OP_CURRENT_RUNG, (0, 0), (3, 1, 2, 0, 1)),
        ((4, 2, 3, 1, 0), oracle.OP_CURRENT_RUNG, (1, 1), (4, 2, 3, 2, 0)),
        ((4, 1, 3, 0, 1), oracle.OP_CURRENT_RUNG, (2, 0), (4, 1, 3, 2, 0)),
        ((3, 1, 2, 0, 0), oracle.OP_CURRENT_LEG, (1, 0), (3, 1, 2, 1, 0)),
        ((4, 2, 3, 1, 1), oracle.OP_CURRENT_LEG, (1, 1), (4, 2, 3, 2, 0)),
        ((4, 1, 3, 0, 0), oracle.OP_CURRENT_LEG, (2, 1), (4, 1, 3, 2, 1)),
        ((3, 2, 2, 1, 0), oracle.OP_CURRENT_LEG, (0, 0), (3, 2, 2, 1, 0))]


def _operator_pair(case, mutate=None):
    """(brute-force operator in the symmetric bases, the oracle's operator matrix)."""
    si, op, (qx, qy), sf = case
    fi, ff = Fock(si[0], si[1], si[2]), Fock(sf[0], sf[1], sf[2])
    _, Bi = fi.symmetric_basis(si[3], si[4])
    _, Bf = ff.symmetric_basis(sf[3], sf[4])
    Os = Bf.conj().T @ _fock_operator(fi, ff, op, qx, qy, mutate) @ Bi
    oi, of = OracleSector(*si), OracleSector(*sf)
    ptr, rows, vals = oi.op_columns(of, op, qx, qy + (1 if op == oracle.OP_CURRENT_RUNG else 0))
    M = np.zeros((of.dim, oi.dim), np.complex128)
    for i in range(oi.dim):
        for k in range(ptr[i], ptr[i + 1]):
            M[rows[k], i] += vals[k]
    return Os, M


@pytest.mark.parametrize("case", OPS)
def test_operators(case):
    Os, M = _operator_pair(case)
    # the reference's q.r uses single-precision division (~1e-7 phase error; pinned bit for bit against the
    # reference's own get_q_dot_r in test_oracle_ref.py)
    assert np.max(np.abs(M - Os)) < 2e-6, case
    assert np.max(np.abs(Os)) > 0.05


@pytest.mark.parametrize("case,mutate", [(OPS[-3], "no_sign"), (OPS[-2], "leg_endpoint"), (OPS[-4], "no_i"),
                                         (OPS[-5], "rung_mirror"), (OPS[-7], "rung_mirror")])
def test_operator_check_catches_planted_mistakes(case, mutate):
    """A dropped fermion sign, a flipped endpoint weight, a missing factor i or a rung current without its
    mirror sign in the brute force moves it away from the oracle: the comparison above has the power to
    see such a slip on either side."""
    Os, M = _operator_pair(case, mutate)
    assert np.max(np.abs(M - Os)) > 1e-2, (case, mutate)


def _site(L, i, leg, rung):
    """get_index_by_relative_shift on the ladder (symmetry_ladder.h:103-111 with symmetry_generator.h:47-59)."""
    return 2 * ((i // 2 + leg) % L) + (i % 2 + rung) % 2


def _apply(st, ops):
    """Apply a product of fermion operators (rightmost first) to a Fock state: (sign, new state) or None.
    ops: list of (site, spin, create) with spin 1 down, 2 up."""
    st = list(st)
    sign = 1
    for site, spin, create in reversed(ops):
        if create:
            if st[site] != 0:
                return None
            sign *= ann_sign(st, site)
            st[site] = spin
        else:
            if st[site] != spin:
                return None
            sign *= ann_sign(st, site)
            st[site] = 0
    return sign, tuple(st)


def _fock_singlet(fk, y, x, j):
    """(0.5 / N) sum_i Delta_i^dagger(y) Delta_{i+j}(x), Delta_i(x) = c_{i up} c_{i+x down} - c_{i down} c_{i+x up},
    from explicit fermion operators."""
    L, N = fk.L, fk.N
    dim = len(fk.states)
    O = np.zeros((dim, dim), np.complex128)
    for c, st in enumerate(fk.states):
        for i in range(N):
            iy = _site(L, i, *y)
            a = _site(L, i, *j)
            ax = _site(L, a, *x)
            creations = [(+1, [(iy, 1, True), (i, 2, True)]), (-1, [(iy, 2, True), (i, 1, True)])]
            annihilations = [(+1, [(a, 2, False), (ax, 1, False)]), (-1, [(a, 1, False), (ax, 2, False)])]
            for s1, cr in creations:
                for s2, an in annihilations:
                    res = _apply(st, cr + an)
                    if res is None:
                        continue
                    sg, new = res
                    O[fk.index[new], c] += 0.5 / N * s1 * s2 * sg
    return O


def _fock_diagonal(fk, kind, shifts):
    """Hole projection (kind 0) or S^z product (kind 1) over positions (r, r + shifts...) summed over r."""
    L, N = fk.L, fk.N
    d = np.zeros(len(fk.states))
    for c, st in enumerate(fk.states):
        for r in range(N):
            pos = [r] + [_site(L, r, *s) for s in shifts]
            if kind == 0:
                d[c] += all(st[p] == 0 for p in pos)
            elif all(st[p] != 0 for p in pos):
                d[c] += np.prod([-1 if st[p] == 1 else 1 for p in pos])
    return np.diag(d).astype(np.complex128)


POSITION_CASES = [
    ((3, 2, 2, 0, 0), 2, [(1, 1), (1, 1), (1, 0)]),
    ((4, 2, 3, 0, 0), 2, [(1, 0), (1, 1), (2, 1)]),
    ((4, 2, 3, 1, 1), 2, [(0, 1), (1, 0), (1, 1)]),
    ((4, 2, 3, 2, 0), 2, [(1, 1), (0, 1), (0, 0)]),
    ((3, 3, 1, 1, 1), 2, [(0, 1), (0, 1), (2, 1)]),
    ((4, 2, 3, 0, 1), 0, [(1, 1)]),
    ((4, 3, 2, 1, 0), 0, [(1, 0), (2, 1)]),
    ((4, 2, 3, 1, 0), 1, [(0, 1)]),
    ((4, 1, 3, 2, 1), 1, [(1, 1), (0, 0)]),
]


@pytest.mark.parametrize("case", POSITION_CASES)
def test_position_operators(case):
    """get_position_correlation_operator (matrix_setup.pyx:409-480) against its second-quantised definition."""
    sec, kind, shifts = case
    fk = Fock(sec[0], sec[1], sec[2])
    _, B = fk.symmetric_basis(sec[3], sec[4])
    if kind == 2:
        Of = _fock_singlet(fk, shifts[0], shifts[1], shifts[2])
    else:
        Of = _fock_diagonal(fk, kind, shifts)
    Os = B.conj().T @ Of @ B
    o = OracleSector(*sec)
    ptr, rows, vals = o.position_columns(kind, shifts)
    M = np.zeros((o.dim, o.dim), np.complex128)
    for i in range(o.dim):
        for k in range(ptr[i], ptr[i + 1]):
            M[rows[k], i] += vals[k]
    assert np.max(np.abs(M - Os)) < 1e-12, case
    assert np.max(np.abs(Os)) > 1e-3, case


@pytest.mark.parametrize("sec,kind,fixed", [((4, 2, 3, 0, 0), 0, []), ((4, 3, 2, 1, 1), 0, [(1, 1)]),
                                            ((4, 1, 3, 2, 0), 1, []), ((3, 2, 2, 1, 1), 1, [(0, 1)])])
def test_position_correlations(sec, kind, fixed):
    """calculate_hole_spin_projection_correlations (position_correlations.pyx:84-127, :186-230) on a random
    symmetric state = <Psi| sum_r P_r(d) |Psi> from the Fock-space diagonal operator, for every node index d."""
    fk = Fock(sec[0], sec[1], sec[2])
    _, B = fk.symmetric_basis(sec[3], sec[4])
    o = OracleSector(*sec)
    rng = np.random.default_rng(5)
    psi = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
    psi /= np.linalg.norm(psi)
    got = o.position_correlations(kind, fixed, psi)
    Psi = B @ psi
    for d in range(2 * sec[0]):
        Od = _fock_diagonal(fk, kind, [(d // 2, d % 2)] + list(fixed))
        ref = np.real(np.vdot(Psi, Od @ Psi))
        assert abs(got[d] - ref) < 1e-12, (sec, kind, d)


# -- the spin flip (no counterpart in the reference; pins the brute-force construction the GPU's spin-flip
#    sectors are checked against) ---------------------------------------------------------------------------

@pytest.mark.parametrize("sector", [(3, 0, 3, 0, 0), (4, 2, 3, 1, 1), (4, 0, 4, 2, 0), (3, 2, 2, 1, 1), (4, 2, 3, 0, 0)])
def test_spin_flip_blocks_split_the_momentum_sector(sector):
    """The spin flip commutes with every term of H, its two parity blocks of a momentum sector hold together
    exactly the sector's states (orthonormal, mutually orthogonal) and their spectra together are the sector's;
    Szq turns parity z into -z and N(q) and the currents keep it."""
    from fock import flip_spins

    L, h, nd, kx, ky = sector
    fk = Fock(L, h, nd)
    P = np.zeros((len(fk.states), len(fk.states)))
    for c, st in enumerate(fk.states):
        P[fk.index[flip_spins(st)], c] = 1.0
    w = [1.0, 0.7, 0.4, 0.3]
    Hf = fk.hamiltonian(w)
    assert np.max(np.abs(P @ Hf @ P.T - Hf)) < 1e-12
    _, B = fk.symmetric_basis(kx, ky)
    parts = [fk.symmetric_basis(kx, ky, z)[1] for z in (1, -1)]
    assert parts[0].shape[1] + parts[1].shape[1] == B.shape[1]
    Bz = np.concatenate(parts, axis=1)
    assert np.max(np.abs(Bz.conj().T @ Bz - np.eye(Bz.shape[1]))) < 1e-12
    assert np.max(np.abs(Bz @ Bz.conj().T - B @ B.conj().T)) < 1e-12  # the same subspace
    ev = np.linalg.eigvalsh(B.conj().T @ Hf @ B)
    evz = np.sort(np.concatenate([np.linalg.eigvalsh(Bp.conj().T @ Hf @ Bp) for Bp in parts]))
    assert np.max(np.abs(ev - evz)) < 1e-10
    for op, qy_shift, parity in ((oracle.OP_SZQ, 0, -1), (oracle.OP_NQ, 0, 1), (oracle.OP_CURRENT_LEG, 0, 1),
                                 (oracle.OP_CURRENT_RUNG, 1, 1)):
        O = _fock_operator(fk, fk, op, 1 % L, 0)
        assert np.max(np.abs(P @ O @ P.T - parity * O)) < 1e-12, op
