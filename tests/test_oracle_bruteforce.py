"""Brute-force second-quantised check of the oracle on tiny ladders.

Independent of the reference's algorithms: Fock states are ordered creation
strings (highest site leftmost), group elements act as site permutations
with the sign of the reordering, symmetric states come from the projector
sum_g exp(-i theta_g) U(g), and the Hamiltonian and operators are applied as
explicit fermion operators.  A dropped term, a wrong sign or a transposed
convention in the oracle fails one of these comparisons.
"""

import itertools
from math import sqrt

import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import OracleSector


def configs_of(N, h, nd):
    out = []
    for holes in itertools.combinations(range(N), h):
        rest = [i for i in range(N) if i not in holes]
        for downs in itertools.combinations(rest, nd):
            st = [2] * N
            for i in holes:
                st[i] = 0
            for i in downs:
                st[i] = 1
            out.append(tuple(st))
    return out


def key(st):
    H = sum(1 << i for i, v in enumerate(st) if v == 0)
    D = sum(1 << i for i, v in enumerate(st) if v == 1)
    return (H << 32) | D


def perm_site(i, L, m, b):
    x, y = divmod(i, 2)
    return 2 * ((x + m) % L) + (y ^ b)


def apply_group(st, L, m, b):
    """U(g)|st> = sign * |g st>, creation operators ordered by descending site."""
    N = 2 * L
    occ = [i for i in range(N - 1, -1, -1) if st[i] != 0]
    img = [perm_site(i, L, m, b) for i in occ]
    inv = sum(1 for a in range(len(img)) for c in range(a + 1, len(img)) if img[a] < img[c])
    new = [0] * N
    for i in range(N):
        new[perm_site(i, L, m, b)] = st[i]
    return tuple(new), (-1) ** inv


def ann_sign(st, p):
    """sign of moving c_p to the operator c^dagger_p in the descending string."""
    return (-1) ** sum(1 for i in range(p + 1, len(st)) if st[i] != 0)


class Fock:
    def __init__(self, L, h, nd):
        self.L, self.N = L, 2 * L
        self.states = configs_of(self.N, h, nd)
        self.index = {s: k for k, s in enumerate(self.states)}

    def hamiltonian(self, w):
        L, N = self.L, self.N
        dim = len(self.states)
        H = np.zeros((dim, dim), np.complex128)
        leg = [(i, (i + 2) % N) for i in range(N)]
        rung = [(2 * x, 2 * x + 1) for x in range(L)]
        for c, st in enumerate(self.states):
            for bonds, wt in ((leg, w[0]), (rung, w[1])):
                for a, b in bonds:
                    for t, s in ((a, b), (b, a)):  # c^dagger_t c_s
                        if st[t] != 0 or st[s] == 0:
                            continue
                        sg = ann_sign(st, s)
                        mid = list(st)
                        mid[t], mid[s] = st[s], 0
                        sg *= ann_sign(mid, t)
                        H[self.index[tuple(mid)], c] += -wt * sg
            for bonds, wt in ((leg, w[2]), (rung, w[3])):
                for a, b in bonds:
                    if st[a] == 0 or st[b] == 0 or st[a] == st[b]:
                        continue
                    H[c, c] += -0.5 * wt
                    sw = list(st)
                    sw[a], sw[b] = st[b], st[a]
                    H[self.index[tuple(sw)], c] += 0.5 * wt
        return H

    def symmetric_basis(self, kx, ky):
        L = self.L
        cols, reps = [], []
        seen = set()
        for st in sorted(self.states, key=key):
            orbit = {apply_group(st, L, m, b)[0] for m in range(L) for b in range(2)}
            r = min(orbit, key=key)
            if r in seen:
                continue
            seen.add(r)
            v = np.zeros(len(self.states), np.complex128)
            for m in range(L):
                for b in range(2):
                    img, sg = apply_group(r, L, m, b)
                    v[self.index[img]] += sg * np.exp(-2j * np.pi * (kx * m / L + ky * b / 2))
            nrm = np.linalg.norm(v)
            if nrm > 1e-9:
                cols.append(v / nrm)
                reps.append(r)
        order = np.argsort([key(r) for r in reps])
        return [reps[k] for k in order], np.stack([cols[k] for k in order], axis=1)


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


def _fock_operator(fk_i, fk_f, op, qx, qy):
    L, N = fk_i.L, fk_i.N
    O = np.zeros((len(fk_f.states), len(fk_i.states)), np.complex128)
    for c, st in enumerate(fk_i.states):
        for p in range(N):
            ph = np.exp(-2j * np.pi * (qx * (p // 2) / L + qy * (p % 2) / 2))
            if op == oracle.OP_SZQ:
                if st[p] != 0:
                    O[fk_f.index[st], c] += ph * (0.5 if st[p] == 2 else -0.5) / sqrt(N)
            elif op in (oracle.OP_APLUS_DOWN, oracle.OP_APLUS_UP):
                if st[p] == 0:
                    new = list(st)
                    new[p] = 1 if op == oracle.OP_APLUS_DOWN else 2
                    O[fk_f.index[tuple(new)], c] += ph * ann_sign(st, p) / sqrt(N)
            elif op in (oracle.OP_AMINUS_DOWN, oracle.OP_AMINUS_UP):
                want = 1 if op == oracle.OP_AMINUS_DOWN else 2
                if st[p] == want:
                    new = list(st)
                    new[p] = 0
                    O[fk_f.index[tuple(new)], c] += ph * ann_sign(st, p) / sqrt(N)
    return O


OPS = [((3, 2, 2, 0, 0), oracle.OP_SZQ, (1, 1), (3, 2, 2, 1, 1)),
       ((4, 2, 3, 1, 0), oracle.OP_SZQ, (2, 0), (4, 2, 3, -1, 0)),
       ((3, 1, 2, 0, 0), oracle.OP_APLUS_DOWN, (1, 0), (3, 0, 3, 1, 0)),
       ((4, 2, 3, 1, 1), oracle.OP_APLUS_UP, (2, 1), (4, 1, 3, -1, 0)),
       ((3, 1, 2, 1, 0), oracle.OP_AMINUS_DOWN, (-1, 1), (3, 2, 1, 0, 1)),
       ((4, 1, 3, 0, 1), oracle.OP_AMINUS_UP, (1, 0), (4, 2, 3, 1, 1))]


@pytest.mark.parametrize("case", OPS)
def test_operators(case):
    si, op, (qx, qy), sf = case
    fi, ff = Fock(si[0], si[1], si[2]), Fock(sf[0], sf[1], sf[2])
    _, Bi = fi.symmetric_basis(si[3], si[4])
    _, Bf = ff.symmetric_basis(sf[3], sf[4])
    Os = Bf.conj().T @ _fock_operator(fi, ff, op, qx, qy) @ Bi
    oi, of = OracleSector(*si), OracleSector(*sf)
    ptr, rows, vals = oi.op_columns(of, op, qx, qy)
    M = np.zeros((of.dim, oi.dim), np.complex128)
    for i in range(oi.dim):
        for k in range(ptr[i], ptr[i + 1]):
            M[rows[k], i] += vals[k]
    # the reference's q.r uses single-precision division (~1e-7 phase error)
    assert np.max(np.abs(M - Os)) < 2e-6, case
    assert np.max(np.abs(Os)) > 0.05
