"""Second-quantised brute force on tiny ladders (test infrastructure).

Independent of the reference's algorithms: Fock states are ordered creation
strings (highest site leftmost), group elements act as site permutations
with the sign of the reordering, symmetric states come from the projector
sum_g exp(-i theta_g) U(g) (optionally times the global spin flip with parity
z, which relabels spins without reordering any operator), and the Hamiltonian
and operators are applied as explicit fermion operators.
"""

import itertools
from math import sqrt

import numpy as np

from oracle import oracle


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


def flip_spins(st):
    """Every spin reversed (1 down <-> 2 up), holes kept."""
    return tuple(0 if v == 0 else 3 - v for v in st)


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

    def symmetric_basis(self, kx, ky, z=0, orbit_sizes=False):
        """Representatives (lowest key of each orbit with a nonzero projection) and the normalised symmetric
        states as columns, in key order; z = +1 or -1 adds the spin flip with that parity to the group."""
        L = self.L
        flips = (0, 1) if z else (0,)
        cols, reps, sizes = [], [], []
        seen = set()

        def images(st):
            for f in flips:
                for m in range(L):
                    for b in range(2):
                        img, sg = apply_group(st, L, m, b)
                        yield (flip_spins(img) if f else img), sg, m, b, f

        for st in sorted(self.states, key=key):
            orbit = {im for im, _, _, _, _ in images(st)}
            r = min(orbit, key=key)
            if r in seen:
                continue
            seen.add(r)
            v = np.zeros(len(self.states), np.complex128)
            for img, sg, m, b, f in images(r):
                v[self.index[img]] += sg * (z if f else 1) * np.exp(-2j * np.pi * (kx * m / L + ky * b / 2))
            nrm = np.linalg.norm(v)
            if nrm > 1e-9:
                cols.append(v / nrm)
                reps.append(r)
                sizes.append(len(orbit))
        order = np.argsort([key(r) for r in reps])
        out = [reps[k] for k in order], np.stack([cols[k] for k in order], axis=1)
        return out + ([sizes[k] for k in order],) if orbit_sizes else out


def _hop(st, t, p):
    """c^dagger_t c_p (same spin) on a Fock state: (sign, new state) or None."""
    if st[t] != 0 or st[p] == 0:
        return None
    sg = ann_sign(st, p)
    mid = list(st)
    mid[t], mid[p] = st[p], 0
    return sg * ann_sign(mid, t), tuple(mid)


def _fock_operator(fk_i, fk_f, op, qx, qy, mutate=None):
    """Excitation operators from explicit fermion operators.  qy is the physical momentum transfer across the
    legs; for the rung current the final sector's ky is shifted by one more (the current is odd under the leg
    mirror).  `mutate` plants a mistake (a dropped fermion sign, a flipped endpoint sign, a missing i, a
    rung current without the mirror sign) to show that the comparison catches it."""
    L, N = fk_i.L, fk_i.N
    O = np.zeros((len(fk_f.states), len(fk_i.states)), np.complex128)
    for c, st in enumerate(fk_i.states):
        for p in range(N):
            ph = np.exp(-2j * np.pi * (qx * (p // 2) / L + qy * (p % 2) / 2))
            if op == oracle.OP_SZQ:
                if st[p] != 0:
                    O[fk_f.index[st], c] += ph * (0.5 if st[p] == 2 else -0.5) / sqrt(N)
            elif op == oracle.OP_NQ:
                # N(q) = N^-1/2 sum_p e^{-iq.r_p} n_p
                if st[p] != 0:
                    O[fk_f.index[st], c] += ph / sqrt(N)
            elif op == oracle.OP_CURRENT_RUNG:
                # i sum_x e^{-i qx x} (c^dag_{x,1} c_{x,0} - (-1)^qy c^dag_{x,0} c_{x,1})
                x, y = divmod(p, 2)
                coef = 1j * np.exp(-2j * np.pi * qx * x / L)
                if y == 1 and mutate != "rung_mirror":
                    coef *= -((-1) ** qy)
                res = _hop(st, 2 * x + (1 - y), p)
                if res is not None:
                    O[fk_f.index[res[1]], c] += coef * res[0]
            elif op == oracle.OP_CURRENT_LEG:
                # i sum_p e^{-iq.r_p} (c^dag_{p+x} c_p - c^dag_{p-x} c_p), fermion signs from the algebra
                x, y = divmod(p, 2)
                for dx, e in ((1, 1.0), (-1, -1.0)):
                    if mutate == "leg_endpoint" and dx == -1:
                        e = 1.0
                    res = _hop(st, 2 * ((x + dx) % L) + y, p)
                    if res is None:
                        continue
                    sg = 1 if mutate == "no_sign" else res[0]
                    i_factor = 1.0 if mutate == "no_i" else 1j
                    O[fk_f.index[res[1]], c] += i_factor * ph * e * sg
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
