"""Bit-exact agreement of the oracle with the reference's own C++ (oracle/_ref).

oracle/_ref/libedref.so is compiled from the reference's
tJ_spin_half_ladder/symmetry_ladder.h (and its includes) by oracle/Makefile;
it exists only where the reference checkout was available at build time.
"""

import itertools
from math import comb

import numpy as np
import pytest

from oracle import oracle
from oracle.oracle import OracleSector, RefSector

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference checkout)")


def all_small_sectors():
    for L in (2, 3, 4, 5):
        N = 2 * L
        for h in range(N + 1):
            for nd in range(N - h + 1):
                for kx, ky in itertools.product(range(-1, L + 1), range(2)):
                    yield (L, h, nd, kx, ky)


def test_basis_bit_exact_all_small_sectors():
    n = 0
    for p in all_small_sectors():
        so, co = OracleSector(*p, threads=1).basis()
        sr, cr = RefSector(*p, threads=1).basis()
        assert np.array_equal(so, sr) and np.array_equal(co, cr), p
        n += 1
    assert n > 500


@pytest.mark.parametrize("p", [(6, 2, 5, 0, 0), (6, 2, 5, 3, 1), (6, 0, 6, 2, 0), (7, 3, 5, 1, 1), (6, 4, 4, 0, 1),
                               (8, 2, 7, 4, 0), (7, 1, 6, -3, 0)])
def test_orbit_lowest_rank_bit_exact(p):
    o, r = OracleSector(*p), RefSector(*p)
    so, co = o.basis()
    sr, cr = r.basis()
    assert np.array_equal(so, sr) and np.array_equal(co, cr)
    L, h, nd = p[:3]
    total = comb(2 * L - h, nd) * comb(2 * L, h)
    rng = np.random.default_rng(7)
    for s in rng.choice(total, size=min(total, 1500), replace=False):
        st = o.unrank(s)
        assert np.array_equal(st, r.unrank(s))
        assert o.rank(st) == s == r.rank(st)
        assert o.lowest(st) == r.lowest(st)
        for cl in (False, True):
            a, b = o.orbit(st, cl), r.orbit(st, cl)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for hr in range(comb(2 * L, h)):
        assert o.hole_eligible(hr) == r.hole_eligible(hr)


def test_reference_sparse_dot_matches_oracle():
    o = OracleSector(6, 2, 5, 1, 0)
    rng = np.random.default_rng(3)
    x = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
    w = [0.45, 0.36, 0.15, 0.12]
    y_oracle = o.hv(w, x)
    y_ref = np.zeros(o.dim, np.complex128)
    first = True
    for t in range(4):
        values, vinds, minor, major = o.term_csr(t)
        oracle.ref_csr_dot(values, vinds, minor, major, x, y_ref, w[t], first, threads=4)
        first = False
    assert np.array_equal(y_oracle, y_ref)
