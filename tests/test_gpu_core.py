"""GPU parity of the hot path against the CPU oracle (bit-exact where integer or exact)."""

import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SECTORS = [
    (2, 0, 2, 0, 0), (2, 1, 1, 1, 1), (3, 1, 2, 0, 0), (3, 2, 2, 0, 0), (3, 0, 3, 1, 1), (4, 2, 3, 1, 1),
    (4, 3, 2, 2, 0), (4, 0, 4, 2, 1), (5, 2, 4, 2, 0), (5, 3, 3, 2, 0), (6, 2, 5, 0, 0), (6, 1, 5, 3, 1),
    (6, 0, 6, 3, 0), (6, 0, 6, 0, 1), (6, 4, 4, 0, 1), (7, 2, 6, 3, 0), (4, 2, 3, -1, 0), (6, 0, 6, 7, 0),
]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _sectors(oracle_mod):
    from exactdiag_amd.sector import Sector

    for p in SECTORS:
        print("sector", p, flush=True)
        yield p, oracle_mod.OracleSector(*p), Sector(*p)


def test_basis_bit_exact(oracle_mod, torch):
    for p, o, g in _sectors(oracle_mod):
        so, co = o.basis()
        sg, cg = g.basis()
        assert o.dim == g.dim, p
        assert np.array_equal(so, sg), p
        assert np.array_equal(co, cg), p


def test_lookup_and_lowest(oracle_mod, torch):
    rng = np.random.default_rng(1)
    from math import comb

    for p, o, g in _sectors(oracle_mod):
        L, h, nd = p[0], p[1], p[2]
        total = comb(2 * L - h, nd) * comb(2 * L, h)
        s = np.unique(rng.integers(0, total, size=min(total, 3000)))
        dense = g.lookup(s)
        ref = np.array([o.dense(x) for x in s])
        assert np.array_equal(dense, ref), p
        rep, amp = g.lowest(s)
        for k, x in enumerate(s):
            st = o.unrank(x)
            ri, ra = o.lowest(st)
            assert rep[k] == ri, (p, x)
            if ri >= 0:
                assert amp[k] == ra, (p, x, amp[k], ra)


@pytest.mark.parametrize("triangle", [True, False])
def test_term_columns_bit_exact(oracle_mod, torch, triangle):
    for p, o, g in _sectors(oracle_mod):
        for term in range(4):
            po, ro, vo = o.term_columns(term, triangle)
            pg, rg, vg = g.term_columns(term, triangle)
            assert np.array_equal(po, pg), (p, term)
            assert np.array_equal(ro, rg), (p, term)
            assert np.array_equal(vo, vg), (p, term)


def test_hv_matches_oracle(oracle_mod, torch):
    rng = np.random.default_rng(2)
    for p, o, g in _sectors(oracle_mod):
        if o.dim == 0:
            continue
        for w in ([1, 1, 0.4, 0.4], [0.45, 0.36, 0.15, 0.12], [0, 0, 1, 1], [1, 0, 0, 0]):
            x = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
            yo = o.hv(w, x)
            yg = g.hv(w, torch.as_tensor(x, device="cuda")).cpu().numpy()
            err = np.max(np.abs(yo - yg)) / max(1.0, np.max(np.abs(yo)))
            assert err < 1e-13, (p, w, err)


def test_hv_dense_hermitian(oracle_mod, torch):
    for p, o, g in _sectors(oracle_mod):
        if o.dim == 0 or o.dim > 600:
            continue
        w = [1, 1, 0.4, 0.4]
        H = np.conj(o.toarray(w))  # the reference's toarray is the conjugate of the dot matrix
        eye = torch.eye(o.dim, dtype=torch.complex128, device="cuda")
        cols = [g.hv(w, eye[:, k].contiguous()).cpu().numpy() for k in range(o.dim)]
        Hg = np.stack(cols, axis=1)
        assert np.max(np.abs(Hg - H)) < 1e-13, p
        assert np.max(np.abs(Hg - Hg.conj().T)) < 1e-13, p
