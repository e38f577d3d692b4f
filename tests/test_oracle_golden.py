"""Pin the CPU oracle against the reference's own fixtures and tests.

Mirrors tests/tJ_spin_half_ladder/test_basis_size.py and test_eigvals.py of the
reference: 68 basis-size fixtures, k <-> -k invariance, k periodicity, the
sum over k and the sum over holes and S^z, and 51 eigenvalue fixtures (8
decimals, tolerance 1e-8).
"""

import glob
import itertools
import json
import pathlib
from math import comb

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle.oracle import OracleSector

GOLD = pathlib.Path(__file__).parent / "golden" / "precalculated_basis_size_eigenvalues"
EIGENVALUE_TOLERANCE = 1e-8


def collect_all():
    return sorted(glob.glob(str(GOLD / "*.txt")))


def load(path):
    with open(path, encoding="utf-8") as fp:
        return json.load(fp)


def sector_of(d, kx=None):
    k = d["symmetry_qs"]
    return (d["num_rungs"], d["num_holes"], d["num_down_spins"], k[0] if kx is None else kx, k[1])


def collect(num_rungs, num_holes, num_down_spins):
    out = []
    for kx in range(num_rungs // 2 + 1):
        for ky in range(2):
            out.append(load(GOLD / f"rungs{num_rungs}_holes{num_holes}_k[{kx},{ky}]_down{num_down_spins}.txt"))
    return out


def test_against_precalculated():
    files = collect_all()
    assert len(files) == 68
    for f in files:
        d = load(f)
        assert OracleSector(*sector_of(d)).dim == d["basis_size"], f


def test_negative_k_leg():
    coll = collect(6, 0, 6)
    assert len(coll) == 8
    for d in coll:
        k = d["symmetry_qs"]
        if k[0] == 0:
            continue
        assert OracleSector(*sector_of(d, kx=-k[0])).dim == d["basis_size"]


def test_k_leg_periodicity():
    for d in collect(6, 0, 6):
        k = d["symmetry_qs"]
        if tuple(k) == (0, 0):
            continue
        for m in range(-4, 5):
            assert OracleSector(*sector_of(d, kx=k[0] + m * 6)).dim == d["basis_size"]


def test_sum_over_k():
    groups = {}
    for f in collect_all():
        d = load(f)
        key = (d["num_rungs"], d["num_holes"], d["num_down_spins"])
        kx = d["symmetry_qs"][0]
        mult = 1 + (kx != 0 and (kx != d["num_rungs"] // 2 or d["num_rungs"] % 2))
        groups.setdefault(key, []).append(mult * d["basis_size"])
    groups = {k: v for k, v in groups.items() if len(v) == 2 * (k[0] // 2 + 1)}
    assert len(groups) == 7
    for (L, h, nd), sizes in groups.items():
        assert comb(2 * L - h, nd) * comb(2 * L, h) == sum(sizes)


def test_sum_over_down_spins():
    L, N = 6, 12
    for h in range(N + 1):
        total = 0
        for nd in range(N - h + 1):
            for kx, ky in itertools.product(range(L // 2 + 1), range(2)):
                mult = 1 + (kx != 0 and (kx != L // 2 or L % 2))
                total += mult * OracleSector(L, h, nd, kx, ky).dim
        assert total == 2 ** (N - h) * comb(N, h), h


def _lowest(o, w, nev):
    if o.dim <= 3000:
        return np.linalg.eigvalsh(o.toarray(w))[:nev]
    op = spla.LinearOperator((o.dim, o.dim), matvec=lambda x: o.hv(w, x), dtype=np.complex128)
    vals = spla.eigsh(op, k=nev, which="SA", tol=1e-12, ncv=40)[0]
    return np.sort(vals)


def test_eigenvalues_against_precalculated():
    n = 0
    for f in collect_all():
        d = load(f)
        if "eigvals" not in d:
            continue
        n += 1
        w = [d["weights"][t] for t in ("tl", "tr", "jl", "jr")]
        o = OracleSector(*sector_of(d))
        vals = _lowest(o, w, len(d["eigvals"]))
        assert np.all(np.abs(vals - d["eigvals"]) < EIGENVALUE_TOLERANCE), (f, vals - d["eigvals"])
    assert n == 51


def test_eigenvalues_negative_k_and_periodicity():
    for d in collect(6, 0, 6):
        k = d["symmetry_qs"]
        w = [d["weights"][t] for t in ("tl", "tr", "jl", "jr")]
        kxs = [-k[0]] + [k[0] + m * 6 for m in (-2, 1, 3)]
        for kx in kxs:
            vals = _lowest(OracleSector(*sector_of(d, kx=kx)), w, len(d["eigvals"]))
            assert np.all(np.abs(vals - d["eigvals"]) < EIGENVALUE_TOLERANCE)
