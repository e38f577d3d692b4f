"""The oracle's position correlations against the published values the reference's tests hold.

tests/golden/precalculated_position_correlations/hole_correlations_paulina.txt
(copied from the reference's tests/tJ_spin_half_ladder/precalculated_position_correlations)
lists 2-hole correlations of 2xL ladders; the reference's test compares them, normalised to the first
value, at 1e-3 (tests/tJ_spin_half_ladder/test_position_correlations.py:16-45).  Here the oracle's
restatement of calculate_hole_spin_projection_correlations is fed the ground states of the oracle's
Hamiltonian (scipy's Lanczos on its CSR dot), for three of the systems small enough for the CPU (L <= 8);
the GPU path is checked on all of them in tests/test_gpu_positions.py.
"""

import json
import pathlib

import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, eigsh

GOLD = pathlib.Path(__file__).parent / "golden" / "precalculated_position_correlations"


def _cases(name, max_rungs, pick=None):
    small = [d for d in json.loads((GOLD / name).read_text()) if d["hamiltonian"]["num_rungs"] <= max_rungs]
    return small if pick is None else [small[i] for i in pick]


def _ground_states(o, w, nev):
    op = LinearOperator((o.dim, o.dim), matvec=lambda x: o.hv(w, x), dtype=np.complex128)
    vals, vecs = eigsh(op, k=nev, which="SA", tol=1e-12, v0=np.ones(o.dim, np.complex128))
    order = np.argsort(vals)
    vals, vecs = vals[order], vecs[:, order]
    keep = [j for j in range(nev) if abs(vals[j] - vals[0]) <= 1e-6]
    return [vecs[:, j] for j in keep]


@pytest.mark.parametrize("case", _cases("hole_correlations_paulina.txt", 8, pick=[0, 1, 4]),
                         ids=lambda d: f"L{d['hamiltonian']['num_rungs']}")
def test_oracle_hole_correlations_against_published(oracle_mod, case):
    ham = case["hamiltonian"]
    L, h = ham["num_rungs"], ham["num_holes"]
    nd = (2 * L - h) // 2
    k = ham["symmetry_qs"]
    o = oracle_mod.OracleSector(L, h, nd, k["leg"], k["rung"])
    w = [ham["weights"][t] for t in ("tl", "tr", "jl", "jr")]
    fixed = [(s["leg"], s["rung"]) for s in case["correlations"]["fixed_distances"]]
    states = _ground_states(o, w, case["eigenpair"]["num_eigenpairs"])
    corr = sum(o.position_correlations(0, fixed, gs) for gs in states) / len(states)
    target = np.asarray(case["probabilities"])
    got = corr / corr[0]
    assert np.all(np.abs(got[: len(target)] - target) < 1e-3), (got[: len(target)], target)
