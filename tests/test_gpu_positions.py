"""Position correlations on the GPU: operators and reductions against the oracle, correlations against the
published values the reference's tests hold.

The oracle's operators are pinned by their second-quantised definitions (tests/test_oracle_bruteforce.py)
and its correlations by the published 2-hole values (tests/test_position_oracle.py).  Here the GPU must
give the oracle's operator columns bit for bit, the same correlations and expectation values to 1e-12, and
the reference's test_position_correlations.py comparisons (normalised to the first value, 1e-3) on every
published system: 9 + 7 hole-correlation systems (2x7 .. 2x12, 2 and 4 holes, with and without a fixed
hole) and 18 singlet-singlet systems (2x10, 4 holes).
"""

import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).parent / "golden" / "precalculated_position_correlations"

CASES = [
    ((4, 2, 3, 0, 0), 0, [(1, 1)]), ((4, 3, 2, 1, 0), 0, [(1, 0), (2, 1)]), ((5, 2, 4, 2, 1), 1, [(0, 1)]),
    ((4, 1, 3, 2, 1), 1, [(1, 1), (0, 0)]), ((4, 2, 3, 0, 0), 2, [(1, 0), (1, 1), (2, 1)]),
    ((5, 2, 4, 1, 1), 2, [(0, 1), (1, 0), (1, 1)]), ((5, 4, 3, 0, 0), 2, [(1, 1), (1, 1), (3, 0)]),
    ((6, 4, 4, 2, 0), 2, [(1, 0), (0, 1), (2, 1)]),
]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.mark.parametrize("case", CASES)
def test_operator_columns_bit_exact(oracle_mod, torch, case):
    from exactdiag_amd.sector import Sector

    sec, kind, shifts = case
    print("case", case, flush=True)
    o, g = oracle_mod.OracleSector(*sec), Sector(*sec)
    po, ro, vo = o.position_columns(kind, shifts)
    pg, rg, vg = g.position_columns(kind, shifts)
    assert np.array_equal(po, pg) and np.array_equal(ro, rg) and np.array_equal(vo, vg)
    # product and expectation from the same elements
    rng = np.random.default_rng(3)
    x = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
    M = np.zeros((o.dim, o.dim), np.complex128)
    for i in range(o.dim):
        for k in range(po[i], po[i + 1]):
            M[ro[k], i] += vo[k]
    y = g.position_apply(kind, shifts, torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert np.max(np.abs(y - M @ x)) <= 1e-13 * max(1.0, np.max(np.abs(M @ x)))
    e = g.position_expectation(kind, shifts, torch.as_tensor(x, device="cuda"))
    ref = np.vdot(x, M @ x)
    assert abs(e - ref) <= 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("sec,kind,fixed", [((6, 2, 5, 0, 0), 0, []), ((6, 4, 4, 1, 1), 0, [(1, 1)]),
                                            ((5, 1, 4, 2, 0), 1, []), ((6, 2, 5, 3, 1), 1, [(0, 1), (2, 0)])])
def test_correlations_match_oracle(oracle_mod, torch, sec, kind, fixed):
    from exactdiag_amd.sector import Sector

    o, g = oracle_mod.OracleSector(*sec), Sector(*sec)
    rng = np.random.default_rng(7)
    psi = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
    psi /= np.linalg.norm(psi)
    ref = o.position_correlations(kind, fixed, psi)
    got = g.position_correlations(kind, fixed, torch.as_tensor(psi, device="cuda"))
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def calc_dir(tmp_path_factory):
    """One calculations folder for the module, so that systems sharing a Hamiltonian share its eigenpairs."""
    return tmp_path_factory.mktemp("positions")


def _published(name, calc_dir, monkeypatch, normalize):
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    monkeypatch.chdir(calc_dir)
    worst = 0.0
    for d in json.loads((GOLD / name).read_text()):
        target = np.asarray(d.pop("probabilities"))
        config = configs.Full_Position_Correlation_Config(**d)
        print(name, d["hamiltonian"]["num_rungs"], d["hamiltonian"]["num_holes"], d["hamiltonian"]["weights"],
              [(s.leg, s.rung) for s in config.correlations.fixed_distances], flush=True)
        probabilities = api.get_position_correlations(config).spectrum.flatten()
        if normalize:
            probabilities = probabilities / probabilities[0]
        diff = np.max(np.abs(probabilities[: len(target)] - target))
        worst = max(worst, diff)
        assert diff < 1e-3, (d, probabilities[: len(target)], target)
    print("largest difference", worst, flush=True)


def test_hole_correlations_against_published_paulina(torch, calc_dir, monkeypatch):
    """test_hole_correlations_against_precalculated_paulina (test_position_correlations.py:16-22)."""
    _published("hole_correlations_paulina.txt", calc_dir, monkeypatch, normalize=True)


def test_hole_correlations_against_published_scco(torch, calc_dir, monkeypatch):
    """test_hole_correlations_against_precalculated_scco (test_position_correlations.py:25-33), 2 and 3 holes."""
    _published("hole_correlations_scco.txt", calc_dir, monkeypatch, normalize=True)


def test_singlet_correlations_against_published(torch, calc_dir, monkeypatch):
    """test_singlet_singlet_correlations_against_precalculated (test_position_correlations.py:48-56)."""
    _published("singlet_correlations_scco.txt", calc_dir, monkeypatch, normalize=False)


def test_projection_operators_select_the_right_states(torch, tmp_path, monkeypatch):
    """test_hole_matrix / test_sz_matrix of the reference (test_position_correlations.py:59-200), on a
    subset of the shifts: the stored operator is diagonal, its values count the reference sites whose
    positions hold holes (or the S^z sign product), and the folder follows the reference's naming."""
    from exactdiag_amd.tJ_spin_half_ladder import configs

    monkeypatch.chdir(tmp_path)
    for name, nh in (("hole_correlations", 2), ("Sz_correlations", 1), ("hole_correlations", 3)):
        config = configs.Limited_Position_Correlation_Config(
            hamiltonian={"num_rungs": 5, "num_holes": nh, "weights": {"tl": 0.45, "tr": 0.36, "jl": 0.15, "jr": 0.12},
                         "symmetry_qs": {"leg": 0, "rung": 0}, "num_threads": 8},
            correlations={"name": name, "fixed_distances": [] if nh < 3 else [{"leg": 1, "rung": 1}]})
        translator, bmap, sym = config.hamiltonian.get_translators()
        for free in ({"leg": 0, "rung": 1}, {"leg": 2, "rung": 0}, {"leg": 1, "rung": 1}):
            free_shift = configs.Position_Shift(**free)
            m = config.setup_excitation_operator(free_shift)[0]()
            shifts = [s.to_npint32() for s in config.correlations.fixed_distances] + [free_shift.to_npint32()]
            assert m.name == f"{name}_shifts{np.array(shifts).tolist()}"
            major, minor, vals = np.asarray(m.major_pointers), np.asarray(m.minor_inds), np.asarray(m.vals)
            empty = vals.tolist() == [0] and major[-1] == 1
            for i in range(m.shape[0]):
                st = translator.sparse_index_to_state(bmap.get_sparse(i))
                value = 0
                for r in range(len(st)):
                    pos = [r] + [sym.get_index_by_relative_shift(r, s) for s in shifts]
                    if name == "hole_correlations":
                        value += all(st[p] == 0 for p in pos)
                    elif all(st[p] != 0 for p in pos):
                        value += np.prod([-1 if st[p] == 1 else 1 for p in pos])
                n = major[i + 1] - major[i]
                if value == 0 or empty:
                    assert empty or n == 0, (name, free, i)
                    continue
                assert n == 1 and minor[major[i]] == i and vals[m.value_inds[major[i]]] == value, (name, free, i)
