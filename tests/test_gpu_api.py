"""The drop-in Python API on the GPU, tested the way the reference tests itself.

Mirrors tests/tJ_spin_half_ladder/test_eigvals.py of the reference (the 51
eigenvalue fixtures through get_lowest_eigenpairs at 1e-8, k <-> -k,
k periodicity) and drives the all-q spectrum entry point
(tJ_spin_half_ladder/api.py:83-132) against the exact Lehmann sum built
from the oracle's dense matrices.  Results are cached under ./calculations
as in the reference; every test runs in its own temporary directory.
"""

import glob
import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).parent / "golden" / "precalculated_basis_size_eigenvalues"
EIGENVALUE_TOLERANCE = 1e-8


@pytest.fixture(autouse=True)
def in_tmp(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _load(path):
    with open(path, encoding="utf-8") as fp:
        return json.load(fp)


def _he_config(d, leg=None):
    """Precalculated.to_he_config of the reference's test_basis_size.py:158-179."""
    from exactdiag_amd.tJ_spin_half_ladder import configs

    num_spins = 2 * d["num_rungs"] - d["num_holes"]
    weights = d.get("weights") or {"tl": 1, "tr": 1, "jl": 1, "jr": 1}
    k = d["symmetry_qs"]
    ham = configs.Hamiltonian_Config(num_rungs=d["num_rungs"], num_holes=d["num_holes"], weights=weights,
                                     symmetry_qs={"leg": k[0] if leg is None else leg, "rung": k[1]},
                                     total_spin_projection=num_spins - 2 * d["num_down_spins"])
    return configs.Eigenpair_Config(hamiltonian=ham, eigenpair={"num_eigenpairs": len(d["eigvals"])})


def test_against_precalculated(torch):
    from exactdiag_amd.general.lanczos_diagonalization import get_lowest_eigenpairs

    n = 0
    for f in sorted(glob.glob(str(GOLD / "*.txt"))):
        d = _load(f)
        if "eigvals" not in d:
            continue
        n += 1
        print(pathlib.Path(f).name, flush=True)
        vals, vecs = get_lowest_eigenpairs(_he_config(d))
        assert np.all(np.abs(vals - d["eigvals"]) < EIGENVALUE_TOLERANCE), (f, vals, d["eigvals"])
        assert vecs.shape[1] == len(d["eigvals"])
    assert n == 51


def _collect(num_rungs, num_holes, num_down_spins):
    out = []
    for kx in range(num_rungs // 2 + 1):
        for ky in range(2):
            out.append(_load(GOLD / f"rungs{num_rungs}_holes{num_holes}_k[{kx},{ky}]_down{num_down_spins}.txt"))
    return out


def test_negative_k_leg(torch):
    from exactdiag_amd.general.lanczos_diagonalization import get_lowest_eigenpairs

    coll = _collect(6, 0, 6)
    assert len(coll) == 8
    for d in coll:
        k = d["symmetry_qs"]
        if k[0] == 0:
            continue
        vals, _ = get_lowest_eigenpairs(_he_config(d, leg=-k[0]))
        assert np.all(np.abs(vals - d["eigvals"]) < EIGENVALUE_TOLERANCE)


def test_k_leg_periodicity(torch):
    from exactdiag_amd.general.lanczos_diagonalization import get_lowest_eigenpairs

    for d in _collect(6, 0, 6):
        k = d["symmetry_qs"]
        if tuple(k) == (0, 0):
            continue
        for m in (-2, 1, 3):
            vals, _ = get_lowest_eigenpairs(_he_config(d, leg=k[0] + m * 6))
            assert np.all(np.abs(vals - d["eigvals"]) < EIGENVALUE_TOLERANCE)


def test_stored_operator_through_the_api(torch, monkeypatch, oracle_mod):
    """With the stored form forced on, the API's eigenpairs are those of the oracle's dense matrix."""
    from exactdiag_amd.general import sparse_matrices
    from exactdiag_amd.general.lanczos_diagonalization import calculate_eigenpairs
    from exactdiag_amd.tJ_spin_half_ladder import configs

    monkeypatch.setattr(sparse_matrices, "STORED_MIN_DIM", 0)
    w = {"tl": 0.45, "tr": 0.36, "jl": 0.15, "jr": 0.12}
    ham = configs.Hamiltonian_Config(num_rungs=6, num_holes=2, weights=w, symmetry_qs={"leg": 2, "rung": 1})
    cfg = configs.Eigenpair_Config(hamiltonian=ham, eigenpair={"num_eigenpairs": 4})
    op = ham.setup()()
    assert op.stored is not None and op.stored.nnz > 0
    vals, vecs = calculate_eigenpairs(cfg)
    exact = np.linalg.eigvalsh(np.conj(oracle_mod.OracleSector(6, 2, 5, 2, 1).toarray([0.45, 0.36, 0.15, 0.12])))
    assert np.max(np.abs(vals - exact[:4])) < 1e-10


def _op_matrix(ptr, rows, vals, n_rows, n_cols):
    M = np.zeros((n_rows, n_cols), np.complex128)
    for i in range(n_cols):
        for k in range(ptr[i], ptr[i + 1]):
            M[rows[k], i] += vals[k]
    return M


def test_all_q_szq_spectra_equal_lehmann_sums(torch, oracle_mod):
    """api.get_excitation_spectrum for S^z(q) at every q of a 2x4 ladder vs the exact spectral sums."""
    from exactdiag_amd.allq import final_momentum
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    L, h, nd = 4, 2, 3
    w = {"tl": 1.0, "tr": 1.0, "jl": 0.4, "jr": 0.4}
    wl = [1.0, 1.0, 0.4, 0.4]
    eta, ws = 0.05, np.linspace(0, 4, 401)
    cfg = configs.Full_Spectrum_Config(
        hamiltonian={"num_rungs": L, "num_holes": h, "weights": w, "symmetry_qs": {"leg": 0, "rung": 0}},
        spectrum={"name": "Szq", "operator_symmetry_qs": {"leg": 0, "rung": 0}, "broadening": eta,
                  "omega_min": 0.0, "omega_max": 4.0, "omega_steps": 401, "num_lanczos_vecs": 120},
        eigenpair={"num_eigenpairs": 3})
    res = api.get_excitation_spectrum(cfg, limited_qs=True)
    assert res.spectrum.shape == (L, 2, 401)
    oi = oracle_mod.OracleSector(L, h, nd, 0, 0)
    Hi = np.conj(oi.toarray(wl))
    ev, evec = np.linalg.eigh(Hi)
    gs = [evec[:, i] for i in range(len(ev)) if ev[i] - ev[0] < 1e-6]
    for qx in range(L):
        for qy in range(2):
            q_used = (L - qx if qx > L // 2 else qx, qy)  # limited_qs, api.py:126-131
            kx, ky = final_momentum(L, 0, 0, *q_used)
            of = oracle_mod.OracleSector(L, h, nd, kx, ky)
            Hf = np.conj(of.toarray(wl))
            fv, fvec = np.linalg.eigh(Hf)
            O = _op_matrix(*oi.op_columns(of, 0, *q_used), of.dim, oi.dim)
            exact = np.zeros(len(ws))
            for g in gs:  # average over degenerate ground states (lanczos_diagonalization.py:158-175)
                wts = np.abs(fvec.conj().T @ (O @ g)) ** 2
                exact += np.array([np.sum(wts * eta / np.pi / ((x + ev[0] - fv) ** 2 + eta**2)) for x in ws])
            exact /= len(gs)
            got = res.spectrum[qx, qy]
            assert np.max(np.abs(got - exact)) < 1e-6 * max(1.0, np.max(exact)), (qx, qy)


def test_translators_write_and_check_the_basis_file(torch, oracle_mod):
    """get_translators() keeps <calc folder>/Sz_local-momentum_map.dat as get_basis_map does
    (general/basis_indexing.h:245-253): written from the GPU basis, then read back and checked."""
    from exactdiag_amd.general import basis_indexing as bi
    from exactdiag_amd.tJ_spin_half_ladder import configs

    ham = configs.Hamiltonian_Config(num_rungs=6, num_holes=2, weights={"tl": 1, "tr": 1, "jl": 0.4, "jr": 0.4},
                                     symmetry_qs={"leg": 1, "rung": 1})
    translator, bmap, sym = ham.get_translators()
    path = ham.get_calc_folder() / bi.BASIS_MAP_NAME
    sparse, orbit = bi.load_basis_map(path)
    so, co = oracle_mod.OracleSector(6, 2, 5, 1, 1).basis()
    assert np.array_equal(sparse, so) and np.array_equal(orbit, co)
    assert bmap.get_num_states() == len(so)
    ham.get_translators()  # an existing, matching file is accepted
    bad = sparse.copy()
    bad[3], bad[4] = bad[4], bad[3]
    bi.save_basis_map(path, bad, orbit)
    with pytest.raises(ValueError):
        ham.get_translators()
