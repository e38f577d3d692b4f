"""The computations of the reference's examples (examples/tJ_spin_half_ladder/*.py) from their own JSON
configs (copied to tests/golden/examples), as the reference's test_examples.py runs them; plotting is left
out.  Results are checked against the oracle where it is cheap (eigenvalues), against exact identities
otherwise: spectra are non-negative (the continued fraction's imaginary part has one sign for eta > 0), the
hole correlations of h holes sum to h^2 over all distances and the S^z correlations of an S^z = 0 state to 0.
"""

import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EX = pathlib.Path(__file__).parent / "golden" / "examples"


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.fixture(autouse=True)
def in_tmp(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)


def test_all_ks_eigvals(torch, oracle_mod):
    """all_ks_eigvals.py: get_all_k_eigenpairs over kx <= L/2, both ky."""
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    config = configs.Eigenpair_Config.load(EX / "all_ks_eigvals.json")
    config.hamiltonian.symmetry_qs = None  # as the example's config is used for every k
    ham = config.hamiltonian
    w = [ham.weights.tl, ham.weights.tr, ham.weights.jl, ham.weights.jr]
    k_to_pairs = api.get_all_k_eigenpairs(config)
    assert sorted(k_to_pairs) == [(kx, ky) for kx in range(ham.num_rungs // 2 + 1) for ky in range(2)]
    for (kx, ky), (vals, vecs) in k_to_pairs.items():
        o = oracle_mod.OracleSector(ham.num_rungs, ham.num_holes, ham.num_down_spins, kx, ky)
        ev = np.linalg.eigvalsh(np.conj(o.toarray(w)))
        assert vecs.shape == (o.dim, len(vals))
        np.testing.assert_allclose(vals, ev[: len(vals)], rtol=0, atol=1e-8)
        for j in range(len(vals)):
            r = o.hv(w, vecs[:, j]) - vals[j] * vecs[:, j]
            assert np.linalg.norm(r) < 1e-6


def test_multiple_spectra(torch):
    """multiple_spectra.py: four excitation spectra of one config over all q (limited_qs=False)."""
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    config = configs.Full_Spectrum_Config.load(EX / "multiple_spectra.json")
    omega_mins = {"current_rung": 0, "current_leg": -0.5, "Szq": 0, "spectral_function_plus": -2}
    for name, omega_min in omega_mins.items():
        config.spectrum.name = name
        config.spectrum.omega_min = omega_min
        res = api.get_excitation_spectrum(config, limited_qs=False)
        s = np.asarray(res.spectrum)
        print(name, s.shape, float(s.max()), flush=True)
        assert np.all(np.isfinite(s)) and s.max() > 0
        assert s.min() > -1e-9 * s.max(), name


def test_rung_current_spectrum(torch):
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    config = configs.Full_Spectrum_Config.load(EX / "rung_current_spectrum.json")
    res = api.get_excitation_spectrum(config)
    s = np.asarray(res.spectrum)
    assert s.shape == (config.spectrum.omega_steps,) and np.all(np.isfinite(s)) and s.min() > -1e-9 * s.max()


@pytest.mark.parametrize("name", ["hole_position_correlations", "untested_Sz_position_correlations",
                                  "singlet-singlet_correlations"])
def test_position_correlation_examples(torch, name):
    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    config = configs.Full_Position_Correlation_Config.load(EX / f"{name}.json")
    res = api.get_position_correlations(config)
    ws, s = np.asarray(res.ws), np.asarray(res.spectrum)
    N = 2 * config.hamiltonian.num_rungs
    assert ws.shape == (N, 2) and s.shape == (N,) and np.all(np.isfinite(s))
    assert np.array_equal(ws, [[d // 2, d % 2] for d in range(N)])
    h = config.hamiltonian.num_holes
    if name.startswith("hole"):
        assert abs(s.sum() - h * h) < 1e-9  # sum over d and r of n_r n_{r+d} = (number of holes)^2
    elif name.startswith("untested_Sz"):
        assert abs(s.sum()) < 1e-9 and abs(s[0] - (N - h)) < 1e-9  # (2 S^z)^2 = 0; d = 0 counts the electrons
    else:
        assert s[0] > 0.1  # the published value is 0.1500 (singlet_correlations_scco.txt, first system)
