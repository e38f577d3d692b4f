"""Eigenpairs and spectra (reference: general/lanczos_diagonalization.py).

Same functions, arguments, caching and return values as the reference:
``get_lowest_eigenpairs`` -> (eigenvalues float64[nev], eigenvectors
complex128[D, nev]); ``get_spectrum`` -> (ws, spectrum).  The dense path for
small blocks (D < 250 or D < 2 (nev + added)) and the averaging over
degenerate ground states follow the reference; the large-block eigensolver is
the GPU thick-restart Lanczos of ``krylov.py`` in place of ARPACK.
"""

from __future__ import annotations

import logging
import os
import pathlib
from os import makedirs

import numpy as np

from exactdiag_amd.general import krylov
from exactdiag_amd.general.cython_lanczos_diagonalization import lanczos_dynamical_response
from exactdiag_amd.utils import load_calculate_save

_logger = logging.getLogger(__name__)

EIGEN_TOL = 1e-10


def get_lowest_eigenpairs(config):
    """Load or calculate and save the eigenpairs (lanczos_diagonalization.py:21-34)."""
    eigvals, eigvecs = load_calculate_save(
        *config.get_eigenpair_paths(),
        load_fun=load_eigenpairs,
        calc_fun=lambda: calculate_eigenpairs(config),
        save_fun=save_eigenpairs,
    )
    _logger.info(f"Got lowest eigenpairs with k={config.hamiltonian.symmetry_qs.to_npint32()!s}, {eigvals=}")
    return eigvals, eigvecs


# Device copies of the ground-state vectors last loaded or computed in this process, keyed by their files
# (path, size, modification time): every q of an all-q run starts from the same ground state, which the
# reference reloads from disk for each q.  One entry; a rewritten file misses.
_GROUND_STATES: dict = {}


def _file_key(paths):
    key = []
    for p in paths:
        st = os.stat(p)
        key.append((str(p), st.st_size, st.st_mtime_ns))
    return tuple(key)


def ground_states_on_device(config):
    """(eigenvalues, (n_degenerate, D) device tensor of the ground states) for calculate_spectrum: the same
    eigenpairs get_lowest_eigenpairs returns (loaded or calculated and saved), kept on the device between
    calls for the same files."""
    import torch

    paths = config.get_eigenpair_paths()
    if all(os.path.isfile(p) for p in paths):
        hit = _GROUND_STATES.get(_file_key(paths))
        if hit is not None:
            return hit
    eigvals, eigvecs = get_lowest_eigenpairs(config)
    eigvals = np.asarray(eigvals)
    if len(eigvals) == 0:
        return eigvals, eigvecs
    n = 0
    while n < len(eigvals) and abs(eigvals[0] - eigvals[n]) <= 1e-6:
        n += 1
    vecs = torch.as_tensor(np.ascontiguousarray(np.asarray(eigvecs)[:, :n].T), device="cuda")
    _GROUND_STATES.clear()
    if all(os.path.isfile(p) for p in paths):
        _GROUND_STATES[_file_key(paths)] = (eigvals[:n], vecs)
    return eigvals[:n], vecs


def load_eigenpairs(vals_path, vecs_path):
    return np.load(vals_path, allow_pickle=False), np.load(vecs_path, allow_pickle=False)


def save_eigenpairs(eigenpairs, vals_path, vecs_path):
    makedirs(pathlib.Path(vals_path).parent, exist_ok=True)
    makedirs(pathlib.Path(vecs_path).parent, exist_ok=True)
    eigvals, eigvecs = eigenpairs
    np.save(vals_path, eigvals)
    np.save(vecs_path, eigvecs)


def calculate_eigenpairs_device(config, tol=EIGEN_TOL):
    """(eigenvalues, device tensor (nev, D) of eigenvectors) without leaving the GPU."""
    import torch

    hamiltonian = config.hamiltonian.setup()()
    num_rows = hamiltonian.shape[0]
    nev = config.eigenpair.num_eigenpairs
    added = config.eigenpair.num_recalculated_vecs_per_iteration
    if num_rows == 0:
        _logger.warning("hamiltonian has 0 rows, skipping eigen problem")
        return np.zeros(0), torch.zeros((0, 0), dtype=torch.complex128, device="cuda")
    if num_rows < 2 * (nev + added) or num_rows < 250:
        # lanczos_diagonalization.py:62-67: dense eigh of toarray()
        eigvals, eigvecs = np.linalg.eigh(hamiltonian.toarray())
        order = eigvals.argsort()[:nev]
        vecs = torch.as_tensor(np.ascontiguousarray(eigvecs[:, order].T), device="cuda")
        return eigvals[order], vecs
    (eigvals, vecs), _ = diagonalize_IRLM(nev, hamiltonian, num_added_vecs=added, tol=tol, device=True)
    return eigvals, vecs


def calculate_eigenpairs(config):
    """eigvecs[:, i] is the normalised eigenvector of eigvals[i] (lanczos_diagonalization.py:53-78)."""
    eigvals, vecs = calculate_eigenpairs_device(config)
    if len(eigvals) == 0:
        return [], []
    return np.real(eigvals), vecs.cpu().numpy().T.copy()


def diagonalize_IRLM(num_eigenpairs, operator, is_symmetric=False, num_added_vecs=5, initial_guess=None, tol=EIGEN_TOL,
                     device=False):
    """Lowest eigenpairs of a Hermitian GPU operator ((eigvals, eigvecs), iterations).

    Replaces lanczos_diagonalization.py:81-121 (ARPACK) by thick-restart
    Lanczos on the device.  ``eigvecs`` is (D, nev) numpy like the
    reference, or a (nev, D) device tensor when device=True.
    """
    D = operator.shape[1]

    def apply(v):
        return operator.apply(v)

    ncv = max(2 * num_eigenpairs + 24, 40, num_eigenpairs + num_added_vecs)
    # a real stored H (momenta kx = 0, L/2): the Krylov basis in float64, half the bytes per vector
    stored = getattr(operator, "stored", None)
    real = bool(stored is not None and stored.is_real and getattr(operator, "use_real_vectors", False))
    vals, vecs, it = krylov.lowest_eigenpairs(apply, D, num_eigenpairs, tol=tol, ncv=ncv, v0=initial_guess,
                                              real=real)
    if device:
        return (vals, vecs), it
    return (vals, vecs.cpu().numpy().T.copy()), it


def get_spectrum(config, suffix: str = ""):
    """Load or calculate and save one spectrum (lanczos_diagonalization.py:124-143)."""

    def calculate(**kwargs):
        eigvals, eigvecs = ground_states_on_device(config)
        get_operator, final_kwargs, get_excited_H = config.setup_excitation_operator()
        return calculate_spectrum(
            ws=config.spectrum.get_omegas(),
            broadening=config.spectrum.broadening,
            excited_state_hamiltonian=get_excited_H(),
            excitation_operator=get_operator(),
            eigvecs=eigvecs,
            eigvals=eigvals,
            num_lanczos_vecs=config.spectrum.num_lanczos_vecs,
            num_threads=config.spectrum.num_threads,
        )

    path = config.get_spectrum_path(suffix)
    return load_calculate_save(path, load_fun=load_spectrum, calc_fun=calculate, save_fun=save_spectrum)


def load_spectrum(path):
    with np.load(path, allow_pickle=False) as npz_file:
        ws, spectrum = (npz_file[k] for k in npz_file.files)
    return ws, spectrum


def save_spectrum(ws_spectrum, path):
    path = pathlib.Path(path)
    makedirs(path.parent, exist_ok=True)
    np.savez(path, *ws_spectrum)


def calculate_spectrum(ws, broadening, num_lanczos_vecs, excited_state_hamiltonian, excitation_operator, eigvecs,
                       eigvals, num_threads=None, precision=0):
    """Average over degenerate ground states (lanczos_diagonalization.py:158-175).

    ``eigvecs`` may be (D, nev) numpy as in the reference or a (nev, D)
    device tensor.
    """
    import torch

    gs_energy = eigvals[0]
    num_degenerate_states = 0
    spectrum = np.zeros(len(ws))
    zs = ws + gs_energy + 1j * broadening
    for j in range(len(eigvals)):
        if abs(gs_energy - eigvals[j]) > 1e-6:
            break
        num_degenerate_states += 1
        gs = eigvecs[j] if isinstance(eigvecs, torch.Tensor) else eigvecs[:, j]
        spectrum += lanczos_dynamical_response(gs, excitation_operator, excited_state_hamiltonian, zs,
                                               num_lanczos_vecs, num_threads, precision=precision)
    spectrum /= num_degenerate_states
    return ws, spectrum


class Dummy_Operator:
    def dot(self, vector):
        return vector
