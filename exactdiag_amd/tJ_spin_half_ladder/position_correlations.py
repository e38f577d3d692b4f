"""Position correlations of ground states on the GPU
(reference: tJ_spin_half_ladder/position_correlations.pyx).

``get_hole_spin_projection_correlations`` and ``get_singlet_singlet_correlations``
keep the reference's signatures, caching (``config.get_spectrum_path()``) and
return values: (varied_shifts int32[N, 2], correlations float64[N]) averaged
over the degenerate ground states (|E_j - E_0| <= 1e-6).

Hole and S^z correlations are one reduction over the basis per state
(``ed_position_correlations``: all N node distances in one pass, |psi|^2 times
the value at the positions of every reference site).  Singlet-singlet
correlations are <psi|O(shift)|psi> of the reference's operator for every
relative position of the two singlets (``ed_position_expectation``: the
operator's columns are generated and contracted with psi in one kernel, no
matrix is stored).
"""

from __future__ import annotations

import numpy as np

from exactdiag_amd import _lib
from exactdiag_amd.general.lanczos_diagonalization import get_lowest_eigenpairs, load_spectrum, save_spectrum
from exactdiag_amd.tJ_spin_half_ladder import configs
from exactdiag_amd.utils import load_calculate_save

DEGENERACY_TOL = 1e-6  # position_correlations.pyx:40, :64


def _device(vec):
    import torch

    return torch.as_tensor(np.ascontiguousarray(vec, np.complex128), device="cuda")


def _unit_shifts(num_nodes):
    """get_unit_shifts_from_index of every node index (symmetry_ladder.h:103-108): (index // 2, index % 2)."""
    return np.array([[d // 2, d % 2] for d in range(num_nodes)], dtype=np.int32)


def _degenerate(eigvals):
    """Indices of the ground states averaged over: the leading run within 1e-6 of the lowest."""
    out = []
    for j in range(len(eigvals)):
        if abs(eigvals[0] - eigvals[j]) > DEGENERACY_TOL:
            break
        out.append(j)
    return out


def get_hole_spin_projection_correlations(config: "configs.Full_Position_Correlation_Config"):
    """Hole or S^z correlations (position_correlations.pyx:29-50), loaded from or saved to the spectrum path."""
    from exactdiag_amd.tJ_spin_half_ladder.matrix_setup import get_sector

    def calculate():
        sector = get_sector(config.hamiltonian)
        eigvals, eigvecs = get_lowest_eigenpairs(config)
        fixed = [shift.to_npint32() for shift in config.correlations.fixed_distances]
        spin_or_hole = config.correlations.name == configs.Position_Correlation_Name.SZ_CORRELATIONS
        states = _degenerate(eigvals)
        spectrum = None
        for j in states:
            varied_shifts, corr = calculate_hole_spin_projection_correlations(sector, eigvecs[:, j], spin_or_hole,
                                                                              fixed)
            spectrum = corr if spectrum is None else spectrum + corr
        return varied_shifts, spectrum / len(states)

    return load_calculate_save(config.get_spectrum_path(), load_fun=load_spectrum, calc_fun=calculate,
                               save_fun=save_spectrum)


def get_singlet_singlet_correlations(config: "configs.Full_Position_Correlation_Config"):
    """Singlet-singlet correlations (position_correlations.pyx:52-74)."""
    from exactdiag_amd.tJ_spin_half_ladder.matrix_setup import get_sector

    def calculate():
        sector = get_sector(config.hamiltonian)
        eigvals, eigvecs = get_lowest_eigenpairs(config)
        fixed = [shift.to_npint32() for shift in config.correlations.fixed_distances]
        states = _degenerate(eigvals)
        spectrum = None
        for j in states:
            varied_shifts, corr = calculate_singlet_singlet_correlations(sector, eigvecs[:, j], fixed)
            spectrum = corr if spectrum is None else spectrum + corr
        return varied_shifts, spectrum / len(states)

    return load_calculate_save(config.get_spectrum_path(), load_fun=load_spectrum, calc_fun=calculate,
                               save_fun=save_spectrum)


def calculate_hole_spin_projection_correlations(sector, eigvec, spin_or_hole: bool, fixed_distances):
    """(varied_shifts, correlations) of one state (position_correlations.pyx:84-127); spin_or_hole True = S^z.

    eigvec: numpy or device vector of the sector.  Correlation d counts, for every reference site r, the
    positions r, r shifted by the unit shifts of node index d, and r shifted by each fixed distance.
    """
    kind = _lib.POS_SZ if spin_or_hole else _lib.POS_HOLES
    corr = sector.position_correlations(kind, np.asarray(fixed_distances, np.int32).reshape(-1, 2), _device(eigvec))
    return _unit_shifts(sector.N), corr


def calculate_singlet_singlet_correlations(sector, eigvec, fixed_distances):
    """(varied_shifts, Re <psi|O(shift)|psi>) for every node index as the free shift
    (position_correlations.pyx:129-142)."""
    psi = _device(eigvec)
    shifts = _unit_shifts(sector.N)
    fixed = np.asarray(fixed_distances, np.int32).reshape(-1, 2)
    corrs = np.empty(sector.N)
    for i, shift in enumerate(shifts):
        corrs[i] = sector.position_expectation(_lib.POS_SINGLET, np.vstack([fixed, shift[None, :]]), psi).real
    return shifts, corrs
