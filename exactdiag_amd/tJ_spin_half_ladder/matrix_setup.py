"""Wiring of sectors, Hamiltonians and excitation operators
(reference: tJ_spin_half_ladder/matrix_setup.pyx).

``setup_hamiltonian`` / ``setup_excitation_operator`` / ``py_get_ladder_translators``
keep the reference's signatures and return values; the objects behind them
are GPU sectors and matrix-free operators instead of CSR files.
"""

from __future__ import annotations

import copy
import functools
import logging
import pathlib

import numpy as np

from exactdiag_amd import _lib
from exactdiag_amd.general.sparse_matrices import Excitation_Operator, Hamiltonian_Operator
from exactdiag_amd.sector import Sector
from exactdiag_amd.tJ_spin_half_ladder import configs

_logger = logging.getLogger(__name__)

_SECTORS: "dict[tuple, Sector]" = {}


def get_sector(config: "configs.Hamiltonian_Config") -> Sector:
    """The GPU sector of a Hamiltonian config (built once per process and key)."""
    key = config.sector_key()
    s = _SECTORS.get(key)
    if s is None:
        s = Sector(*key)
        _SECTORS[key] = s
    return s


def clear_sector_cache():
    for s in _SECTORS.values():
        s.close()
    _SECTORS.clear()


class Symmetries:
    """Inspection view of the ladder symmetry group (symmetry_ladder.h:55-111)."""

    def __init__(self, num_rungs, symmetry_qs):
        self.num_rungs = num_rungs
        self.symmetry_qs = list(symmetry_qs)

    def get_basis_length(self):
        return 2 * self.num_rungs

    def get_shift_periodicities(self):
        return [self.num_rungs, 2]

    def get_num_shifts(self):
        return 2

    def get_index_by_relative_shift(self, initial_index, relative_shift):
        N = 2 * self.num_rungs
        i = initial_index % N
        leg = i // 2 + int(relative_shift[0])
        s1 = int(relative_shift[1])
        rung = (i + int(np.fmod(s1, 2))) % 2
        return (2 * leg + rung) % N

    def get_unit_shifts_from_index(self, index):
        """(leg, rung) shifts of a node index (symmetry_ladder.h:103-108)."""
        index = int(index) % (2 * self.num_rungs)
        return np.array([index // 2, index % 2], dtype=np.int32)

    def get_index_by_relative_index(self, initial_index, relative_index):
        """initial_index moved by the shifts that take node 0 to relative_index (symmetry_generator.h:36-45)."""
        return self.get_index_by_relative_shift(initial_index, self.get_unit_shifts_from_index(relative_index))


class Translator:
    """Sparse index <-> state and lowest-state lookups (State_Index_Amplitude_Translator)."""

    def __init__(self, sector: Sector, config):
        self.sector = sector
        self.num_nodes = sector.N
        self.num_holes = sector.h
        self.num_down_spins = sector.nd
        self.comb = configs.set_combinatorics(sector.N)
        self.num_spin_states = int(self.comb[sector.nd, sector.N - sector.h])

    def get_num_minor_sparse_states(self):
        return self.num_spin_states

    def sparse_index_to_state(self, sparse_index):
        """symmetry_ladder.h:314-344."""
        N, comb = self.num_nodes, self.comb
        rh, rs = self.num_holes, self.num_down_spins
        hi, si = divmod(int(sparse_index), self.num_spin_states)
        state = np.empty(N, np.int16)
        for j in range(N - 1, -1, -1):
            if rh > 0 and hi >= comb[rh, j]:
                state[j] = 0
                hi -= comb[rh, j]
                rh -= 1
                continue
            if rs > 0 and si >= comb[rs, j - rh]:
                state[j] = 1
                si -= comb[rs, j - rh]
                rs -= 1
                continue
            state[j] = 2
        return state

    def state_to_sparse_index(self, state):
        """symmetry_ladder.h:346-361."""
        hi = si = 0
        ah, ad = int(state[0] == 0), int(state[0] == 1)
        for i in range(1, self.num_nodes):
            if state[i] == 0:
                ah += 1
                hi += self.comb[ah, i]
            elif state[i] == 1:
                ad += 1
                si += self.comb[ad, i - ah]
        return int(si + hi * self.num_spin_states)

    def get_lowest_sparse_amplitude_from_state(self, state):
        """(sparse index of the lowest image or -1, its amplitude), general/symmetry.h:240-269."""
        rep, amp = self.sector.lowest([self.state_to_sparse_index(state)])
        return int(rep[0]), complex(amp[0])


class Basis_Map:
    """Dense <-> sparse index map of a sector (Basis_Index_Map, basis_indexing.h:31-185)."""

    def __init__(self, sector: Sector):
        self.sector = sector
        self._host = None

    def _arrays(self):
        if self._host is None:
            self._host = self.sector.basis()
        return self._host

    def get_num_states(self):
        return self.sector.dim

    def get_sparse(self, dense_index):
        return int(self._arrays()[0][dense_index])

    def get_dense(self, sparse_index):
        return int(self.sector.lookup([sparse_index])[0])

    def get_num_sparse_in_dense(self, dense_index):
        return int(self._arrays()[1][dense_index])

    def get_num_sparse_in_dense_from_sparse(self, sparse_index):
        d = self.get_dense(sparse_index)
        return self.get_num_sparse_in_dense(d) if d > -1 else 0


def py_get_ladder_translators(config: "configs.Hamiltonian_Config", basis_file=True):
    """(translator, basis map, symmetries), matrix_setup.pyx:377-406.

    As the reference's get_basis_map (general/basis_indexing.h:245-253) does with
    <calc folder>/Sz_local-momentum_map.dat: an existing file is read and must equal the
    GPU-built basis (ValueError otherwise); a missing one is written.
    """
    from exactdiag_amd.general import basis_indexing

    s = get_sector(config)
    bmap = Basis_Map(s)
    if basis_file:
        path = config.get_calc_folder() / basis_indexing.BASIS_MAP_NAME
        sparse, orbit = bmap._arrays()
        if path.exists():
            basis_indexing.check_basis_map(path, sparse, orbit)
        else:
            basis_indexing.save_basis_map(path, sparse, orbit)
    return Translator(s, config), bmap, Symmetries(config.num_rungs, config.symmetry_qs.to_npint32())


def _weights_dict(config):
    return {n: getattr(config.weights, n, 0) for n in ("tl", "tr", "jl", "jr")}


def setup_hamiltonian(config: "configs.Hamiltonian_Config"):
    """Callable returning the Hamiltonian (matrix_setup.pyx:41-129)."""
    weights = _weights_dict(config)

    def get():
        return Hamiltonian_Operator(get_sector(config), weights)

    return get


OPERATOR_IDS = {
    configs.Spectrum_Name.SZQ: _lib.OP_SZQ,
    configs.Spectrum_Name.DENSITY_Q: _lib.OP_NQ,
    configs.Spectrum_Name.CURRENT_RUNG: _lib.OP_CURRENT_RUNG,
    configs.Spectrum_Name.CURRENT_LEG: _lib.OP_CURRENT_LEG,
}


def excitation_setup(initial_config):
    """Final-sector config, operator id and q of an excitation (matrix_setup.pyx:195-330)."""
    ham = initial_config.hamiltonian
    num_rungs = ham.num_rungs
    kx, ky = (int(v) for v in ham.symmetry_qs.to_npint32())
    qs = initial_config.spectrum.operator_symmetry_qs
    qx, qy = (int(v) for v in qs.to_npint32()) if qs is not None else (0, 0)
    final = copy.deepcopy(initial_config)
    name = configs.Spectrum_Name(initial_config.spectrum.name)
    if name in (configs.Spectrum_Name.SZQ, configs.Spectrum_Name.DENSITY_Q, configs.Spectrum_Name.CURRENT_LEG):
        op = OPERATOR_IDS[name]
    elif name == configs.Spectrum_Name.CURRENT_RUNG:
        qy += 1  # matrix_setup.pyx:251
        op = _lib.OP_CURRENT_RUNG
    elif name == configs.Spectrum_Name.SPECTRAL_FUNCTION_PLUS:
        final.hamiltonian.num_holes -= 1
        if ham.total_spin_projection > 0:
            final.hamiltonian.total_spin_projection -= 1
            op = _lib.OP_APLUS_DOWN
        else:
            final.hamiltonian.total_spin_projection += 1
            op = _lib.OP_APLUS_UP
    elif name == configs.Spectrum_Name.SPECTRAL_FUNCTION_MINUS:
        final.hamiltonian.num_holes += 1
        if ham.total_spin_projection <= 0:
            final.hamiltonian.total_spin_projection += 1
            op = _lib.OP_AMINUS_DOWN
        else:
            final.hamiltonian.total_spin_projection -= 1
            op = _lib.OP_AMINUS_UP
    else:
        raise ValueError(f"{name} not supported")
    half = (num_rungs - 1) // 2
    final.hamiltonian.symmetry_qs.leg = (kx + qx + half) % num_rungs - half
    final.hamiltonian.symmetry_qs.rung = (ky + qy) % 2
    return final, op, qx, qy


def setup_excitation_operator(initial_config):
    """(get_operator, final Hamiltonian config, get_excited_H), matrix_setup.pyx:195-374."""
    final, op, qx, qy = excitation_setup(initial_config)
    get_excited_H = final.hamiltonian.setup()
    name = str(initial_config.spectrum.name)

    def get_operator():
        return Excitation_Operator(get_sector(initial_config.hamiltonian), get_sector(final.hamiltonian), op, qx,
                                   qy, name)

    _logger.debug(f"Finished setting up operator {name} with {initial_config}.")
    return get_operator, final.hamiltonian, get_excited_H


# keep functools imported for callers that build partials like the reference
partial = functools.partial


# ---------------------------------------------------------------------------
# The reference's matrix folders under the calc folder
# (general/sparse_matrices.py:246-274, matrix_setup.pyx:41-129 and :195-374).

TERM_IDS = {"tl": 0, "tr": 1, "jl": 2, "jr": 3}


def term_max_values_per_column(config, name):
    """max_num_values_per_column of a term, as the reference sets it (matrix_setup.pyx:80-108)."""
    h, L = config.num_holes, config.num_rungs
    spins = 2 * L - h
    return {"tl": 2 * h, "tr": min(h, L), "jl": min(spins, L) + 1, "jr": min(spins, L) + 1}[name]


def hamiltonian_term_signature(config, name):
    """The Signature setup_hamiltonian gives term `name` (matrix_setup.pyx:54-71, :176-190)."""
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Signature

    D = get_sector(config).dim
    strings = ["hermitian", "sorted"] + (["triangle_only"] if config.triangle_only else [])
    return Signature(matrix_name=name, shape=(D, D), symmetry_strings=tuple(strings),
                     max_values_per_column=term_max_values_per_column(config, name), major=config.major,
                     initial_system_info=config.get_symmetry_info(), final_system_info=None,
                     folder_name=str(config.get_calc_folder() / name), **FILE_NAMES)


def write_hamiltonian_term(config, name):
    """Build term `name` on the GPU in the config's layout, write its folder, return it opened."""
    from exactdiag_amd.csr import StoredCSR
    from exactdiag_amd.general.sparse_matrices import Sparse_Matrix

    if name not in TERM_IDS:
        raise ValueError(f"{name} is not a Hamiltonian term")
    m = StoredCSR.from_sector(get_sector(config), TERM_IDS[name], triangle_only=config.triangle_only,
                              row_major=config.major == "row")
    limit = term_max_values_per_column(config, name)
    if m.max_count > limit:
        _logger.warning(f"{name}: a column has {m.max_count} elements, more than the {limit} the reference "
                        f"allows for it (its build raises IndexError for this sector); written anyway.")
    try:
        return Sparse_Matrix.from_device(m, hamiltonian_term_signature(config, name))
    finally:
        m.close()


def get_hamiltonian_matrices(config, overwrite=False):
    """setup_hamiltonian(config)() of the reference, on disk (general/sparse_matrices.py:246-274).

    Every term with a nonzero weight is opened from <calc folder>/<term> when saved there, otherwise built
    on the GPU and saved; returns their weighted sum (``Added_Sparse_Matrices``, applied on the GPU).
    """
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Added_Sparse_Matrices, Sparse_Matrix

    mats, weights = [], []
    for name, w in _weights_dict(config).items():
        if w == 0:
            continue
        folder = config.get_calc_folder() / name
        if overwrite or not (folder / FILE_NAMES["info_name"]).exists():
            mats.append(write_hamiltonian_term(config, name))
        else:
            mats.append(Sparse_Matrix(hamiltonian_term_signature(config, name), num_threads=config.num_threads))
        weights.append(w)
    if not mats:
        raise ValueError("all Hamiltonian weights are zero")
    return Added_Sparse_Matrices(mats, weights)


def _operator_max_values_per_column(initial_config, op):
    """max_num_values_per_column of an excitation operator (matrix_setup.pyx:240-323)."""
    ham = initial_config.hamiltonian
    h, nodes = ham.num_holes, 2 * ham.num_rungs
    if op in (_lib.OP_SZQ, _lib.OP_NQ):
        return 1
    if op == _lib.OP_CURRENT_RUNG:
        return h
    if op == _lib.OP_CURRENT_LEG:
        return 2 * h
    if op in (_lib.OP_APLUS_DOWN, _lib.OP_APLUS_UP):
        return h
    if op == _lib.OP_AMINUS_DOWN:
        return ham.num_down_spins
    return nodes - h - ham.num_down_spins


def excitation_operator_signature(initial_config):
    """The Signature setup_excitation_operator gives the operator (matrix_setup.pyx:350-364)."""
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Signature

    final, op, _, _ = excitation_setup(initial_config)
    qs = initial_config.spectrum.operator_symmetry_qs
    q_string = f"_{qs.to_npint32()}" if qs is not None else ""
    full_name = f"{initial_config.spectrum.name}{q_string}"
    shape = (get_sector(final.hamiltonian).dim, get_sector(initial_config.hamiltonian).dim)
    return Signature(matrix_name=full_name, shape=shape, symmetry_strings=("sorted",),
                     max_values_per_column=_operator_max_values_per_column(initial_config, op), major="column",
                     initial_system_info=initial_config.hamiltonian.get_symmetry_info(),
                     final_system_info=final.hamiltonian.get_symmetry_info(),
                     folder_name=str(initial_config.hamiltonian.get_calc_folder() / full_name), **FILE_NAMES)


def get_excitation_operator_matrix(initial_config, overwrite=False):
    """get_operator() of the reference's setup_excitation_operator, on disk: opened when saved, else built
    on the GPU and saved (column-major, final x initial)."""
    from exactdiag_amd.csr import StoredCSR
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Sparse_Matrix

    sig = excitation_operator_signature(initial_config)
    folder = initial_config.hamiltonian.get_calc_folder() / sig.matrix_name
    if not overwrite and (folder / FILE_NAMES["info_name"]).exists():
        return Sparse_Matrix(sig)
    final, op, qx, qy = excitation_setup(initial_config)
    m = StoredCSR.from_operator(get_sector(initial_config.hamiltonian), get_sector(final.hamiltonian), op, qx, qy)
    try:
        return Sparse_Matrix.from_device(m, sig)
    finally:
        m.close()


# ---------------------------------------------------------------------------
# Position-correlation operators (matrix_setup.pyx:409-520).

POSITION_KINDS = {
    configs.Position_Correlation_Name.HOLE_CORRELATIONS: _lib.POS_HOLES,
    configs.Position_Correlation_Name.SZ_CORRELATIONS: _lib.POS_SZ,
    configs.Position_Correlation_Name.SINGLET_SINGLET: _lib.POS_SINGLET,
}


def position_operator_signature(config, free_shift):
    """The Signature get_position_correlation_operator gives the operator (matrix_setup.pyx:492-508)."""
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Signature

    shifts = np.array([s.to_npint32() for s in config.correlations.fixed_distances] + [free_shift.to_npint32()],
                      dtype=np.int32)
    full_name = f"{config.correlations.name}_shifts{shifts.tolist()}"
    name = configs.Position_Correlation_Name(config.correlations.name)
    n_spins = 2 * config.hamiltonian.num_rungs - config.hamiltonian.num_holes
    D = get_sector(config.hamiltonian).dim
    return Signature(matrix_name=full_name, shape=(D, D), symmetry_strings=("sorted",),
                     max_values_per_column=4 * n_spins if name == configs.Position_Correlation_Name.SINGLET_SINGLET
                     else 1,
                     major="column", initial_system_info=config.hamiltonian.get_symmetry_info(),
                     final_system_info=None, folder_name=str(config.hamiltonian.get_calc_folder() / full_name),
                     **FILE_NAMES), shifts


def get_position_correlation_operator(config, free_shift):
    """(get_operator, Hamiltonian config, get_excited_H) as the reference returns them.

    get_operator() opens the operator's folder <calc folder>/<name>_shifts[...] when it is saved there,
    otherwise builds the operator on the GPU and saves it, as the reference does; the returned matrix is
    applied on the GPU.  (The correlations themselves do not go through the folder:
    position_correlations.py contracts the operator with the state directly.)
    """
    from exactdiag_amd.csr import StoredCSR
    from exactdiag_amd.general.sparse_matrices import FILE_NAMES, Sparse_Matrix

    name = configs.Position_Correlation_Name(config.correlations.name)
    kind = POSITION_KINDS[name]

    def get_operator():
        sig, shifts = position_operator_signature(config, free_shift)
        if (pathlib.Path(sig.folder_name) / FILE_NAMES["info_name"]).exists():
            return Sparse_Matrix(sig)
        m = StoredCSR.from_position_operator(get_sector(config.hamiltonian), kind, shifts)
        try:
            return Sparse_Matrix.from_device(m, sig)
        finally:
            m.close()

    get_excited_H = config.hamiltonian.setup()
    return get_operator, config.hamiltonian, get_excited_H
