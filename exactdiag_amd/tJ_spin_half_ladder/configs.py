"""Configuration of the anisotropic t-J model on a two-leg ladder
(reference: tJ_spin_half_ladder/configs.py).

Same classes, field names, defaults, validation errors and cache paths as the
reference; the GPU objects are created by ``get_translators`` / ``setup``.
"""

from __future__ import annotations

import dataclasses
import enum
import logging
import pathlib
import typing

import numpy as np

from exactdiag_amd.general import configs as gc

_logger = logging.getLogger(__name__)


def set_combinatorics(size):
    """comb[j, i] = C(i, j) (general/matrix_rules_utils.pyx:122-134)."""
    comb = np.zeros((size + 1, size + 1), dtype=np.int64)
    for i in range(size + 1):
        t = 1
        for j in range(i // 2 + 1):
            if i != j and i * j != 0:
                t *= (i + 1 - j) / j
            comb[j, i] = int(np.around(t))
            comb[i - j, i] = int(np.around(t))
    return comb


@dataclasses.dataclass
class Weights(gc.Weights_Base):
    """Weights of the Hamiltonian terms (configs.py:24-31)."""

    tl: float
    tr: float
    jl: float
    jr: float


@dataclasses.dataclass
class Quantum_Numbers(gc.Quantum_Numbers_Base):
    """[momentum along the legs in units of 2pi/L, leg-mirror parity] (configs.py:34-37)."""

    leg: int
    rung: int

    @staticmethod
    def _convert_leg(v):
        return int(v)

    @staticmethod
    def _convert_rung(v):
        v = int(v)
        if v not in (0, 1):
            raise ValueError(f"rung quantum number must be 0 or 1, got {v}")
        return v


def _as_weights(v):
    if isinstance(v, Weights):
        return v
    gc._check_fields(Weights, v)
    return Weights(**v)


def _as_qs(v):
    if v is None or isinstance(v, Quantum_Numbers):
        return v
    if isinstance(v, (list, tuple, np.ndarray)):
        return Quantum_Numbers(leg=int(v[0]), rung=int(v[1]))
    gc._check_fields(Quantum_Numbers, v)
    return Quantum_Numbers(**v)


@dataclasses.dataclass
class Hamiltonian_Config(gc._Validated):
    """Information necessary to construct a Hamiltonian block (configs.py:40-147)."""

    num_rungs: int
    num_holes: int
    weights: typing.Any
    symmetry_qs: typing.Any = None
    num_threads: int = 1
    total_spin_projection: typing.Optional[int] = None
    major: str = "row"
    triangle_only: bool = True
    CALCULATIONS_FOLDER_NAME: typing.ClassVar[str] = gc.CALCULATIONS_FOLDER_NAME
    FIGURES_FOLDER_NAME: typing.ClassVar[str] = gc.FIGURES_FOLDER_NAME

    _convert_weights = staticmethod(_as_weights)
    _convert_symmetry_qs = staticmethod(_as_qs)

    def __post_init__(self):
        if self.total_spin_projection is None:
            self.total_spin_projection = self._get_default_spin_projection()
        elif (self.num_spins - self.total_spin_projection) % 2 != 0:
            raise ValueError(
                f"Incompatible total_spin_projection {self.total_spin_projection}/2 with {self.num_spins} spins."
            )
        _logger.info(f"Config created: {self.num_rungs} rungs, {self.num_holes} holes, k={self.symmetry_qs}")

    @property
    def combinatorics_table(self):
        return set_combinatorics(self.num_nodes)

    def _get_default_spin_projection(self) -> int:
        return self.num_spins % 2

    @property
    def num_spin_states(self) -> int:
        return int(self.combinatorics_table[self.num_down_spins, self.num_spins])

    @property
    def num_spins(self) -> int:
        return self.num_nodes - self.num_holes

    @property
    def num_down_spins(self) -> int:
        return (self.num_spins - self.total_spin_projection) // 2

    @property
    def periodicities(self):
        return (self.num_rungs, 2)

    @property
    def num_nodes(self) -> int:
        return 2 * self.num_rungs

    def _get_relative_path(self) -> pathlib.Path:
        if self.symmetry_qs is None:
            raise ValueError("symmetry_qs must not be None to specify a path.")
        rungs_holes = f"rungs{self.num_rungs}_holes{self.num_holes}"
        qs_spins = f"kx{self.symmetry_qs.leg}_ky{self.symmetry_qs.rung}_down_spins{self.num_down_spins}"
        return pathlib.Path(rungs_holes) / qs_spins

    def get_calc_folder(self) -> pathlib.Path:
        return gc.calc_folder(self._get_relative_path(), self.CALCULATIONS_FOLDER_NAME)

    def get_figures_folder(self) -> pathlib.Path:
        return gc.calc_folder(self._get_relative_path(), self.FIGURES_FOLDER_NAME)

    def sector_key(self):
        return (self.num_rungs, self.num_holes, self.num_down_spins, int(self.symmetry_qs.leg),
                int(self.symmetry_qs.rung))

    def get_translators(self):
        """(translator, basis map, symmetries) backed by the GPU sector (matrix_setup.pyx:377-406)."""
        from exactdiag_amd.tJ_spin_half_ladder import matrix_setup

        return matrix_setup.py_get_ladder_translators(self)

    def setup(self):
        """Callable returning the GPU Hamiltonian operator (matrix_setup.pyx:41-129)."""
        from exactdiag_amd.tJ_spin_half_ladder import matrix_setup

        return matrix_setup.setup_hamiltonian(self)

    def get_symmetry_info(self):
        symmetry_qs = None if self.symmetry_qs is None else [int(i) for i in self.symmetry_qs.to_npint32()]
        dic = {"num_rungs": self.num_rungs, "num_holes": self.num_holes, "symmetry_qs": symmetry_qs}
        if self.total_spin_projection != self._get_default_spin_projection():
            dic["total_spin_projection"] = self.total_spin_projection
        return dic


def _as_hamiltonian(v):
    if isinstance(v, Hamiltonian_Config):
        return v
    gc._check_fields(Hamiltonian_Config, v)
    return Hamiltonian_Config(**v)


@dataclasses.dataclass
class Eigenpair_Config(gc._Validated):
    """Hamiltonian plus diagonalisation settings (general/configs.py:134-160)."""

    hamiltonian: typing.Any
    eigenpair: typing.Any

    _convert_hamiltonian = staticmethod(_as_hamiltonian)
    _convert_eigenpair = staticmethod(gc.as_eigenpair_part)

    def __post_init__(self):
        if self.eigenpair.num_threads is None:
            self.eigenpair.num_threads = self.hamiltonian.num_threads

    def get_eigenpair_paths(self):
        hamiltonian_string = "_".join(f"{n}{w:g}" for n, w in dataclasses.asdict(self.hamiltonian.weights).items())
        suffix = f"_{hamiltonian_string}_eigenpairs{self.eigenpair.num_eigenpairs}.npy"
        eigen_folder = self.hamiltonian.get_calc_folder() / "eigenpairs"
        return eigen_folder / f"eigenvals{suffix}", eigen_folder / f"eigenvecs{suffix}"


class Spectrum_Name(str, enum.Enum):
    """Names of the available operators (configs.py:163-172); DENSITY_Q is new."""

    CURRENT_RUNG = "current_rung"
    CURRENT_LEG = "current_leg"
    SZQ = "Szq"
    SPECTRAL_FUNCTION_PLUS = "spectral_function_plus"
    SPECTRAL_FUNCTION_MINUS = "spectral_function_minus"
    OFFIDAG_SPEC_FUNC_PLUS = "offdiagonal_spectral_function_plus"
    OFFIDAG_SPEC_FUNC_MINUS = "offdiagonal_spectral_function_minus"
    DENSITY_Q = "Nq"

    def __str__(self):
        return self.value


def _as_name(v):
    return v if isinstance(v, Spectrum_Name) else Spectrum_Name(v)


@dataclasses.dataclass
class Spectrum_Part(gc.Spectrum_Part_Base):
    _convert_name = staticmethod(_as_name)
    _convert_operator_symmetry_qs = staticmethod(_as_qs)


def _as_spectrum(v):
    if isinstance(v, Spectrum_Part):
        return v
    gc._check_fields(Spectrum_Part, v)
    return Spectrum_Part(**v)


@dataclasses.dataclass
class Limited_Spectrum_Config(gc._Validated):
    hamiltonian: typing.Any
    spectrum: typing.Any

    _convert_hamiltonian = staticmethod(_as_hamiltonian)
    _convert_spectrum = staticmethod(_as_spectrum)

    def __post_init__(self):
        if self.spectrum.num_threads is None:
            self.spectrum.num_threads = self.hamiltonian.num_threads

    def setup_excitation_operator(self):
        from exactdiag_amd.tJ_spin_half_ladder import matrix_setup

        return matrix_setup.setup_excitation_operator(self)


@dataclasses.dataclass
class Full_Spectrum_Config(gc._Validated):
    """Hamiltonian, spectrum and eigenpair settings (configs.py:214-233)."""

    hamiltonian: typing.Any
    spectrum: typing.Any
    eigenpair: typing.Any

    _convert_hamiltonian = staticmethod(_as_hamiltonian)
    _convert_spectrum = staticmethod(_as_spectrum)
    _convert_eigenpair = staticmethod(gc.as_eigenpair_part)

    def __post_init__(self):
        if self.spectrum.num_threads is None:
            self.spectrum.num_threads = self.hamiltonian.num_threads
        if self.eigenpair.num_threads is None:
            self.eigenpair.num_threads = self.hamiltonian.num_threads

    get_eigenpair_paths = Eigenpair_Config.get_eigenpair_paths
    setup_excitation_operator = Limited_Spectrum_Config.setup_excitation_operator

    def get_spectrum_path(self, operator_name_suffix=""):
        folder = self.hamiltonian.get_calc_folder() / f"{self.spectrum.name}_spectra"
        suffix = _get_spectrum_figure_suffix(self.hamiltonian, self.spectrum, operator_name_suffix)
        return folder / f"{self.spectrum.name}{suffix}.npz"


def _get_spectrum_figure_suffix(hamiltonian_config, spectrum_config, operator_name_suffix):
    """configs.py:366-379."""
    weight_string = "_".join(f"{n}{w}" for n, w in dataclasses.asdict(hamiltonian_config.weights).items())
    operator_name_suffix = f"_{operator_name_suffix}" if operator_name_suffix else ""
    if spectrum_config.operator_symmetry_qs is None:
        q_string = ""
    else:
        q_string = f"_q{spectrum_config.operator_symmetry_qs.to_npint32()}"
    ws_string = f"_ws{spectrum_config.omega_min:g},{spectrum_config.omega_max:g},{spectrum_config.omega_steps:g}"
    broad_string = f"_broad{spectrum_config.broadening}"
    vecs_string = f"_vecs{spectrum_config.num_lanczos_vecs}"
    return f"{operator_name_suffix}{q_string}{ws_string}{broad_string}{vecs_string}_{weight_string}"


class Position_Correlation_Name(str, enum.Enum):
    """Names of the position correlations (configs.py:236-241)."""

    HOLE_CORRELATIONS = "hole_correlations"
    SZ_CORRELATIONS = "Sz_correlations"
    SINGLET_SINGLET = "singlet-singlet"

    def __str__(self):
        return self.value


@dataclasses.dataclass
class Position_Shift(gc._Validated):
    """A relative node position (leg shift, rung shift), configs.py:244-247 with general/configs.py:47-52."""

    leg: int
    rung: int

    @staticmethod
    def _convert_leg(v):
        return int(v)

    @staticmethod
    def _convert_rung(v):
        v = int(v)
        if v not in (0, 1):
            raise ValueError("rung shift must be 0 or 1")
        return v

    def to_npint32(self):
        return np.array([self.leg, self.rung], dtype=np.int32)


def _as_shift(v):
    if isinstance(v, Position_Shift):
        return v
    gc._check_fields(Position_Shift, v)
    return Position_Shift(**v)


@dataclasses.dataclass
class Position_Correlation_Part(gc._Validated):
    """Operator part of a position correlation (configs.py:250-274): name, fixed relative shifts between the
    operator terms, threads (None: copied from the Hamiltonian config)."""

    name: typing.Any
    fixed_distances: typing.Any
    num_threads: typing.Optional[int] = None

    @staticmethod
    def _convert_name(v):
        return v if isinstance(v, Position_Correlation_Name) else Position_Correlation_Name(v)

    @staticmethod
    def _convert_fixed_distances(v):
        return [_as_shift(e) for e in v]


def _as_correlations(v):
    if isinstance(v, Position_Correlation_Part):
        return v
    gc._check_fields(Position_Correlation_Part, v)
    return Position_Correlation_Part(**v)


@dataclasses.dataclass
class Limited_Position_Correlation_Config(gc._Validated):
    """Hamiltonian and position-correlation settings (configs.py:277-309)."""

    hamiltonian: typing.Any
    correlations: typing.Any

    _convert_hamiltonian = staticmethod(_as_hamiltonian)
    _convert_correlations = staticmethod(_as_correlations)

    def __post_init__(self):
        if self.correlations.num_threads is None:
            self.correlations.num_threads = self.hamiltonian.num_threads

    def setup_excitation_operator(self, free_shift):
        """(get_operator, Hamiltonian config, get_excited_H) of the operator at relative position free_shift."""
        from exactdiag_amd.tJ_spin_half_ladder import matrix_setup

        return matrix_setup.get_position_correlation_operator(self, _as_shift(free_shift))


@dataclasses.dataclass
class Full_Position_Correlation_Config(gc._Validated):
    """Hamiltonian, eigenpair and position-correlation settings (configs.py:312-331)."""

    hamiltonian: typing.Any
    eigenpair: typing.Any
    correlations: typing.Any

    _convert_hamiltonian = staticmethod(_as_hamiltonian)
    _convert_eigenpair = staticmethod(gc.as_eigenpair_part)
    _convert_correlations = staticmethod(_as_correlations)

    def __post_init__(self):
        if self.correlations.num_threads is None:
            self.correlations.num_threads = self.hamiltonian.num_threads
        if self.eigenpair.num_threads is None:
            self.eigenpair.num_threads = self.hamiltonian.num_threads

    get_eigenpair_paths = Eigenpair_Config.get_eigenpair_paths
    setup_excitation_operator = Limited_Position_Correlation_Config.setup_excitation_operator

    def get_spectrum_path(self, operator_name_suffix=""):
        folder = self.hamiltonian.get_calc_folder() / f"{self.correlations.name}_spectra"
        suffix = _get_position_correlation_figure_suffix(self.hamiltonian, self.correlations, operator_name_suffix)
        return folder / f"{self.correlations.name}{suffix}.npz"


def _get_position_correlation_figure_suffix(hamiltonian_config, correlation_config, operator_name_suffix):
    """configs.py:382-389."""
    weight_string = "_".join(f"{n}{w}" for n, w in dataclasses.asdict(hamiltonian_config.weights).items())
    operator_name_suffix = f"_{operator_name_suffix}" if operator_name_suffix else ""
    fixed_string = f"{[shift.to_npint32().tolist() for shift in correlation_config.fixed_distances]}"
    return f"{operator_name_suffix}_fixed{fixed_string}_{weight_string}"
