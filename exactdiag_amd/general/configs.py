"""Model-independent configuration classes (reference: general/configs.py).

The reference builds these with pydantic dataclasses and Python 3.12
generics; here they are plain dataclasses with the same field names,
defaults, JSON loading and assignment-time conversion of nested dicts
(``config.spectrum.operator_symmetry_qs = {"leg": 1, "rung": 0}`` works as in
the reference, general/configs.py:19-24).
"""

from __future__ import annotations

import dataclasses
import json
import os
import pathlib
import typing

import numpy as np


class _Validated:
    """Dataclass mixin: converts assigned values through ``_convert_<field>``."""

    def __setattr__(self, name, value):
        conv = getattr(type(self), f"_convert_{name}", None)
        if conv is not None:
            value = conv(value)
        object.__setattr__(self, name, value)

    @classmethod
    def load(cls, path):
        """Load from a JSON file (general/configs.py:60-66)."""
        with open(path, mode="r", encoding="utf-8") as fp:
            d = json.load(fp)
        return cls(**d)


def _check_fields(cls, kwargs):
    names = {f.name for f in dataclasses.fields(cls)}
    extra = set(kwargs) - names
    if extra:
        raise TypeError(f"{cls.__name__} got unexpected fields {sorted(extra)}")


@dataclasses.dataclass
class Weights_Base(_Validated):
    """Weights of the Hamiltonian terms."""


@dataclasses.dataclass
class Quantum_Numbers_Base(_Validated):
    """Quantum numbers conserved in the dense basis but not the sparse one."""

    def to_npint32(self):
        return np.array(list(dataclasses.asdict(self).values()), dtype=np.int32)


@dataclasses.dataclass
class Eigenpair_Part(_Validated):
    """general/configs.py:116-128."""

    num_eigenpairs: int
    num_recalculated_vecs_per_iteration: int = 4
    num_threads: typing.Optional[int] = None

    @staticmethod
    def _convert_num_eigenpairs(v):
        v = int(v)
        if v < 1:
            raise ValueError("num_eigenpairs must be >= 1")
        return v


def as_eigenpair_part(v):
    if isinstance(v, Eigenpair_Part):
        return v
    _check_fields(Eigenpair_Part, v)
    return Eigenpair_Part(**v)


@dataclasses.dataclass
class Operator_Part_Base(_Validated):
    name: typing.Any
    num_threads: typing.Optional[int] = None


@dataclasses.dataclass
class Spectrum_Part_Base(_Validated):
    """general/configs.py:174-191."""

    name: typing.Any
    operator_symmetry_qs: typing.Any
    broadening: float
    omega_max: float
    omega_min: float = 0
    omega_steps: int = 1000
    num_lanczos_vecs: int = 250
    num_threads: typing.Optional[int] = None

    def get_omegas(self):
        """Uniformly spaced energies (general/configs.py:189-191)."""
        return np.linspace(self.omega_min, self.omega_max, self.omega_steps)


CALCULATIONS_FOLDER_NAME = "calculations"
FIGURES_FOLDER_NAME = "figures"


def calc_folder(relative: pathlib.Path, root: str) -> pathlib.Path:
    path = pathlib.Path(root) / relative
    os.makedirs(path, exist_ok=True)
    return path
