"""The reference's basis-map file (reference: general/basis_indexing.h).

``Basis_Index_Map::save`` (basis_indexing.h:128-146) writes
``[size_t D][int64 dense_to_sparse[D]][int32 num_sparse_in_dense[D]]`` in the
machine's byte order (little-endian here) and ``Basis_Index_Map(path)``
(basis_indexing.h:72-99) reads it back, failing on a missing or truncated
file.  ``get_basis_map`` (basis_indexing.h:245-253) loads the file when it
exists and otherwise computes the map and saves it, under
``<calc folder>/Sz_local-momentum_map.dat`` (tJ_spin_half_ladder/matrix_setup.pyx:390).

Here the basis is always built on the GPU (it takes milliseconds); the file
is written for reuse by either package and, when present, read and checked
against the GPU basis, so a cache written by the reference is validated and
a corrupt or foreign one is reported.
"""

from __future__ import annotations

import os
import pathlib

import numpy as np

BASIS_MAP_NAME = "Sz_local-momentum_map.dat"
_HEADER = np.dtype("<u8")
_SPARSE = np.dtype("<i8")
_ORBIT = np.dtype("<i4")


def save_basis_map(path, dense_to_sparse, num_sparse_in_dense) -> None:
    """Write the map in the reference's layout (creates the parent folders, basis_indexing.h:128-146)."""
    sparse = np.ascontiguousarray(dense_to_sparse, _SPARSE)
    orbit = np.ascontiguousarray(num_sparse_in_dense, _ORBIT)
    if sparse.shape != orbit.shape or sparse.ndim != 1:
        raise ValueError("dense_to_sparse and num_sparse_in_dense must be 1-d arrays of one length")
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(path.name + ".tmp")
    with open(tmp, "wb") as fp:
        fp.write(np.array([len(sparse)], _HEADER).tobytes())
        fp.write(sparse.tobytes())
        fp.write(orbit.tobytes())
    os.replace(tmp, path)


def load_basis_map(path):
    """(dense_to_sparse int64[D], num_sparse_in_dense int32[D]) from a map file (basis_indexing.h:72-99).

    Raises FileNotFoundError when the file is missing and ValueError when it is shorter or longer than its
    header says (the reference throws on both).
    """
    path = pathlib.Path(path)
    with open(path, "rb") as fp:
        head = fp.read(_HEADER.itemsize)
        if len(head) < _HEADER.itemsize:
            raise ValueError(f"Failed to read {path}: no header")
        D = int(np.frombuffer(head, _HEADER)[0])
        expected = _HEADER.itemsize + D * (_SPARSE.itemsize + _ORBIT.itemsize)
        size = path.stat().st_size
        if size != expected:
            raise ValueError(f"Failed to read {path}: {size} bytes, header says {D} states ({expected} bytes)")
        sparse = np.frombuffer(fp.read(D * _SPARSE.itemsize), _SPARSE).copy()
        orbit = np.frombuffer(fp.read(D * _ORBIT.itemsize), _ORBIT).copy()
    return sparse, orbit


def check_basis_map(path, dense_to_sparse, num_sparse_in_dense) -> None:
    """Raise ValueError unless the file holds exactly this map."""
    sparse, orbit = load_basis_map(path)
    if len(sparse) != len(dense_to_sparse):
        raise ValueError(f"{path} holds {len(sparse)} states, the sector has {len(dense_to_sparse)}")
    if not (np.array_equal(sparse, dense_to_sparse) and np.array_equal(orbit, num_sparse_in_dense)):
        bad = int(np.argmax((sparse != dense_to_sparse) | (orbit != num_sparse_in_dense)))
        raise ValueError(f"{path} differs from the sector's basis first at dense index {bad}")
