"""The reference's matrix folders on the CPU: signature, file layout, reopening, dense form.

A folder is matrix_info.json plus four raw arrays written with np.memmap
(general/cython_sparse_matrices.pyx:110-184): values complex128, value_inds
int16, minor_inds int64, major_pointers int64, no header.  The content here
is the oracle's restatement of calc_sparse_matrix; the GPU build is compared
with it in tests/test_gpu_csr.py.
"""

import json
import logging

import numpy as np
import pytest

from exactdiag_amd.general import sparse_matrices as sm


def _signature(folder, D, name="jl", strings=("hermitian", "sorted", "triangle_only"), major="row"):
    return sm.Signature(matrix_name=name, shape=(D, D), symmetry_strings=strings, max_values_per_column=7,
                        major=major, initial_system_info={"num_rungs": 5, "num_holes": 2, "symmetry_qs": [1, 0]},
                        folder_name=str(folder), **sm.FILE_NAMES)


@pytest.fixture(scope="module")
def sector(oracle_mod):
    return oracle_mod.OracleSector(5, 2, 4, 1, 0)


def test_signature_rules():
    with pytest.raises(ValueError):
        _signature("x", 3, strings=("sorted", "triangle_only"))
    with pytest.raises(ValueError):
        _signature("x", 3, major="diagonal")
    with pytest.raises(ValueError):
        sm._as_signature({"matrix_name": "a", "colour": 1})
    sig = _signature("x", 3)
    assert sig["matrix_name"] == "jl" and sig.is_row_major and sig.is_hermitian_triangle
    assert list(sig.as_json())[0] == "matrix_name" and sig.as_json()["shape"] == [3, 3]


def test_folder_layout_and_round_trip(tmp_path, sector):
    values, vinds, minor, major = sector.term_csr(2)
    folder = tmp_path / "calc" / "jl"
    sm.write_matrix_folder(_signature(folder, sector.dim), values, vinds, minor, major)
    info = json.loads((folder / "matrix_info.json").read_text())
    assert info["folder_name"] == str(folder) and info["final_system_info"] is None
    assert (folder / "values.npy").stat().st_size == 16 * len(values)
    assert (folder / "value_inds.npy").stat().st_size == 2 * len(vinds)
    assert (folder / "minor_inds.npy").stat().st_size == 8 * len(minor)
    assert (folder / "major_pointers.npy").stat().st_size == 8 * (sector.dim + 1)
    back = sm.read_matrix_folder(sm.read_signature(folder))
    for a, b in zip(back, (values, vinds, minor, major)):
        assert np.array_equal(np.asarray(a), b)
    with pytest.raises(ValueError):
        sm.write_matrix_folder(_signature(folder, sector.dim + 1), values, vinds, minor, major)


def test_reopen_prefers_the_saved_signature(tmp_path, sector, caplog):
    values, vinds, minor, major = sector.term_csr(0)
    folder = tmp_path / "tl"
    sm.write_matrix_folder(_signature(folder, sector.dim, name="tl"), values, vinds, minor, major)
    requested = _signature(folder, sector.dim, name="tl")
    requested.max_values_per_column = 99
    with caplog.at_level(logging.WARNING):
        m = sm.Sparse_Matrix(requested)
    assert "different signature" in caplog.text
    assert m.signature.max_values_per_column == 7
    assert m.name == "tl" and m.is_row_major and m.is_sorted and m.is_hermitian_triangle
    assert list(m.shape) == [sector.dim, sector.dim]


def test_dense_form_is_the_references(tmp_path, sector):
    """Sparse_Matrix.toarray / Added_Sparse_Matrices.toarray as the reference builds them
    (general/sparse_matrices.py:110-125, :239-243): the oracle's dense matrix of the same weights."""
    w = [0.45, 0.36, 0.15, 0.12]
    mats = []
    for t, name in enumerate(("tl", "tr", "jl", "jr")):
        folder = tmp_path / name
        sm.write_matrix_folder(_signature(folder, sector.dim, name=name), *sector.term_csr(t))
        mats.append(sm.Sparse_Matrix(_signature(folder, sector.dim, name=name)))
    added = sm.Added_Sparse_Matrices(mats, w)
    assert np.max(np.abs(added.toarray() - sector.toarray(w))) < 1e-14
    with pytest.raises(ValueError):
        sm.Added_Sparse_Matrices(mats, w[:3])


def test_empty_matrix_folder(tmp_path, oracle_mod):
    """A term without elements is stored as one explicit zero in the last row (pyx:123-129)."""
    o = oracle_mod.OracleSector(4, 0, 4, 0, 0)  # no holes: no hopping
    values, vinds, minor, major = o.term_csr(0)
    assert len(values) == 1 and values[0] == 0 and list(vinds) == [0] and list(minor) == [0]
    assert major[-1] == 1 and not major[:-1].any()
    folder = tmp_path / "tl"
    sm.write_matrix_folder(_signature(folder, o.dim, name="tl"), values, vinds, minor, major)
    m = sm.Sparse_Matrix(_signature(folder, o.dim, name="tl"))
    assert not m.toarray().any()
