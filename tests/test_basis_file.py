"""The reference's basis-map file format (general/basis_indexing.h:72-146), on the CPU.

The layout is [size_t D][long long dense_to_sparse[D]][int dense orbit sizes[D]] in the host's byte order
(x86-64: little-endian, 8/8/4 bytes; MATRIX_INDEX_TYPE is long long, general/types_and_constants.h:23).
The reference's own reader cannot be built here (basis_indexing.h needs <format>), so the bytes are checked
against that layout field by field, with the oracle's basis as the content.
"""

import numpy as np
import pytest

from exactdiag_amd.general import basis_indexing as bi


@pytest.fixture(scope="module")
def basis(oracle_mod):
    return oracle_mod.OracleSector(6, 2, 5, 1, 1).basis()


def test_layout_field_by_field(tmp_path, basis):
    sparse, orbit = basis
    path = tmp_path / "a" / "b" / bi.BASIS_MAP_NAME
    bi.save_basis_map(path, sparse, orbit)
    raw = path.read_bytes()
    D = len(sparse)
    assert len(raw) == 8 + 8 * D + 4 * D
    assert int.from_bytes(raw[:8], "little") == D
    assert np.array_equal(np.frombuffer(raw[8:8 + 8 * D], "<i8"), sparse)
    assert np.array_equal(np.frombuffer(raw[8 + 8 * D:], "<i4"), orbit)
    s2, o2 = bi.load_basis_map(path)
    assert s2.dtype == np.int64 and o2.dtype == np.int32
    assert np.array_equal(s2, sparse) and np.array_equal(o2, orbit)


def test_missing_truncated_and_long_files_fail(tmp_path, basis):
    sparse, orbit = basis
    with pytest.raises(FileNotFoundError):
        bi.load_basis_map(tmp_path / "none.dat")
    path = tmp_path / bi.BASIS_MAP_NAME
    bi.save_basis_map(path, sparse, orbit)
    raw = path.read_bytes()
    path.write_bytes(raw[:-1])
    with pytest.raises(ValueError):
        bi.load_basis_map(path)
    path.write_bytes(raw + b"\0")
    with pytest.raises(ValueError):
        bi.load_basis_map(path)
    path.write_bytes(raw[:5])
    with pytest.raises(ValueError):
        bi.load_basis_map(path)


def test_check_reports_a_foreign_map(tmp_path, basis):
    sparse, orbit = basis
    path = tmp_path / bi.BASIS_MAP_NAME
    bi.save_basis_map(path, sparse, orbit)
    bi.check_basis_map(path, sparse, orbit)
    bad = sparse.copy()
    bad[17] += 1
    bi.save_basis_map(path, bad, orbit)
    with pytest.raises(ValueError, match="dense index 17"):
        bi.check_basis_map(path, sparse, orbit)
    bi.save_basis_map(path, sparse[:-1], orbit[:-1])
    with pytest.raises(ValueError, match="states"):
        bi.check_basis_map(path, sparse, orbit)


def test_empty_map(tmp_path):
    path = tmp_path / bi.BASIS_MAP_NAME
    bi.save_basis_map(path, np.zeros(0, np.int64), np.zeros(0, np.int32))
    s, o = bi.load_basis_map(path)
    assert len(s) == 0 and len(o) == 0 and path.stat().st_size == 8
