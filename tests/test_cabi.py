"""The C-ABI library builds, loads without a GPU and exports every declared symbol."""

import ctypes as ct
import pathlib
import re
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "exactdiag_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:char\s*\*|int64_t|int)\s+(ed_\w+)\s*\(", text, re.M)))


def test_header_declares_the_api():
    syms = declared_symbols()
    for must in ("ed_sector_create", "ed_basis_build", "ed_hv", "ed_lanczos_run", "ed_cf_eval", "ed_op_apply"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    from exactdiag_amd import _lib

    lib = _lib.lib()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert sorted(_lib.EXPORTED) == declared_symbols()
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ed_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_host_side_errors_without_gpu():
    from exactdiag_amd import _lib

    lib = _lib.lib()
    h = ct.c_void_p()
    assert lib.ed_sector_create(6, 2, 5, 0, 0, ct.byref(h)) == 0
    assert lib.ed_basis_size(h) == -1  # basis not built yet
    assert lib.ed_sector_destroy(h) == 0
    with pytest.raises(ValueError, match="num_down_spins"):
        _lib.check(lib.ed_sector_create(6, 2, 11, 0, 0, ct.byref(h)))
    with pytest.raises(ValueError, match="rung quantum number"):
        _lib.check(lib.ed_sector_create(6, 2, 5, 0, 2, ct.byref(h)))
    with pytest.raises(ValueError, match="num_rungs"):
        _lib.check(lib.ed_sector_create(17, 2, 5, 0, 0, ct.byref(h)))
    with pytest.raises(RuntimeError, match="basis not built"):
        _lib.check(lib.ed_hv(None, None, None, None, None, 0, None))
