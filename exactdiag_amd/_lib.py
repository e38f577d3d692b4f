"""ctypes binding of ``libexactdiag_hip.so`` (the C ABI in include/exactdiag_hip.h).

The library is the only compute path: importing a function that needs it
raises when the library has not been built or cannot be loaded, there is no
CPU fallback.
"""

from __future__ import annotations

import ctypes as ct
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libexactdiag_hip.so"

ED_OK, ED_ERR_ARG, ED_ERR_HIP, ED_ERR_STATE, ED_ERR_CAPACITY = 0, -1, -2, -3, -4

(OP_SZQ, OP_NQ, OP_APLUS_DOWN, OP_APLUS_UP, OP_AMINUS_DOWN, OP_AMINUS_UP, OP_CURRENT_RUNG,
 OP_CURRENT_LEG) = range(8)
POS_HOLES, POS_SZ, POS_SINGLET = range(3)
# vector types of products and Lanczos runs (enum ed_precision)
PREC_C128, PREC_C64, PREC_R64, PREC_R32 = range(4)

EXPORTED = [
    "ed_last_error", "ed_version", "ed_cache_release", "ed_sector_create", "ed_sector_create_flip", "ed_sector_flip", "ed_sector_destroy",
    "ed_basis_build", "ed_basis_size",
    "ed_basis_export", "ed_lookup", "ed_lowest", "ed_term_columns", "ed_op_columns", "ed_hv", "ed_op_apply",
    "ed_lanczos_run", "ed_cf_eval", "ed_cf_eval_tail", "ed_vec_dots", "ed_vec_dots_noalloc", "ed_vec_axpys", "ed_vec_combine", "ed_vec_norm2", "ed_vec_scale",
    "ed_mlp_create", "ed_mlp_destroy", "ed_mlp_num_params", "ed_mlp_get_params", "ed_mlp_set_params",
    "ed_mlp_forward", "ed_mlp_gradients", "ed_mlp_train_step", "ed_gemm_bf16",
    "ed_batch_create", "ed_batch_destroy", "ed_batch_offsets", "ed_batch_hv", "ed_batch_lanczos", "ed_batch_tune",
    "ed_batch_set_config", "ed_batch_get_config", "ed_hv_tune", "ed_lanczos_ritz",
    "ed_hmat_create", "ed_hmat_destroy", "ed_hmat_info", "ed_hmat_hv", "ed_hmat_lanczos_run", "ed_hmat_lanczos_ritz",
    "ed_hmat_tune", "ed_hmat_set_config", "ed_hmat_line_probe", "ed_hmat_element_bytes", "ed_hmat_is_real", "ed_batch_create_stored",
    "ed_hmat_lanczos_run_tol", "ed_spectrum_change", "ed_hmat_lanczos_run_basis",
    "ed_csr_from_sector", "ed_csr_from_operator", "ed_csr_create", "ed_csr_info", "ed_csr_copy_out", "ed_csr_destroy",
    "ed_csr_sum_create", "ed_csr_sum_info", "ed_csr_sum_hv", "ed_csr_sum_destroy",
    "ed_position_op_columns", "ed_position_op_apply", "ed_position_expectation", "ed_position_correlations",
    "ed_csr_from_position_operator",
    "ed_sector_create_union", "ed_mk_create", "ed_mk_destroy", "ed_mk_info", "ed_mk_attach", "ed_mk_extract",
    "ed_mk_hv", "ed_mk_lanczos_run", "ed_mk_set_config", "ed_mk_lanczos_ritz", "ed_mk_residual_norms", "ed_mk_scale",
]

_vp = ct.c_void_p
_i64 = ct.c_int64
_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)

_SIGNATURES = {
    "ed_last_error": (ct.c_char_p, []),
    "ed_version": (ct.c_int, []),
    "ed_cache_release": (ct.c_int, []),
    "ed_sector_create": (ct.c_int, [ct.c_int] * 5 + [ct.POINTER(_vp)]),
    "ed_sector_create_flip": (ct.c_int, [ct.c_int] * 6 + [ct.POINTER(_vp)]),
    "ed_sector_flip": (ct.c_int, [_vp]),
    "ed_sector_destroy": (ct.c_int, [_vp]),
    "ed_basis_build": (ct.c_int, [_vp, _vp]),
    "ed_basis_size": (_i64, [_vp]),
    "ed_basis_export": (ct.c_int, [_vp, _vp, _vp, _vp]),
    "ed_lookup": (ct.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "ed_lowest": (ct.c_int, [_vp, _vp, _vp, _vp, _i64, _vp]),
    "ed_term_columns": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp]),
    "ed_op_columns": (ct.c_int, [_vp, _vp, ct.c_int, ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp]),
    "ed_hv": (ct.c_int, [_vp, _dp, _vp, _vp, _vp, ct.c_int, _vp]),
    "ed_op_apply": (ct.c_int, [_vp, _vp, ct.c_int, ct.c_int, ct.c_int, _vp, _vp, _vp]),
    "ed_lanczos_run": (ct.c_int, [_vp, _dp, _vp, ct.c_int, _dp, _dp, _ip, _vp, ct.c_int, _vp]),
    "ed_cf_eval": (ct.c_int, [_dp, _dp, ct.c_int, _dp, ct.c_int, ct.c_double, ct.c_double, ct.c_double, _dp]),
    "ed_cf_eval_tail": (ct.c_int, [_dp, _dp, ct.c_int, _dp, ct.c_int] + [ct.c_double] * 5 + [_dp]),
    "ed_vec_dots": (ct.c_int, [_vp, ct.c_int, _i64, _vp, _i64, _dp, _vp]),
    "ed_vec_dots_noalloc": (ct.c_int, [_vp, ct.c_int, _i64, _vp, _i64, _dp, _vp]),
    "ed_vec_axpys": (ct.c_int, [_vp, ct.c_int, _i64, _dp, _vp, _i64, _vp]),
    "ed_vec_combine": (ct.c_int, [_vp, ct.c_int, _i64, _dp, _vp, _i64, _vp]),
    "ed_vec_norm2": (ct.c_int, [_vp, _i64, _dp, _vp]),
    "ed_vec_scale": (ct.c_int, [_vp, _i64, ct.c_double, ct.c_double, _vp]),
    "ed_mlp_create": (ct.c_int, [ct.POINTER(ct.c_int), ct.c_int, ct.c_uint64, ct.POINTER(_vp)]),
    "ed_mlp_destroy": (ct.c_int, [_vp]),
    "ed_mlp_num_params": (_i64, [_vp]),
    "ed_mlp_get_params": (ct.c_int, [_vp, ct.POINTER(ct.c_float), _i64]),
    "ed_mlp_set_params": (ct.c_int, [_vp, ct.POINTER(ct.c_float), _i64]),
    "ed_mlp_forward": (ct.c_int, [_vp, _vp, ct.c_int, _vp, _vp]),
    "ed_mlp_gradients": (ct.c_int, [_vp, _vp, _vp, ct.c_int, ct.POINTER(ct.c_float), _dp, _vp]),
    "ed_mlp_train_step": (ct.c_int, [_vp, _vp, _vp, ct.c_int, ct.c_float, _dp, _vp]),
    "ed_gemm_bf16": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, _vp, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp]),
    "ed_batch_create": (ct.c_int, [ct.POINTER(_vp), _dp, ct.c_int, ct.POINTER(_vp)]),
    "ed_batch_destroy": (ct.c_int, [_vp]),
    "ed_batch_offsets": (ct.c_int, [_vp, ct.POINTER(_i64)]),
    "ed_batch_hv": (ct.c_int, [_vp, _vp, _vp, _dp, _vp]),
    "ed_batch_lanczos": (ct.c_int, [_vp, _vp, _ip, ct.c_int, _dp, _vp, _dp, _dp, _ip, ct.c_int, _vp]),
    "ed_batch_tune": (ct.c_int, [_vp, _vp, _vp, _ip, _ip, _dp, _vp]),
    "ed_batch_set_config": (ct.c_int, [_vp, ct.c_int, ct.c_int]),
    "ed_batch_get_config": (ct.c_int, [_vp, _ip, _ip, _ip]),
    "ed_hv_tune": (ct.c_int, [_vp, _dp, ct.c_int, _vp, _vp, _ip, _ip, _dp, _vp]),
    "ed_lanczos_ritz": (ct.c_int, [_vp, _dp, _vp, ct.c_int, _dp, _vp, ct.c_int, _vp]),
    "ed_hmat_create": (ct.c_int, [_vp, _dp, ct.POINTER(_vp), _vp]),
    "ed_hmat_destroy": (ct.c_int, [_vp]),
    "ed_hmat_info": (ct.c_int, [_vp, ct.POINTER(_i64), ct.POINTER(_i64), ct.POINTER(_i64), _ip]),
    "ed_hmat_hv": (ct.c_int, [_vp, _vp, _vp, _vp, ct.c_int, _vp]),
    "ed_hmat_lanczos_run": (ct.c_int, [_vp, _vp, ct.c_int, _dp, _dp, _ip, ct.c_int, _vp]),
    "ed_hmat_lanczos_ritz": (ct.c_int, [_vp, _vp, ct.c_int, _dp, _vp, ct.c_int, _vp]),
    "ed_hmat_tune": (ct.c_int, [_vp, _vp, _vp, ct.c_int, _ip, _ip, _ip, _dp, _vp]),
    "ed_hmat_line_probe": (ct.c_int, [ct.c_int, _i64, _dp, _vp]),
    "ed_hmat_set_config": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int]),
    "ed_hmat_element_bytes": (ct.c_int, [_vp]),
    "ed_hmat_is_real": (ct.c_int, [_vp]),
    "ed_batch_create_stored": (ct.c_int, [ct.POINTER(_vp), ct.c_int, ct.POINTER(_vp)]),
    "ed_csr_from_sector": (ct.c_int, [_vp, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(_vp), _vp]),
    "ed_csr_from_operator": (ct.c_int, [_vp, _vp, ct.c_int, ct.c_int, ct.c_int, ct.POINTER(_vp), _vp]),
    "ed_csr_create": (ct.c_int, [_i64, _i64, ct.c_int, ct.c_int, _i64, ct.c_int, _vp, _vp, _vp, _vp,
                                 ct.POINTER(_vp), _vp]),
    "ed_csr_info": (ct.c_int, [_vp, ct.POINTER(_i64), ct.POINTER(_i64), ct.POINTER(_i64), _ip, _ip]),
    "ed_csr_copy_out": (ct.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ed_csr_destroy": (ct.c_int, [_vp]),
    "ed_csr_sum_create": (ct.c_int, [ct.POINTER(_vp), _dp, ct.c_int, ct.POINTER(_vp), _vp]),
    "ed_csr_sum_info": (ct.c_int, [_vp, ct.POINTER(_i64), ct.POINTER(_i64), ct.POINTER(_i64), ct.POINTER(_i64)]),
    "ed_csr_sum_hv": (ct.c_int, [_vp, _vp, _vp, _vp]),
    "ed_csr_sum_destroy": (ct.c_int, [_vp]),
    "ed_position_op_columns": (ct.c_int, [_vp, ct.c_int, _vp, ct.c_int, ct.c_int, _vp, _vp, _vp, _vp]),
    "ed_position_op_apply": (ct.c_int, [_vp, ct.c_int, _vp, ct.c_int, _vp, _vp, _vp]),
    "ed_position_expectation": (ct.c_int, [_vp, ct.c_int, _vp, ct.c_int, _vp, _dp, _vp]),
    "ed_position_correlations": (ct.c_int, [_vp, ct.c_int, _vp, ct.c_int, _vp, _dp, _vp]),
    "ed_csr_from_position_operator": (ct.c_int, [_vp, ct.c_int, _vp, ct.c_int, ct.POINTER(_vp), _vp]),
    "ed_sector_create_union": (ct.c_int, [ct.c_int] * 3 + [ct.POINTER(_vp)]),
    "ed_mk_create": (ct.c_int, [_vp, _dp, ct.c_int, _ip, _ip, ct.POINTER(_vp), _vp]),
    "ed_mk_destroy": (ct.c_int, [_vp]),
    "ed_mk_info": (ct.c_int, [_vp, ct.POINTER(_i64), ct.POINTER(_i64), ct.POINTER(_i64)]),
    "ed_mk_attach": (ct.c_int, [_vp, ct.c_int, _vp, _vp, _vp, _vp]),
    "ed_mk_extract": (ct.c_int, [_vp, ct.c_int, _vp, _vp, _vp, _vp]),
    "ed_mk_hv": (ct.c_int, [_vp, _vp, _vp, _vp, ct.c_int, _vp]),
    "ed_mk_lanczos_run": (ct.c_int, [_vp, _vp, ct.c_int, _dp, _dp, _ip, ct.c_int, _vp]),
    "ed_mk_set_config": (ct.c_int, [_vp, ct.c_int, ct.c_int]),
    "ed_mk_lanczos_ritz": (ct.c_int, [_vp, _vp, ct.c_int, _dp, _vp, ct.c_int, _vp]),
    "ed_mk_residual_norms": (ct.c_int, [_vp, _vp, _vp, _dp, _dp, _vp]),
    "ed_mk_scale": (ct.c_int, [_vp, _vp, _dp, _vp]),
}



class LanczosStop(ct.Structure):
    """struct ed_lanczos_stop (include/exactdiag_hip.h)."""

    _fields_ = [("ws", _dp), ("n_w", ct.c_int), ("stride", ct.c_int), ("e0", ct.c_double), ("eta", ct.c_double),
                ("tol", ct.c_double), ("min_steps", ct.c_int), ("check_every", ct.c_int), ("delta", ct.c_int)]


_SIGNATURES["ed_hmat_lanczos_run_tol"] = (ct.c_int, [_vp, _vp, ct.c_int, ct.POINTER(LanczosStop), _dp, _dp, _ip, _ip,
                                                     ct.c_int, _vp])
_SIGNATURES["ed_hmat_lanczos_run_basis"] = (ct.c_int, [_vp, _vp, ct.c_int, _dp, _dp, _ip, _vp, _i64, ct.c_int, _vp])
_SIGNATURES["ed_spectrum_change"] = (ct.c_int, [_dp, _dp, ct.c_int, ct.c_int, ct.POINTER(LanczosStop), _dp])

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load the HIP library once (raises LibraryMissing when it is absent)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        # one HIP runtime per process: torch's own libamdhip64 must be loaded before the library, whose
        # dependency then resolves to it; loaded first, the library would pull in the system runtime and torch
        # would see no device
        import torch  # noqa: F401

        handle = ct.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def release_cached_memory() -> None:
    """Return the device blocks the library keeps for reuse to HIP (ed_cache_release)."""
    check(lib().ed_cache_release())


def check(rc: int) -> None:
    """Turn a status code into the exception the reference raises for it."""
    if rc == ED_OK:
        return
    msg = lib().ed_last_error().decode(errors="replace")
    if rc == ED_ERR_ARG:
        raise ValueError(msg)
    if rc == ED_ERR_CAPACITY:
        raise IndexError(msg)
    raise RuntimeError(f"exactdiag_hip status {rc}: {msg}")
