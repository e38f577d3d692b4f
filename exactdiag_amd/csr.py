"""Matrices in the reference's stored CSR layout, held in HBM (ed_csr_* of the C ABI).

The reference writes every Hamiltonian term and excitation operator it
computes as a CSR folder (general/cython_sparse_matrices.pyx:110-184):
values complex128[U], value_inds int16[nnz], minor_inds int64[nnz],
major_pointers int64[major + 1].  ``StoredCSR`` holds one such matrix on the
GPU, built there from a sector (``from_sector``, ``from_operator``) or
uploaded from arrays read from a folder (``from_arrays``); ``CSRSum`` is a
weighted sum of them applied on the GPU as Added_Sparse_Matrices.dot does
(general/sparse_matrices.py:233-237).
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from exactdiag_amd._lib import check, lib
from exactdiag_amd.sector import Sector, _ptr, _stream, _torch


class StoredCSR:
    """One matrix in the reference's stored layout on the device."""

    def __init__(self, handle, keep=None):
        self._lib = lib()
        self._h = handle
        self._keep = keep  # sectors the matrix was built from (kept alive with it)
        rows, cols, nnz = ct.c_int64(), ct.c_int64(), ct.c_int64()
        nv, mc = ct.c_int(), ct.c_int()
        check(self._lib.ed_csr_info(self._h, ct.byref(rows), ct.byref(cols), ct.byref(nnz), ct.byref(nv),
                                    ct.byref(mc)))
        self.shape = (rows.value, cols.value)
        self.nnz, self.n_values, self.max_count = nnz.value, nv.value, mc.value

    @classmethod
    def from_sector(cls, sector: Sector, term: int, triangle_only: bool = True, row_major: bool = True):
        """Hamiltonian term (0 tl, 1 tr, 2 jl, 3 jr) as setup_hamiltonian stores it (default: row-major upper
        triangle, the reference's default config)."""
        h = ct.c_void_p()
        check(lib().ed_csr_from_sector(sector._h, int(term), int(bool(triangle_only)), int(bool(row_major)),
                                       ct.byref(h), _stream()))
        m = cls(h, keep=(sector,))
        m.row_major, m.hermitian_triangle = bool(row_major), bool(triangle_only)
        return m

    @classmethod
    def from_operator(cls, initial: Sector, final: Sector, op: int, qx: int, qy: int):
        """Excitation operator as setup_excitation_operator stores it (column-major, final x initial)."""
        h = ct.c_void_p()
        check(lib().ed_csr_from_operator(initial._h, final._h, int(op), int(qx), int(qy), ct.byref(h), _stream()))
        m = cls(h, keep=(initial, final))
        m.row_major, m.hermitian_triangle = False, False
        return m

    @classmethod
    def from_position_operator(cls, sector: Sector, kind: int, shifts):
        """Position-correlation operator (0 holes, 1 S^z, 2 singlet-singlet) as
        get_position_correlation_operator stores it (column-major, sorted)."""
        sh = np.ascontiguousarray(np.asarray(shifts, np.int32).reshape(-1, 2))
        h = ct.c_void_p()
        check(lib().ed_csr_from_position_operator(sector._h, int(kind), sh.ctypes.data, len(sh), ct.byref(h),
                                                  _stream()))
        m = cls(h, keep=(sector,))
        m.row_major, m.hermitian_triangle = False, False
        return m

    @classmethod
    def from_arrays(cls, shape, row_major: bool, hermitian_triangle: bool, values, value_inds, minor, major):
        """Upload arrays in the reference's layout (checked: pointers span the elements, indices in range)."""
        values = np.ascontiguousarray(values, np.complex128)
        value_inds = np.ascontiguousarray(value_inds, np.int16)
        minor = np.ascontiguousarray(minor, np.int64)
        major = np.ascontiguousarray(major, np.int64)
        rows, cols = (int(v) for v in shape)
        n_major = rows if row_major else cols
        if len(major) != n_major + 1:
            raise ValueError(f"major_pointers has {len(major)} entries, the shape needs {n_major + 1}")
        if len(value_inds) != len(minor):
            raise ValueError("value_inds and minor_inds differ in length")
        h = ct.c_void_p()
        check(lib().ed_csr_create(rows, cols, int(bool(row_major)), int(bool(hermitian_triangle)), len(minor),
                                  len(values), values.ctypes.data, value_inds.ctypes.data, minor.ctypes.data,
                                  major.ctypes.data, ct.byref(h), _stream()))
        m = cls(h)
        m.row_major, m.hermitian_triangle = bool(row_major), bool(hermitian_triangle)
        return m

    def arrays(self):
        """(values complex128, value_inds int16, minor_inds int64, major_pointers int64) on the host."""
        n_major = self.shape[0] if self.row_major else self.shape[1]
        values = np.empty(self.n_values, np.complex128)
        vinds = np.empty(self.nnz, np.int16)
        minor = np.empty(self.nnz, np.int64)
        major = np.empty(n_major + 1, np.int64)
        check(self._lib.ed_csr_copy_out(self._h, values.ctypes.data, vinds.ctypes.data, minor.ctypes.data,
                                        major.ctypes.data, _stream()))
        return values, vinds, minor, major

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ed_csr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CSRSum:
    """sum_T w_T M_T over stored matrices of one shape, applied on the GPU."""

    def __init__(self, mats, weights):
        if len(mats) == 0:
            raise ValueError("no matrices")
        if len(weights) != len(mats):
            raise ValueError("wrong weights supplied")
        self._lib = lib()
        w = np.ascontiguousarray(np.asarray(weights, np.complex128)).view(np.float64)
        handles = (ct.c_void_p * len(mats))(*[m._h for m in mats])
        h = ct.c_void_p()
        check(self._lib.ed_csr_sum_create(handles, w.ctypes.data_as(ct.POINTER(ct.c_double)), len(mats),
                                          ct.byref(h), _stream()))
        self._h = h
        rows, cols, nnz, slots = ct.c_int64(), ct.c_int64(), ct.c_int64(), ct.c_int64()
        check(self._lib.ed_csr_sum_info(self._h, ct.byref(rows), ct.byref(cols), ct.byref(nnz), ct.byref(slots)))
        self.shape = (rows.value, cols.value)
        self.nnz, self.slots = nnz.value, slots.value  # elements; slots with the slice padding
        self.device = _torch().device("cuda", _torch().cuda.current_device())

    def hv(self, x, out=None):
        """y = sum_T w_T M_T x for a complex128 device vector x."""
        torch = _torch()
        if x.dtype != torch.complex128 or not x.is_cuda:
            raise TypeError("x must be a complex128 device tensor")
        if x.numel() != self.shape[1]:
            raise ValueError(f"x has {x.numel()} entries, the matrix has {self.shape[1]} columns")
        x = x.contiguous()
        y = torch.empty(self.shape[0], dtype=torch.complex128, device=x.device) if out is None else out
        check(self._lib.ed_csr_sum_hv(self._h, _ptr(x), _ptr(y), _stream()))
        return y

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ed_csr_sum_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
