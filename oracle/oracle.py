"""ctypes wrapper of the CPU oracle (test infrastructure only).

Loads ``oracle/_build/libedoracle.so`` (the step-by-step restatement of the
reference's CPU algorithms in ``ed_oracle.cpp``) and, when it was built in a
container that holds the reference checkout, ``oracle/_ref/libedref.so`` (a
harness compiled from the reference's own ``symmetry_ladder.h`` and
``cpp_sparse_dot.h``).  Only tests, the smoke check and the bench's CPU
baseline import this module; the product path never does.
"""

from __future__ import annotations

import ctypes as ct
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_SO = HERE / "_build" / "libedoracle.so"
REF_SO = HERE / "_ref" / "libedref.so"

_i64p = ct.POINTER(ct.c_int64)
_i32p = ct.POINTER(ct.c_int32)
_i16p = ct.POINTER(ct.c_int16)
_dp = ct.POINTER(ct.c_double)

# Operator ids shared with ed_oracle.cpp and the HIP library.
OP_SZQ, OP_NQ, OP_APLUS_DOWN, OP_APLUS_UP, OP_AMINUS_DOWN, OP_AMINUS_UP, OP_CURRENT_RUNG, OP_CURRENT_LEG = range(8)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def _load(path):
    if not path.exists():
        raise FileNotFoundError(f"{path} is not built; run `make -C oracle`")
    return ct.CDLL(str(path))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(ORACLE_SO)
        L = _lib
        L.orc_sector_new.restype = ct.c_void_p
        L.orc_sector_new.argtypes = [ct.c_int] * 6
        L.orc_sector_free.argtypes = [ct.c_void_p]
        L.orc_sector_dim.restype = ct.c_int64
        L.orc_sector_dim.argtypes = [ct.c_void_p]
        L.orc_sector_basis.argtypes = [ct.c_void_p, _i64p, _i32p]
        L.orc_comb.restype = ct.c_int64
        L.orc_comb.argtypes = [ct.c_int] * 3
        L.orc_unrank.argtypes = [ct.c_void_p, ct.c_int64, _i16p]
        L.orc_rank.restype = ct.c_int64
        L.orc_rank.argtypes = [ct.c_void_p, _i16p]
        L.orc_dense.restype = ct.c_int64
        L.orc_dense.argtypes = [ct.c_void_p, ct.c_int64]
        L.orc_hole_eligible.argtypes = [ct.c_void_p, ct.c_int64]
        L.orc_orbit.argtypes = [ct.c_void_p, _i16p, ct.c_int, _i16p, _dp]
        L.orc_lowest.argtypes = [ct.c_void_p, _i16p, _i64p, _dp]
        L.orc_term_columns.restype = ct.c_int64
        L.orc_term_columns.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, _i64p, _i64p, _dp]
        L.orc_op_columns.restype = ct.c_int64
        L.orc_op_columns.argtypes = [ct.c_void_p, ct.c_void_p] + [ct.c_int] * 4 + [_i64p, _i64p, _dp]
        L.orc_op_weights.argtypes = [ct.c_int] * 4 + [_dp]
        L.orc_q_dot_r.restype = ct.c_double
        L.orc_q_dot_r.argtypes = [ct.c_int] * 4
        L.orc_build_term_csr.argtypes = [ct.c_void_p, ct.c_int, ct.c_int]
        L.orc_build_term_csr_rows.argtypes = [ct.c_void_p, ct.c_int, ct.c_int64, ct.c_int64, ct.c_int]
        L.orc_term_csr_sizes.restype = ct.c_int64
        L.orc_term_csr_sizes.argtypes = [ct.c_void_p, ct.c_int, _i64p, _i64p]
        L.orc_term_csr_copy.argtypes = [ct.c_void_p, ct.c_int, _dp, _i16p, _i64p, _i64p]
        L.orc_build_layout_csr.argtypes = [ct.c_void_p] + [ct.c_int] * 4
        L.orc_position_columns.restype = ct.c_int64
        L.orc_position_columns.argtypes = [ct.c_void_p, ct.c_int, _i32p, ct.c_int, ct.c_int, _i64p, _i64p, _dp]
        L.orc_position_correlations.argtypes = [ct.c_void_p, ct.c_int, _i32p, ct.c_int, _dp, _dp]
        L.orc_build_op_csr.argtypes = [ct.c_void_p, ct.c_void_p] + [ct.c_int] * 4
        L.orc_hv.argtypes = [ct.c_void_p, _dp, _dp, _dp, ct.c_int]
        L.orc_cf.argtypes = [_dp, _dp, ct.c_int, _dp, ct.c_int, _dp]
        L.orc_lanczos.argtypes = [ct.c_void_p, _dp, _dp, ct.c_int, _dp, _dp, ct.c_int]
        L.orc_lanczos.restype = ct.c_int
        L.orc_last_error.restype = ct.c_char_p
    return _lib


def ref_available() -> bool:
    return REF_SO.exists()


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_SO)
        R = _ref
        R.ref_sector_new.restype = ct.c_void_p
        R.ref_sector_new.argtypes = [ct.c_int] * 6
        R.ref_sector_free.argtypes = [ct.c_void_p]
        R.ref_sector_dim.restype = ct.c_int64
        R.ref_sector_dim.argtypes = [ct.c_void_p]
        R.ref_sector_basis.argtypes = [ct.c_void_p, _i64p, _i32p]
        R.ref_hole_eligible.restype = ct.c_int64
        R.ref_hole_eligible.argtypes = [ct.c_void_p, ct.c_int64]
        R.ref_unrank.argtypes = [ct.c_void_p, ct.c_int64, _i16p]
        R.ref_rank.restype = ct.c_int64
        R.ref_rank.argtypes = [ct.c_void_p, _i16p]
        R.ref_orbit.argtypes = [ct.c_void_p, _i16p, ct.c_int, _i16p, _dp]
        R.ref_lowest.argtypes = [ct.c_void_p, _i16p, _i64p, _dp]
        R.ref_q_dot_r.argtypes = [ct.c_int] * 4 + [_dp]
        R.ref_csr_dot_rows.argtypes = [ct.c_int64, ct.c_int64, ct.c_int64, _i64p, _i16p, _i64p, _dp, _dp, _dp,
                                       ct.c_double, ct.c_double, ct.c_int]
        R.ref_csr_dot.argtypes = [ct.c_int64, _i64p, _i16p, _i64p, _dp, _dp, _dp, ct.c_double, ct.c_double,
                                  ct.c_int, ct.c_int]
    return _ref


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


class OracleSector:
    """One symmetry sector (num_rungs L, holes h, down spins nd, k=(kx, ky))."""

    def __init__(self, L, h, nd, kx, ky, threads=None):
        self.L, self.h, self.nd, self.kx, self.ky = L, h, nd, kx, ky
        self.N = 2 * L
        self._lib = lib()
        self._h = self._lib.orc_sector_new(L, h, nd, kx, ky, threads or default_threads())
        if not self._h:
            raise RuntimeError(self._lib.orc_last_error().decode())
        self.dim = int(self._lib.orc_sector_dim(self._h))
        self._terms_built = set()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.orc_sector_free(h)
            self._h = None

    def basis(self):
        sp = np.empty(self.dim, np.int64)
        cn = np.empty(self.dim, np.int32)
        self._lib.orc_sector_basis(self._h, _ptr(sp, _i64p), _ptr(cn, _i32p))
        return sp, cn

    def unrank(self, sparse):
        st = np.empty(self.N, np.int16)
        self._lib.orc_unrank(self._h, int(sparse), _ptr(st, _i16p))
        return st

    def rank(self, st):
        st = np.ascontiguousarray(st, np.int16)
        return int(self._lib.orc_rank(self._h, _ptr(st, _i16p)))

    def dense(self, sparse):
        return int(self._lib.orc_dense(self._h, int(sparse)))

    def hole_eligible(self, hole_index):
        return bool(self._lib.orc_hole_eligible(self._h, int(hole_index)))

    def orbit(self, st, check_lowest):
        st = np.ascontiguousarray(st, np.int16)
        out = np.empty((self.N, self.N), np.int16)
        amp = np.empty(2 * self.N, np.float64)
        n = self._lib.orc_orbit(self._h, _ptr(st, _i16p), int(check_lowest), _ptr(out, _i16p), _ptr(amp, _dp))
        return out[:n].copy(), amp[: 2 * n].view(np.complex128).copy()

    def lowest(self, st):
        st = np.ascontiguousarray(st, np.int16)
        ind = np.zeros(1, np.int64)
        amp = np.zeros(2, np.float64)
        self._lib.orc_lowest(self._h, _ptr(st, _i16p), _ptr(ind, _i64p), _ptr(amp, _dp))
        return int(ind[0]), complex(amp[0], amp[1])

    def term_columns(self, term, triangle=True, threads=None):
        t = threads or default_threads()
        nnz = self._lib.orc_term_columns(self._h, term, int(triangle), t, None, None, None)
        ptr = np.empty(self.dim + 1, np.int64)
        rows = np.empty(max(nnz, 1), np.int64)
        vals = np.empty(max(nnz, 1), np.complex128)
        self._lib.orc_term_columns(self._h, term, int(triangle), t, _ptr(ptr, _i64p), _ptr(rows, _i64p),
                                   _ptr(vals.view(np.float64), _dp))
        return ptr, rows[:nnz], vals[:nnz]

    def op_columns(self, final: "OracleSector", op, qx, qy, threads=None):
        t = threads or default_threads()
        nnz = self._lib.orc_op_columns(self._h, final._h, op, qx, qy, t, None, None, None)
        ptr = np.empty(self.dim + 1, np.int64)
        rows = np.empty(max(nnz, 1), np.int64)
        vals = np.empty(max(nnz, 1), np.complex128)
        self._lib.orc_op_columns(self._h, final._h, op, qx, qy, t, _ptr(ptr, _i64p), _ptr(rows, _i64p),
                                 _ptr(vals.view(np.float64), _dp))
        return ptr, rows[:nnz], vals[:nnz]

    def position_columns(self, kind, shifts, threads=None):
        """Columns of a position-correlation operator (kind 0 holes, 1 S^z, 2 singlet-singlet); shifts =
        fixed distances then the free shift, as (leg, rung) pairs."""
        sh = np.ascontiguousarray(np.asarray(shifts, np.int32).reshape(-1, 2))
        t = threads or default_threads()
        nnz = self._lib.orc_position_columns(self._h, kind, _ptr(sh, _i32p), len(sh), t, None, None, None)
        if nnz < 0:
            raise ValueError(self._lib.orc_last_error().decode())
        ptr = np.empty(self.dim + 1, np.int64)
        rows = np.empty(max(nnz, 1), np.int64)
        vals = np.empty(max(nnz, 1), np.complex128)
        self._lib.orc_position_columns(self._h, kind, _ptr(sh, _i32p), len(sh), t, _ptr(ptr, _i64p),
                                       _ptr(rows, _i64p), _ptr(vals.view(np.float64), _dp))
        return ptr, rows[:nnz], vals[:nnz]

    def position_correlations(self, kind, fixed, psi):
        """calculate_hole_spin_projection_correlations: one value per node index (kind 0 holes, 1 S^z)."""
        fx = np.ascontiguousarray(np.asarray(fixed, np.int32).reshape(-1, 2))
        psi = np.ascontiguousarray(psi, np.complex128)
        out = np.empty(2 * self.L, np.float64)
        if self._lib.orc_position_correlations(self._h, kind, _ptr(fx, _i32p), len(fx),
                                               _ptr(psi.view(np.float64), _dp), _ptr(out, _dp)) != 0:
            raise ValueError(self._lib.orc_last_error().decode())
        return out

    def build_csr(self, terms=(0, 1, 2, 3), threads=None):
        for t in terms:
            if t in self._terms_built:
                continue
            if self._lib.orc_build_term_csr(self._h, t, threads or default_threads()) != 0:
                raise IndexError(self._lib.orc_last_error().decode())
            self._terms_built.add(t)

    def build_csr_rows(self, term, lo, hi, threads=None):
        """Stored CSR of one term for rows [lo, hi) only (replaces any full build of that term)."""
        if self._lib.orc_build_term_csr_rows(self._h, term, int(lo), int(hi), threads or default_threads()) != 0:
            raise IndexError(self._lib.orc_last_error().decode())
        self._terms_built.add(term)

    def layout_csr(self, term, triangle=True, row_major=True, threads=None):
        """Stored CSR of one term in any of the reference's layouts (triangle_only, major)."""
        if self._lib.orc_build_layout_csr(self._h, term, int(triangle), int(row_major),
                                          threads or default_threads()) != 0:
            raise IndexError(self._lib.orc_last_error().decode())
        return self._slot_csr(4)

    def op_csr(self, final: "OracleSector", op, qx, qy, threads=None):
        """Stored CSR of an excitation operator (column-major, final rows x initial columns)."""
        if self._lib.orc_build_op_csr(self._h, final._h, op, qx, qy, threads or default_threads()) != 0:
            raise IndexError(self._lib.orc_last_error().decode())
        return self._slot_csr(4)

    def term_csr(self, term):
        self.build_csr((term,))
        return self._slot_csr(term)

    def _slot_csr(self, term):
        nu = np.zeros(1, np.int64)
        nr = np.zeros(1, np.int64)
        nnz = self._lib.orc_term_csr_sizes(self._h, term, _ptr(nu, _i64p), _ptr(nr, _i64p))
        values = np.empty(int(nu[0]), np.complex128)
        vinds = np.empty(nnz, np.int16)
        minor = np.empty(nnz, np.int64)
        major = np.empty(int(nr[0]) + 1, np.int64)  # the rows the stored CSR covers
        self._lib.orc_term_csr_copy(self._h, term, _ptr(values.view(np.float64), _dp), _ptr(vinds, _i16p),
                                    _ptr(minor, _i64p), _ptr(major, _i64p))
        return values, vinds, minor, major

    def hv(self, weights, x, threads=None):
        w = np.ascontiguousarray(np.asarray(weights, np.complex128))
        self.build_csr(tuple(t for t in range(4) if w[t] != 0), threads)
        x = np.ascontiguousarray(x, np.complex128)
        y = np.empty(self.dim, np.complex128)
        self._lib.orc_hv(self._h, _ptr(w.view(np.float64), _dp), _ptr(x.view(np.float64), _dp),
                         _ptr(y.view(np.float64), _dp), threads or default_threads())
        return y

    def toarray(self, weights):
        """Dense H as the reference's Added_Sparse_Matrices.toarray builds it (the conjugate of the dot matrix)."""
        w = np.asarray(weights, np.complex128)
        H = np.zeros((self.dim, self.dim), np.complex128)
        for t in range(4):
            if w[t] == 0:
                continue
            values, vinds, minor, major = self.term_csr(t)
            M = np.zeros_like(H)
            # general/sparse_matrices.py:116-123: entry (major i, minor u) is
            # written at [u, i], so this is the transpose of the matrix `dot` applies
            for i in range(self.dim):
                for k in range(major[i], major[i + 1]):
                    u = minor[k]
                    M[u, i] += values[vinds[k]]
                    if u != i:
                        M[i, u] += np.conj(values[vinds[k]])
            H += w[t] * M
        return H

    def lanczos(self, weights, v0, nsteps, threads=None):
        w = np.ascontiguousarray(np.asarray(weights, np.complex128))
        self.build_csr(tuple(t for t in range(4) if w[t] != 0), threads)
        v0 = np.ascontiguousarray(v0, np.complex128)
        a = np.zeros(nsteps, np.float64)
        b = np.zeros(nsteps + 1, np.float64)
        n = self._lib.orc_lanczos(self._h, _ptr(w.view(np.float64), _dp), _ptr(v0.view(np.float64), _dp), nsteps,
                                  _ptr(a, _dp), _ptr(b, _dp), threads or default_threads())
        return a[:n].copy(), b[:n].copy()


def q_dot_r(L, site, qx, qy):
    """The oracle's q.r (radians) of one site."""
    return lib().orc_q_dot_r(L, site, qx, qy)


def ref_q_dot_r(L, site, qx, qy):
    """The reference's own get_q_dot_r (compiled in oracle/_ref): (real, imag)."""
    out = np.zeros(2)
    ref().ref_q_dot_r(L, site, qx, qy, _ptr(out, _dp))
    return out[0], out[1]


def op_weights(L, op, qx, qy):
    """(diagonal, off-diagonal) weight of every combination of an operator: 2L rows, 4L for the leg current."""
    n = 2 * L * (2 if op == OP_CURRENT_LEG else 1)
    out = np.empty(4 * n, np.float64)
    lib().orc_op_weights(L, op, qx, qy, _ptr(out, _dp))
    return out.view(np.complex128).reshape(n, 2)


def cf_imag(a, b, zs):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    z = np.ascontiguousarray(zs, np.complex128)
    out = np.empty(len(z), np.float64)
    lib().orc_cf(_ptr(a, _dp), _ptr(b, _dp), len(a), _ptr(z.view(np.float64), _dp), len(z), _ptr(out, _dp))
    return out


def comb(size, j, i):
    return int(lib().orc_comb(size, j, i))


class RefSector:
    """The reference's own C++ (symmetry_ladder.h) through oracle/_ref."""

    def __init__(self, L, h, nd, kx, ky, threads=None):
        self.N = 2 * L
        self._lib = ref()
        self._h = self._lib.ref_sector_new(L, h, nd, kx, ky, threads or default_threads())
        self.dim = int(self._lib.ref_sector_dim(self._h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.ref_sector_free(h)
            self._h = None

    def basis(self):
        sp = np.empty(self.dim, np.int64)
        cn = np.empty(self.dim, np.int32)
        self._lib.ref_sector_basis(self._h, _ptr(sp, _i64p), _ptr(cn, _i32p))
        return sp, cn

    def hole_eligible(self, hole_index):
        return self._lib.ref_hole_eligible(self._h, int(hole_index)) != -1

    def unrank(self, sparse):
        st = np.empty(self.N, np.int16)
        self._lib.ref_unrank(self._h, int(sparse), _ptr(st, _i16p))
        return st

    def rank(self, st):
        st = np.ascontiguousarray(st, np.int16)
        return int(self._lib.ref_rank(self._h, _ptr(st, _i16p)))

    def orbit(self, st, check_lowest):
        st = np.ascontiguousarray(st, np.int16)
        out = np.empty((self.N, self.N), np.int16)
        amp = np.empty(2 * self.N, np.float64)
        n = self._lib.ref_orbit(self._h, _ptr(st, _i16p), int(check_lowest), _ptr(out, _i16p), _ptr(amp, _dp))
        return out[:n].copy(), amp[: 2 * n].view(np.complex128).copy()

    def lowest(self, st):
        st = np.ascontiguousarray(st, np.int16)
        ind = np.zeros(1, np.int64)
        amp = np.zeros(2, np.float64)
        self._lib.ref_lowest(self._h, _ptr(st, _i16p), _ptr(ind, _i64p), _ptr(amp, _dp))
        return int(ind[0]), complex(amp[0], amp[1])


def ref_csr_dot(values, vinds, minor, major, x, y, weight, first, threads=None):
    R = ref()
    D = len(major) - 1
    R.ref_csr_dot(D, _ptr(major, _i64p), _ptr(vinds, _i16p), _ptr(minor, _i64p),
                  _ptr(values.view(np.float64), _dp), _ptr(x.view(np.float64), _dp), _ptr(y.view(np.float64), _dp),
                  float(np.real(weight)), float(np.imag(weight)), int(first), threads or default_threads())


def ref_csr_dot_rows(D, row_lo, values, vinds, minor, major, x, y, weight, threads=None):
    R = ref()
    R.ref_csr_dot_rows(D, row_lo, len(major) - 1, _ptr(major, _i64p), _ptr(vinds, _i16p), _ptr(minor, _i64p),
                       _ptr(values.view(np.float64), _dp), _ptr(x.view(np.float64), _dp),
                       _ptr(y.view(np.float64), _dp), float(np.real(weight)), float(np.imag(weight)),
                       threads or default_threads())
