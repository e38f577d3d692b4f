"""GPU symmetry sector: basis, lookup, matrix elements, H·v, operators, Lanczos.

A :class:`Sector` owns one ``ed_sector`` handle of the HIP library: the
sector (num_rungs, num_holes, num_down_spins, k) of the reference's
``Hamiltonian_Config`` (tJ_spin_half_ladder/configs.py:40-147) with its basis
built on the GPU.  Vectors are torch tensors on the current HIP device
(complex128, or complex64 for the reduced-precision form); torch provides
device memory and the stream, all arithmetic runs in the library's kernels.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from exactdiag_amd import _lib
from exactdiag_amd._lib import PREC_C64, PREC_C128, PREC_R32, PREC_R64, check, lib

TERMS = ("tl", "tr", "jl", "jr")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("exactdiag_amd needs a HIP GPU (torch.cuda.is_available() is False)")
    return torch


def _stream():
    return ct.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return ct.c_void_p(t.data_ptr())


def _dbl(arr):
    arr = np.ascontiguousarray(arr, np.float64)
    return arr, arr.ctypes.data_as(ct.POINTER(ct.c_double))


def weights_array(weights) -> np.ndarray:
    """Four complex term weights (tl, tr, jl, jr) as 8 doubles."""
    if isinstance(weights, dict):
        weights = [weights.get(t, 0.0) for t in TERMS]
    if isinstance(weights, np.ndarray) and weights.dtype == np.float64 and weights.shape == (8,):
        return weights.copy()  # already interleaved (re, im) pairs
    w = np.zeros(4, np.complex128)
    w[: len(weights)] = np.asarray(weights, np.complex128)
    return w.view(np.float64).copy()


class Sector:
    """Basis of one (L, h, nd, kx, ky) sector, built on the GPU."""

    def __init__(self, num_rungs: int, num_holes: int, num_down_spins: int, kx: int, ky: int, *, flip: int = 0,
                 _union=False):
        """``flip`` = +1 or -1 adds the global spin flip with that parity to the symmetry group (S^z = 0 sectors
        only; no counterpart in the reference, see ed_sector_create_flip); 0 keeps the reference's group."""
        self.L, self.h, self.nd, self.kx, self.ky = num_rungs, num_holes, num_down_spins, int(kx), int(ky)
        self.N = 2 * num_rungs
        self.any_k = bool(_union)
        self.flip = int(flip)
        self._lib = lib()
        torch = _torch()
        self.device = torch.device("cuda", torch.cuda.current_device())
        h = ct.c_void_p()
        if _union:
            check(self._lib.ed_sector_create_union(num_rungs, num_holes, num_down_spins, ct.byref(h)))
        elif self.flip:
            check(self._lib.ed_sector_create_flip(num_rungs, num_holes, num_down_spins, self.kx, self.ky, self.flip,
                                                  ct.byref(h)))
        else:
            check(self._lib.ed_sector_create(num_rungs, num_holes, num_down_spins, self.kx, self.ky, ct.byref(h)))
        self._h = h
        check(self._lib.ed_basis_build(self._h, _stream()))
        self.dim = int(self._lib.ed_basis_size(self._h))

    @classmethod
    def union(cls, num_rungs: int, num_holes: int, num_down_spins: int) -> "Sector":
        """Every lowest image of the ladder's states: the union of its momentum sectors
        (ed_sector_create_union), the row space of :class:`exactdiag_amd.multik.MultiK`."""
        return cls(num_rungs, num_holes, num_down_spins, 0, 0, _union=True)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.ed_sector_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def shape(self):
        return (self.dim, self.dim)

    # -- basis ---------------------------------------------------------
    def basis(self):
        """(dense_to_sparse int64, num_sparse_in_dense int32) as numpy arrays."""
        torch = _torch()
        sp = torch.empty(max(self.dim, 1), dtype=torch.int64, device=self.device)
        orb = torch.empty(max(self.dim, 1), dtype=torch.int32, device=self.device)
        check(self._lib.ed_basis_export(self._h, _ptr(sp), _ptr(orb), _stream()))
        torch.cuda.synchronize()
        return sp[: self.dim].cpu().numpy(), orb[: self.dim].cpu().numpy()

    def lookup(self, sparse):
        torch = _torch()
        s = torch.as_tensor(np.ascontiguousarray(sparse, np.int64), device=self.device)
        out = torch.empty_like(s)
        check(self._lib.ed_lookup(self._h, _ptr(s), _ptr(out), s.numel(), _stream()))
        return out.cpu().numpy()

    def lowest(self, sparse):
        torch = _torch()
        s = torch.as_tensor(np.ascontiguousarray(sparse, np.int64), device=self.device)
        rep = torch.empty_like(s)
        amp = torch.empty(s.numel(), dtype=torch.complex128, device=self.device)
        check(self._lib.ed_lowest(self._h, _ptr(s), _ptr(rep), _ptr(amp), s.numel(), _stream()))
        return rep.cpu().numpy(), amp.cpu().numpy()

    # -- matrix elements ---------------------------------------------------
    def _columns(self, fn, args, max_per_col, dim):
        torch = _torch()
        n = max(dim, 1)
        cnt = torch.zeros(n, dtype=torch.int32, device=self.device)
        rows = torch.empty(n * max_per_col, dtype=torch.int64, device=self.device)
        vals = torch.empty(n * max_per_col, dtype=torch.complex128, device=self.device)
        check(fn(*args, max_per_col, _ptr(cnt), _ptr(rows), _ptr(vals), _stream()))
        cnt = cnt[:dim].cpu().numpy()
        if dim and cnt.max(initial=0) > max_per_col:
            raise IndexError(f"max_values_per_column={max_per_col} set too low, {int(cnt.max())} found")
        rows = rows.view(-1, max_per_col)[:dim].cpu().numpy()
        vals = vals.view(-1, max_per_col)[:dim].cpu().numpy()
        ptr = np.zeros(dim + 1, np.int64)
        np.cumsum(cnt, out=ptr[1:])
        mask = np.arange(max_per_col)[None, :] < cnt[:, None]
        return ptr, rows[mask], vals[mask]

    def term_columns(self, term: int, triangle: bool = True, max_per_col: int | None = None):
        """Columns of one Hamiltonian term before de-duplication (CSC: ptr, rows, vals)."""
        mpc = max_per_col or (2 * self.N + 2)
        return self._columns(self._lib.ed_term_columns, (self._h, int(term), int(triangle)), mpc, self.dim)

    def op_columns(self, final: "Sector", op: int, qx: int, qy: int, max_per_col: int | None = None):
        mpc = max_per_col or (2 * self.N + 2)
        return self._columns(self._lib.ed_op_columns, (self._h, final._h, int(op), int(qx), int(qy)), mpc, self.dim)

    # -- products ------------------------------------------------------------
    def hv(self, weights, x, out=None, with_dot=False):
        """y = H x (matrix-free).  x: complex128 or complex64 tensor on the device."""
        torch = _torch()
        prec = 0 if x.dtype == torch.complex128 else 1
        if x.dtype not in (torch.complex128, torch.complex64):
            raise TypeError("x must be complex128 or complex64")
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        w, wp = _dbl(weights_array(weights))
        dot = torch.zeros(1, dtype=torch.complex128, device=self.device) if with_dot else None
        check(self._lib.ed_hv(self._h, wp, _ptr(x), _ptr(y), _ptr(dot) if dot is not None else None, prec,
                              _stream()))
        if with_dot:
            return y, complex(dot.cpu().numpy()[0])
        return y

    def lanczos_ritz(self, weights, v0, ritz, precision: int = 0):
        """Second pass of two-pass Lanczos: sum_n ritz[n] v_n (complex128) from the same start vector."""
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        r, rp = _dbl(ritz)
        x = torch.empty(max(self.dim, 1), dtype=torch.complex128, device=self.device)[: self.dim]
        w, wp = _dbl(weights_array(weights))
        check(self._lib.ed_lanczos_ritz(self._h, wp, _ptr(v0), len(r), rp, _ptr(x), int(precision), _stream()))
        return x

    def store(self, weights):
        """The stored Hamiltonian of this sector for ``weights`` (built once on the GPU)."""
        return StoredH(self, weights)

    def tune(self, weights, precision: int = 0):
        """Autotune this sector's H·v (tiles per workgroup, gathers in flight); returns (tiles, gather, ms)."""
        torch = _torch()
        dt = torch.complex128 if precision == 0 else torch.complex64
        x = torch.ones(max(self.dim, 1), dtype=dt, device=self.device)
        y = torch.empty_like(x)
        w, wp = _dbl(weights_array(weights))
        t, g, ms = ct.c_int(0), ct.c_int(0), ct.c_double(0.0)
        check(self._lib.ed_hv_tune(self._h, wp, int(precision), _ptr(x), _ptr(y), ct.byref(t), ct.byref(g),
                                   ct.byref(ms), _stream()))
        return t.value, g.value, ms.value

    # -- position correlations (tJ_spin_half_ladder/position_correlations.pyx) --------
    @staticmethod
    def _shifts(shifts):
        sh = np.ascontiguousarray(np.asarray(shifts, np.int32).reshape(-1, 2))
        return sh, len(sh)

    def position_columns(self, kind: int, shifts, max_per_col: int | None = None):
        """Columns of a position-correlation operator (kind 0 holes, 1 S^z, 2 singlet-singlet); shifts =
        the fixed distances then the free shift, as (leg, rung) pairs."""
        sh, n = self._shifts(shifts)
        mpc = max_per_col or (2 * self.N + 2)
        return self._columns(self._lib.ed_position_op_columns, (self._h, int(kind), sh.ctypes.data, n), mpc,
                             self.dim)

    def position_apply(self, kind: int, shifts, x):
        torch = _torch()
        sh, n = self._shifts(shifts)
        x = x.to(torch.complex128).contiguous()
        y = torch.empty(max(self.dim, 1), dtype=torch.complex128, device=self.device)[: self.dim]
        check(self._lib.ed_position_op_apply(self._h, int(kind), sh.ctypes.data, n, _ptr(x), _ptr(y), _stream()))
        return y

    def position_expectation(self, kind: int, shifts, psi) -> complex:
        """<psi|O|psi> of a position-correlation operator (fused: no matrix, no O psi vector)."""
        torch = _torch()
        sh, n = self._shifts(shifts)
        psi = psi.to(torch.complex128).contiguous()
        out = np.zeros(2, np.float64)
        check(self._lib.ed_position_expectation(self._h, int(kind), sh.ctypes.data, n, _ptr(psi),
                                                out.ctypes.data_as(ct.POINTER(ct.c_double)), _stream()))
        return complex(out[0], out[1])

    def position_correlations(self, kind: int, fixed, psi) -> np.ndarray:
        """Hole (kind 0) or S^z (kind 1) correlations of psi at every node index (2L values)."""
        torch = _torch()
        fx, n = self._shifts(fixed)
        psi = psi.to(torch.complex128).contiguous()
        out = np.zeros(self.N, np.float64)
        check(self._lib.ed_position_correlations(self._h, int(kind), fx.ctypes.data, n, _ptr(psi),
                                                 out.ctypes.data_as(ct.POINTER(ct.c_double)), _stream()))
        return out

    def op_apply(self, final: "Sector", op: int, qx: int, qy: int, x):
        torch = _torch()
        x = x.to(torch.complex128).contiguous()
        y = torch.empty(max(final.dim, 1), dtype=torch.complex128, device=self.device)[: final.dim]
        check(self._lib.ed_op_apply(self._h, final._h, int(op), int(qx), int(qy), _ptr(x), _ptr(y), _stream()))
        return y

    def lanczos(self, weights, v0, nsteps: int, keep_basis: bool = False, precision: int = 0):
        """Plain Lanczos coefficients (a, b) from v0; optionally the vectors (nsteps x D)."""
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        a = np.zeros(nsteps, np.float64)
        b = np.zeros(nsteps + 1, np.float64)
        n = ct.c_int(0)
        w, wp = _dbl(weights_array(weights))
        basis = None
        if keep_basis:
            dt = torch.complex128 if precision == 0 else torch.complex64
            basis = torch.empty((nsteps, self.dim), dtype=dt, device=self.device)
        check(self._lib.ed_lanczos_run(self._h, wp, _ptr(v0), int(nsteps), a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                       b.ctypes.data_as(ct.POINTER(ct.c_double)), ct.byref(n),
                                       _ptr(basis) if basis is not None else None, int(precision), _stream()))
        k = n.value
        if keep_basis:
            return a[:k].copy(), b[:k].copy(), basis[:k]
        return a[:k].copy(), b[:k].copy()


def real_phase(v, tol=1e-12):
    """v (a complex128 tensor on any device) times the unit phase that makes its largest element real and positive,
    or None when that leaves an imaginary part above tol of the norm.  The Lanczos coefficients and spectra
    do not depend on a global phase, so the rotated vector can run in real vectors."""
    import torch

    if v.numel() == 0:
        return None
    i = int(torch.argmax(v.abs()))
    p = v[i]
    if float(p.abs()) == 0.0:
        return None
    r = v * (p.conj() / p.abs())
    if float(torch.linalg.vector_norm(r.imag)) > tol * float(torch.linalg.vector_norm(r)):
        return None
    return r


class StoredH:
    """H of one sector and one set of weights stored on the device (sliced ELL, 6 bytes per element).

    The same products as :meth:`Sector.hv` and :meth:`Sector.lanczos`, from
    the stored elements instead of recomputing them; the sector must stay
    open while the matrix is used.
    """

    def __init__(self, sector: Sector, weights):
        self.sector = sector
        self.dim = sector.dim
        self.device = sector.device
        w, wp = _dbl(weights_array(weights))
        h = ct.c_void_p()
        check(lib().ed_hmat_create(sector.handle, wp, ct.byref(h), _stream()))
        self._h = h
        b, nnz, slots, nv = ct.c_int64(0), ct.c_int64(0), ct.c_int64(0), ct.c_int(0)
        check(lib().ed_hmat_info(self._h, ct.byref(b), ct.byref(nnz), ct.byref(slots), ct.byref(nv)))
        self.bytes, self.nnz, self.slots, self.n_values = b.value, nnz.value, slots.value, nv.value
        self.element_bytes = int(lib().ed_hmat_element_bytes(self._h))
        # every stored value real (momenta kx = 0, L/2): products and Lanczos runs may use real vectors
        self.is_real = bool(lib().ed_hmat_is_real(self._h))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().ed_hmat_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hv(self, x, out=None, with_dot=False):
        """y = H x for complex128 or complex64 x, or float64 / float32 x when :attr:`is_real`."""
        torch = _torch()
        precs = {torch.complex128: PREC_C128, torch.complex64: PREC_C64, torch.float64: PREC_R64,
                 torch.float32: PREC_R32}
        if x.dtype not in precs:
            raise TypeError("x must be complex128, complex64, float64 or float32")
        prec = precs[x.dtype]
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        dot = torch.zeros(1, dtype=torch.complex128, device=self.device) if with_dot else None
        check(lib().ed_hmat_hv(self._h, _ptr(x), _ptr(y), _ptr(dot) if dot is not None else None, prec, _stream()))
        if with_dot:
            return y, complex(dot.cpu().numpy()[0])
        return y

    def tune(self, precision: int = 0):
        """Autotune the stored product; returns (tiles per workgroup, gathers in flight, non-temporal, ms)."""
        torch = _torch()
        dt = {PREC_C128: torch.complex128, PREC_C64: torch.complex64, PREC_R64: torch.float64,
              PREC_R32: torch.float32}[precision]
        x = torch.ones(max(self.dim, 1), dtype=dt, device=self.device)
        y = torch.empty_like(x)
        t, u, nt, ms = ct.c_int(0), ct.c_int(0), ct.c_int(0), ct.c_double(0.0)
        check(lib().ed_hmat_tune(self._h, _ptr(x), _ptr(y), int(precision), ct.byref(t), ct.byref(u), ct.byref(nt),
                                 ct.byref(ms), _stream()))
        return t.value, u.value, nt.value, ms.value

    def set_config(self, tiles=0, unroll=4, nontemporal=-1):
        """tiles 0 = default; nontemporal -1 = default (non-temporal), 0 or 1 (bools accepted)."""
        check(lib().ed_hmat_set_config(self._h, int(tiles), int(unroll), int(nontemporal)))

    def lanczos(self, v0, nsteps: int, precision: int = 0):
        """Plain Lanczos coefficients from v0 in vectors of `precision` (PREC_*).  The real precisions need
        :attr:`is_real` and a real v0 (see :func:`real_phase`)."""
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        a = np.zeros(nsteps, np.float64)
        b = np.zeros(nsteps + 1, np.float64)
        n = ct.c_int(0)
        check(lib().ed_hmat_lanczos_run(self._h, _ptr(v0), int(nsteps), a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                        b.ctypes.data_as(ct.POINTER(ct.c_double)), ct.byref(n), int(precision),
                                        _stream()))
        k = n.value
        return a[:k].copy(), b[:k].copy()

    def lanczos_ritz(self, v0, ritz, precision: int = 0):
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        r, rp = _dbl(ritz)
        x = torch.empty(max(self.dim, 1), dtype=torch.complex128, device=self.device)[: self.dim]
        check(lib().ed_hmat_lanczos_ritz(self._h, _ptr(v0), len(r), rp, _ptr(x), int(precision), _stream()))
        return x

    def lanczos_basis(self, v0, nsteps: int, precision: int = 0, ld: int | None = None):
        """(a, b, V): the coefficients of :meth:`lanczos` and the normalised Lanczos vectors as the first rows of V,
        a device tensor (nsteps, ld) of the precision's vector type (ld >= dim; columns past dim stay zero)."""
        torch = _torch()
        dt = {PREC_C128: torch.complex128, PREC_C64: torch.complex64, PREC_R64: torch.float64,
              PREC_R32: torch.float32}[precision]
        ld = self.dim if ld is None else int(ld)
        v0 = v0.to(torch.complex128).contiguous()
        V = torch.zeros((nsteps, ld), dtype=dt, device=self.device)
        a = np.zeros(nsteps, np.float64)
        b = np.zeros(nsteps + 1, np.float64)
        n = ct.c_int(0)
        check(lib().ed_hmat_lanczos_run_basis(self._h, _ptr(v0), int(nsteps), a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                              b.ctypes.data_as(ct.POINTER(ct.c_double)), ct.byref(n), _ptr(V), ld,
                                              int(precision), _stream()))
        k = n.value
        return a[:k].copy(), b[:k].copy(), V

    def lanczos_to_tol(self, v0, max_steps: int, rule: "StopRule", precision: int = 0):
        """Lanczos coefficients from v0 until the terminated spectrum changes by less than rule.tol over
        rule.delta steps (ed_hmat_lanczos_run_tol).  Returns (a, b, reason): reason 0 = max_steps reached,
        1 = the rule was met, 2 = the Krylov space was exhausted (the fraction is exact, no terminator)."""
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        a = np.zeros(max_steps, np.float64)
        b = np.zeros(max_steps + 1, np.float64)
        n, why = ct.c_int(0), ct.c_int(0)
        check(lib().ed_hmat_lanczos_run_tol(self._h, _ptr(v0), int(max_steps), ct.byref(rule.c),
                                            a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                            b.ctypes.data_as(ct.POINTER(ct.c_double)), ct.byref(n), ct.byref(why),
                                            int(precision), _stream()))
        k = n.value
        return a[:k].copy(), b[:k].copy(), why.value


class StopRule:
    """The stop rule of :meth:`StoredH.lanczos_to_tol` (struct ed_lanczos_stop): from min_steps on, every
    check_every steps, the spectrum of the first n levels closed by the square-root terminator is compared with
    those of the first n - delta and n - delta // 2 on every stride-th point of ws (z = w + e0 + i eta); the
    run stops when the larger relative L1 distance falls below tol."""

    def __init__(self, ws, e0, eta, tol, min_steps=60, check_every=5, delta=20, max_points=1024):
        from exactdiag_amd._lib import LanczosStop

        self.ws = np.ascontiguousarray(ws, np.float64)
        stride = max(1, -(-len(self.ws) // max_points))
        self.c = LanczosStop(self.ws.ctypes.data_as(ct.POINTER(ct.c_double)), len(self.ws), stride, float(e0),
                             float(eta), float(tol), int(min_steps), int(check_every), int(delta))

    def change(self, a, b, n=None):
        """The rule's estimate for the first n coefficients (host only, ed_spectrum_change)."""
        n = len(a) if n is None else n
        a = np.ascontiguousarray(a[:n], np.float64)
        b = np.ascontiguousarray(b[:n], np.float64)
        out = ct.c_double(0.0)
        check(lib().ed_spectrum_change(a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                       b.ctypes.data_as(ct.POINTER(ct.c_double)), int(n), int(self.c.delta),
                                       ct.byref(self.c), ct.byref(out)))
        return out.value


def cf_eval(a, b, ws, e0, eta, factor, tail=None):
    """factor * Im(1/det(z)) for z = ws + e0 + i*eta on the GPU.

    ``tail=(a_inf, b_inf)`` closes the fraction with the square-root terminator.
    """
    a, ap = _dbl(a)
    b, bp = _dbl(b)
    ws, wsp = _dbl(ws)
    out = np.empty(len(ws), np.float64)
    a_inf, b_inf = (0.0, 0.0) if tail is None else (float(tail[0]), float(tail[1]))
    check(lib().ed_cf_eval_tail(ap, bp, len(a), wsp, len(ws), float(e0), float(eta), float(factor), a_inf, b_inf,
                                out.ctypes.data_as(ct.POINTER(ct.c_double))))
    return out
