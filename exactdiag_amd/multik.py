"""Several momentum sectors of one ladder in one stored product (``ed_mk_*``).

The reference computes all-q spectra one q at a time, each with its own
final-sector Hamiltonian (tJ_spin_half_ladder/api.py:116-132 calling
get_spectrum, general/lanczos_diagonalization.py:124-143).  The sectors of one
ladder share their representatives and differ only in the momentum phase of
each element, so :class:`MultiK` stores H once over the union of the sectors
(:meth:`Sector.union`) and applies it to up to 16 sectors in one pass: the
matrix is read once for all of them and each element's gathers of the 16
slots fall in one contiguous line.

Blocks are complex tensors of shape (D_union, 16) in scaled coordinates
x'_f = x_f / sqrt(n_f) (n_f the orbit size); :meth:`attach` and
:meth:`extract` convert from and to a sector's own vectors.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from exactdiag_amd._lib import check, lib
from exactdiag_amd.sector import Sector, _dbl, _ptr, _stream, _torch, weights_array

SLOTS = 16


def _per_slot_weights(weights, n):
    """One set of four term weights for every slot, or a list of n sets."""
    if isinstance(weights, (list, tuple)) and len(weights) == n and all(
            isinstance(w, (list, tuple, np.ndarray, dict)) for w in weights):
        return np.concatenate([weights_array(w) for w in weights])
    return np.tile(weights_array(weights), n)


class MultiK:
    """H of up to 16 momentum sectors (kx, ky) of one ladder, stored once over the union basis.

    ``weights``: the four term weights (tl, tr, jl, jr) for every slot, or one set per slot."""

    def __init__(self, union: Sector, weights, momenta):
        if not getattr(union, "any_k", False):
            raise ValueError("MultiK needs the union sector (Sector.union)")
        momenta = [(int(kx), int(ky)) for kx, ky in momenta]
        if not 1 <= len(momenta) <= SLOTS:
            raise ValueError(f"1..{SLOTS} momentum slots required")
        self.union = union
        self.momenta = momenta
        self.dim = union.dim
        self.device = union.device
        n = len(momenta)
        kx = (ct.c_int * n)(*[m[0] for m in momenta])
        ky = (ct.c_int * n)(*[m[1] for m in momenta])
        _, wp = _dbl(_per_slot_weights(weights, n))
        h = ct.c_void_p()
        check(lib().ed_mk_create(union.handle, wp, n, kx, ky, ct.byref(h), _stream()))
        self._h = h
        d, nnz, b = ct.c_int64(0), ct.c_int64(0), ct.c_int64(0)
        check(lib().ed_mk_info(self._h, ct.byref(d), ct.byref(nnz), ct.byref(b)))
        self.nnz, self.bytes = nnz.value, b.value

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().ed_mk_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_config(self, waves: int = 8, unroll: int = 4):
        """Launch shape of the product: 4 or 8 wavefronts per workgroup, 4 or 8 gathers in flight per lane."""
        check(lib().ed_mk_set_config(self._h, int(waves), int(unroll)))

    def block(self, dtype=None):
        """A zero block (D_union, 16) on the device (complex128 unless dtype says otherwise)."""
        torch = _torch()
        return torch.zeros((max(self.dim, 1), SLOTS), dtype=dtype or torch.complex128, device=self.device)

    def attach(self, slot: int, sector: Sector, phi, block):
        """Write the sector vector phi (complex128, the sector's dense order) into slot `slot` of block."""
        torch = _torch()
        phi = phi.to(torch.complex128).contiguous()
        check(lib().ed_mk_attach(self._h, int(slot), sector.handle, _ptr(phi), _ptr(block), _stream()))

    def extract(self, slot: int, sector: Sector, block):
        """Slot `slot` of a complex128 block as a vector of the sector (its dense order)."""
        torch = _torch()
        out = torch.empty(max(sector.dim, 1), dtype=torch.complex128, device=self.device)[: sector.dim]
        check(lib().ed_mk_extract(self._h, int(slot), sector.handle, _ptr(block), _ptr(out), _stream()))
        return out

    def hv(self, x, out=None, with_dot=False):
        """y = H x for every slot of the block x (complex128 or complex64); with_dot also <x_k, H x_k> per slot."""
        torch = _torch()
        if x.dtype not in (torch.complex128, torch.complex64):
            raise TypeError("x must be complex128 or complex64")
        prec = 0 if x.dtype == torch.complex128 else 1
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        dot = torch.zeros(SLOTS, dtype=torch.complex128, device=self.device) if with_dot else None
        check(lib().ed_mk_hv(self._h, _ptr(x), _ptr(y), _ptr(dot) if dot is not None else None, prec, _stream()))
        if with_dot:
            return y, dot.cpu().numpy()[: len(self.momenta)]
        return y

    def lanczos(self, v0, nsteps: int, precision: int = 0):
        """Plain Lanczos from the complex128 block v0 for every slot: a list of (a, b) per slot, as
        :meth:`StoredH.lanczos` returns them for one sector."""
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        n = len(self.momenta)
        a = np.zeros((n, nsteps), np.float64)
        b = np.zeros((n, nsteps + 1), np.float64)
        cnt = (ct.c_int * n)()
        check(lib().ed_mk_lanczos_run(self._h, _ptr(v0), int(nsteps), a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                      b.ctypes.data_as(ct.POINTER(ct.c_double)), cnt, int(precision), _stream()))
        return [(a[k, : cnt[k]].copy(), b[k, : cnt[k]].copy()) for k in range(n)]

    def lanczos_ritz(self, v0, ritz, precision: int = 0):
        """Second pass of two-pass Lanczos: the complex128 block sum_n ritz[k][n] v_n(k) for every slot (ritz: one
        coefficient list per slot, zero-padded to a common length)."""
        torch = _torch()
        v0 = v0.to(torch.complex128).contiguous()
        n = len(self.momenta)
        nsteps = max(1, max(len(r) for r in ritz))
        r = np.zeros((n, nsteps), np.float64)
        for k, rk in enumerate(ritz):
            r[k, : len(rk)] = rk
        x = self.block()
        check(lib().ed_mk_lanczos_ritz(self._h, _ptr(v0), nsteps, r.ctypes.data_as(ct.POINTER(ct.c_double)), _ptr(x),
                                       int(precision), _stream()))
        return x

    def residual_norms2(self, x, hx=None, e=None):
        """sum_f n_f |hx_k - e_k x_k|^2 per slot (complex128 blocks); without hx the squared norms of x."""
        n = len(self.momenta)
        out = np.zeros(n, np.float64)
        ev = np.ascontiguousarray(e if e is not None else np.zeros(n), np.float64)
        check(lib().ed_mk_residual_norms(self._h, _ptr(x), _ptr(hx) if hx is not None else None,
                                         ev.ctypes.data_as(ct.POINTER(ct.c_double)),
                                         out.ctypes.data_as(ct.POINTER(ct.c_double)), _stream()))
        return out

    def scale(self, x, s):
        """Slot k of the complex128 block x times s[k], in place."""
        sv = np.ascontiguousarray(s, np.float64)
        check(lib().ed_mk_scale(self._h, _ptr(x), sv.ctypes.data_as(ct.POINTER(ct.c_double)), _stream()))
        return x

    def lowest(self, v0, tol=1e-10, nsteps=120, max_restarts=30):
        """Lowest eigenpair of every slot's sector by restarted two-pass Lanczos, all slots in one recurrence.

        Pass one gives each slot's coefficients; the shortest leading tridiagonal whose Ritz residual estimate
        meets the tolerance gives its lowest eigenvector; pass two accumulates the Ritz vectors in FP64; restart
        from them until every slot has ||H x - E x|| <= tol * max(1, |E|).  A slot whose recurrence stops before
        nsteps (b < 1e-8, the reference's threshold: its Krylov space is invariant to that precision) is final
        after that pass.  Slots whose start vector is zero are skipped.  Returns (energies, block of unit
        eigenvectors, restarts, residuals)."""
        from exactdiag_amd.forms import _shortest_converged_prefix

        n = len(self.momenta)
        live = self.residual_norms2(v0) > 0.0
        v = v0
        e = np.full(n, np.nan)
        res = np.full(n, np.inf)
        x = v0
        for restart in range(max_restarts):
            coeffs = self.lanczos(v, nsteps)
            exhausted = [len(a) < nsteps for a, _ in coeffs]
            ritz = [(_shortest_converged_prefix(a, b, tol) if (live[k] and len(a)) else np.zeros(0))
                    for k, (a, b) in enumerate(coeffs)]
            x = self.lanczos_ritz(v, ritz)
            nrm = np.sqrt(self.residual_norms2(x))
            self.scale(x, np.where(nrm > 0, 1.0 / np.where(nrm > 0, nrm, 1.0), 0.0))
            hx, dots = self.hv(x, with_dot=True)
            e = np.where(live, dots.real, np.nan)
            res = np.sqrt(self.residual_norms2(x, hx, np.where(live, dots.real, 0.0)))
            done = [(not live[k]) or exhausted[k] or res[k] <= tol * max(1.0, abs(e[k])) for k in range(n)]
            del hx
            if all(done):
                return e, x, restart, res
            v = x
        return e, x, max_restarts, res

