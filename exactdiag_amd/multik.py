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


class MultiK:
    """H of up to 16 momentum sectors (kx, ky) of one ladder, stored once over the union basis."""

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
        _, wp = _dbl(weights_array(weights))
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
