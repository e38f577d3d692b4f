"""Batched Lanczos: many sectors (all k sectors, all q of a spectrum) at once.

The reference loops over momentum sectors (tJ_spin_half_ladder/api.py:45-59)
and over q (api.py:116-161), building matrices and running one Lanczos per
sector.  A :class:`Batch` hands every sector to the library's batched engine
(``ed_batch_*`` in include/exactdiag_hip.h): one pool of vectors, one
launch per kernel for all sectors, the whole recurrence in one HIP graph.

* :func:`lowest_eigenpairs` -- lowest eigenpair of every instance by
  restarted two-pass Lanczos (pass one: coefficients; pass two: the Ritz
  vector accumulated from the same recurrence), restarted from the Ritz
  vector until ||H x - E x|| <= tol * max(1, |E|).
* :func:`spectra` -- the continued-fraction spectra of many (operator, q)
  excitations of one ground state, all Lanczos runs in one batch.
"""

from __future__ import annotations

import ctypes as ct
import dataclasses

import numpy as np
from scipy.linalg import eigh_tridiagonal

from exactdiag_amd._lib import check, lib
from exactdiag_amd.sector import Sector, _dbl, _ptr, _stream, _torch, cf_eval, weights_array


class Batch:
    """A set of (sector, weights) instances sharing one vector pool."""

    def __init__(self, sectors, weights, stored=False):
        """``stored``: every instance's H·v reads a stored H (built here, one per distinct sector and
        weights) instead of recomputing its elements."""
        torch = _torch()
        self.sectors = list(sectors)
        n = len(self.sectors)
        if n == 0:
            raise ValueError("a batch needs at least one sector")
        per_instance = (isinstance(weights, (list, tuple)) and len(weights) == n
                        and all(isinstance(w, (dict, list, tuple, np.ndarray)) for w in weights))
        wl = [weights_array(w) for w in weights] if per_instance else [weights_array(weights)] * n
        self._w = np.ascontiguousarray(np.concatenate(wl), np.float64)
        h = ct.c_void_p()
        self._stored = []
        if stored:
            built = {}
            for s, w in zip(self.sectors, wl):
                key = (id(s), w.tobytes())
                if key not in built:
                    built[key] = s.store(w)
                self._stored.append(built[key])
            mats = (ct.c_void_p * n)(*[m._h.value for m in self._stored])
            check(lib().ed_batch_create_stored(mats, n, ct.byref(h)))
        else:
            handles = (ct.c_void_p * n)(*[s.handle.value for s in self.sectors])
            check(lib().ed_batch_create(handles, self._w.ctypes.data_as(ct.POINTER(ct.c_double)), n, ct.byref(h)))
        self._h = h
        off = np.zeros(n + 1, np.int64)
        check(lib().ed_batch_offsets(self._h, off.ctypes.data_as(ct.POINTER(ct.c_int64))))
        self.offsets = off
        self.total = int(off[-1])
        self.dims = np.diff(off)
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __len__(self):
        return len(self.sectors)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().ed_batch_destroy(h)
            self._h = None
        for m in getattr(self, "_stored", []):
            m.close()
        self._stored = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- pool helpers ----------------------------------------------------------
    def empty(self):
        torch = _torch()
        return torch.empty(max(self.total, 1), dtype=torch.complex128, device=self.device)[: self.total]

    def zeros(self):
        torch = _torch()
        return torch.zeros(max(self.total, 1), dtype=torch.complex128, device=self.device)[: self.total]

    def part(self, pool, i):
        return pool[int(self.offsets[i]): int(self.offsets[i + 1])]

    def pack(self, vectors):
        pool = self.empty()
        for i, v in enumerate(vectors):
            self.part(pool, i).copy_(v)
        return pool

    def random(self, seed=0):
        """Seeded complex normal start vectors, one stream of numbers per instance."""
        torch = _torch()
        pool = self.empty()
        for i in range(len(self)):
            rng = np.random.default_rng([seed, i])
            d = int(self.dims[i])
            v = rng.normal(size=d) + 1j * rng.normal(size=d)
            self.part(pool, i).copy_(torch.as_tensor(v))
        return pool

    # -- engine ------------------------------------------------------------------
    def config(self):
        t, g, nj = ct.c_int(0), ct.c_int(0), ct.c_int(0)
        check(lib().ed_batch_get_config(self._h, ct.byref(t), ct.byref(g), ct.byref(nj)))
        return t.value, g.value, nj.value

    def set_config(self, tiles, gather):
        check(lib().ed_batch_set_config(self._h, int(tiles), int(gather)))

    def tune(self):
        """Autotune the batched H·v; returns (tiles, gather, ms)."""
        x = self.random(seed=12345)
        y = self.empty()
        t, g, ms = ct.c_int(0), ct.c_int(0), ct.c_double(0.0)
        check(lib().ed_batch_tune(self._h, _ptr(x), _ptr(y), ct.byref(t), ct.byref(g), ct.byref(ms), _stream()))
        return t.value, g.value, ms.value

    def hv(self, x, out=None, with_dots=False):
        y = self.empty() if out is None else out
        dots = np.zeros(2 * len(self), np.float64) if with_dots else None
        check(lib().ed_batch_hv(self._h, _ptr(x), _ptr(y),
                                dots.ctypes.data_as(ct.POINTER(ct.c_double)) if with_dots else None, _stream()))
        if with_dots:
            return y, dots.view(np.complex128)
        return y

    def lanczos(self, v0, nsteps, ritz=None, graph=True):
        """Lanczos on every instance from the pool v0.

        Returns (a list, b list, acc): a[i] has n_i values, b[i] has n_i + 1
        (b[i][0] = 0, b[i][n_i] the norm of the last residual).  With ``ritz``
        (list of coefficient arrays) acc is the pool sum_n ritz[i][n] v_n.
        """
        n = len(self)
        steps = np.broadcast_to(np.asarray(nsteps, np.int32), (n,)).copy()
        ld = max(int(steps.max()), 1)
        a = np.zeros(n * ld, np.float64)
        b = np.zeros(n * (ld + 1), np.float64)
        cnt = np.zeros(n, np.int32)
        acc = None
        rz = None
        if ritz is not None:
            rz = np.zeros((n, ld), np.float64)
            for i, r in enumerate(ritz):
                r = np.asarray(r, np.float64)[:ld]
                rz[i, : len(r)] = r
            rz = np.ascontiguousarray(rz.reshape(-1))
            acc = self.empty()
        check(lib().ed_batch_lanczos(self._h, _ptr(v0), steps.ctypes.data_as(ct.POINTER(ct.c_int)), ld,
                                     rz.ctypes.data_as(ct.POINTER(ct.c_double)) if rz is not None else None,
                                     _ptr(acc) if acc is not None else None,
                                     a.ctypes.data_as(ct.POINTER(ct.c_double)),
                                     b.ctypes.data_as(ct.POINTER(ct.c_double)),
                                     cnt.ctypes.data_as(ct.POINTER(ct.c_int)), int(bool(graph)), _stream()))
        a = a.reshape(n, ld)
        b = b.reshape(n, ld + 1)
        return ([a[i, : cnt[i]].copy() for i in range(n)], [b[i, : cnt[i] + 1].copy() for i in range(n)], acc)


def _norm(v):
    from exactdiag_amd.general.krylov import vec_norm

    return vec_norm(v) if v.numel() else 0.0


def lowest_eigenpairs(batch: Batch, tol=1e-10, nsteps=120, max_restarts=40, seed=0, graph=True):
    """Lowest eigenpair of every instance: (energies, pool of unit eigenvectors, restarts).

    Restarted two-pass Lanczos; an instance stops restarting when
    ||H x - E x|| <= tol * max(1, |E|), E the Rayleigh quotient.
    """
    from exactdiag_amd.general.krylov import vec_axpys, vec_scale

    n = len(batch)
    v = batch.random(seed)
    x = batch.zeros()
    energies = np.full(n, np.nan)
    done = batch.dims == 0
    energies[done] = np.nan
    for restart in range(max_restarts):
        steps = np.where(done, 0, np.minimum(nsteps, batch.dims)).astype(np.int32)
        a, b, _ = batch.lanczos(v, steps, graph=graph)
        ritz = []
        for i in range(n):
            c = len(a[i])
            if done[i] or c == 0:
                ritz.append(np.zeros(0))
                continue
            if c == 1:
                ritz.append(np.ones(1))
                continue
            _, y = eigh_tridiagonal(a[i], b[i][1:c], select="i", select_range=(0, 0))
            ritz.append(y[:, 0])
        _, _, acc = batch.lanczos(v, np.array([len(r) for r in ritz], np.int32), ritz=ritz, graph=graph)
        for i in range(n):
            if done[i] or len(ritz[i]) == 0:
                continue
            xi = batch.part(x, i)
            xi.copy_(batch.part(acc, i))
            vec_scale(xi, 1.0 / _norm(xi))
        y, dots = batch.hv(x, with_dots=True)
        for i in range(n):
            if done[i] or len(ritz[i]) == 0:
                continue
            e = dots[i].real
            energies[i] = e
            r = batch.part(y, i).clone()
            vec_axpys(batch.part(x, i).view(1, -1), 1, [e], r)
            if _norm(r) <= tol * max(1.0, abs(e)) or len(a[i]) < steps[i]:
                done[i] = True  # converged, or an invariant subspace was reached (b < 1e-8)
        if np.all(done):
            return energies, x, restart
        v = x.clone()
    raise RuntimeError(f"batched Lanczos: {int(np.sum(~done))} instances did not converge")


@dataclasses.dataclass
class Excitation:
    """One Lanczos instance of :func:`spectra`: operator ``op`` with momentum (qx, qy)."""

    name: str
    op: int
    qx: int
    qy: int
    sector: Sector
    norm2: float = 0.0


def spectra(excitations, gs_sector, psi, e0, weights, ws, eta, nsteps, graph=True, stored=False):
    """All excitations' spectra from one ground state in one batched Lanczos run.

    Follows lanczos_dynamical_response (general/cython_lanczos_diagonalization.pyx:18-40)
    per excitation: phi = O psi, Lanczos from phi, S(w) = -||phi||^2/pi Im 1/(z - a_0 - ...).
    Returns (list of spectra, list of (a, b)).
    """
    from exactdiag_amd.general.krylov import vec_norm

    phis = []
    for ex in excitations:
        phi = gs_sector.op_apply(ex.sector, ex.op, ex.qx, ex.qy, psi)
        ex.norm2 = vec_norm(phi) ** 2 if ex.sector.dim else 0.0
        phis.append(phi)
    live = [i for i, ex in enumerate(excitations) if ex.norm2 >= 1e-16]
    out = [np.zeros(len(ws)) for _ in excitations]
    coeffs = [(np.zeros(0), np.zeros(0)) for _ in excitations]
    if not live:
        return out, coeffs
    batch = Batch([excitations[i].sector for i in live], weights, stored=stored)
    try:
        v0 = batch.pack([phis[i] for i in live])
        a, b, _ = batch.lanczos(v0, nsteps, graph=graph)
    finally:
        batch.close()
    for k, i in enumerate(live):
        ak, bk = a[k], b[k][: len(a[k])]
        coeffs[i] = (ak, bk)
        out[i] = cf_eval(ak, bk, ws, e0, eta, -excitations[i].norm2 / np.pi)
    return out, coeffs
