"""All-q dynamical spin structure factor S(q, w) on one GPU.

The reference computes all-q S(q, w) by looping over q and calling
``get_spectrum`` for each (tJ_spin_half_ladder/api.py:116-132): one ground
state, then per q a final-sector basis, the Szq operator, a plain Lanczos
run and the continued fraction.  This module runs the same sequence with
every step on the device and without disk caches, and measures it:

* ``mode="plain"``: FP64 vectors, the configured number of Lanczos steps,
  one q after the other (the plain FP64 GPU Lanczos of the benchmark);
* ``mode="fast"``: FP32 vectors with FP64 coefficient reductions, fewer
  Lanczos steps and a continued-fraction tail (``closure``) in place of the
  missing levels, which is the time-to-accuracy form.

Spectral L1 error = sum_w |S - S_ref| dw / sum_w |S_ref| dw on the same grid.
"""

from __future__ import annotations

import dataclasses
import time

import numpy as np

from exactdiag_amd import _lib
from exactdiag_amd.general import krylov
from exactdiag_amd.sector import Sector, cf_eval


def final_momentum(L, kx, ky, qx, qy):
    """tJ_spin_half_ladder/matrix_setup.pyx:329-330."""
    half = (L - 1) // 2
    return (kx + qx + half) % L - half, (ky + qy) % 2


def distinct_qs(L, limited=True):
    """(qx, qy) pairs computed; with ``limited`` qx > L//2 reuses L - qx (api.py:126-131)."""
    qxs = range(L // 2 + 1) if limited else range(L)
    return [(qx, qy) for qx in qxs for qy in (0, 1)]


def spectral_l1(S, S_ref, ws):
    """sum |S - S_ref| dw / sum |S_ref| dw (0 when both vanish, inf when only S_ref does)."""
    dw = np.gradient(np.asarray(ws))
    num = float(np.sum(np.abs(S - S_ref) * dw))
    den = float(np.sum(np.abs(S_ref) * dw))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return num / den


@dataclasses.dataclass
class AllQResult:
    ws: np.ndarray
    qs: list
    spectra: dict          # (qx, qy) -> S(w)
    coefficients: dict     # (qx, qy) -> (a, b, norm2)
    e0: float
    timings: dict


def ground_state(L, h, nd, kx, ky, weights, tol=1e-12, seed=0):
    import torch

    s = Sector(L, h, nd, kx, ky)
    w = np.asarray(weights, np.complex128)
    from exactdiag_amd.sector import weights_array

    wa = weights_array(w)

    def apply(v):
        return s.hv(wa, v)

    vals, vecs, _ = krylov.lowest_eigenpairs(apply, s.dim, 1, tol=tol, v0=krylov.start_vector(s.dim, seed))
    torch.cuda.synchronize()
    return s, float(vals[0]), vecs[0]


def run_allq(L, h, nd, weights, ws, eta, nsteps, k0=(0, 0), precision=0, limited=True, closure=None, gs=None,
             keep_sectors=False):
    """All-q S(q, w) for the Szq operator.  Returns AllQResult with timings in seconds."""
    import torch

    timings = {}
    t0 = time.perf_counter()
    if gs is None:
        s0, e0, psi = ground_state(L, h, nd, k0[0], k0[1], weights)
    else:
        s0, e0, psi = gs
    torch.cuda.synchronize()
    timings["ground_state"] = time.perf_counter() - t0
    from exactdiag_amd.sector import weights_array

    wa = weights_array(np.asarray(weights, np.complex128))
    spectra, coeffs = {}, {}
    t_basis = t_lz = t_cf = t_op = 0.0
    sectors = {}
    for qx, qy in distinct_qs(L, limited):
        t1 = time.perf_counter()
        kx, ky = final_momentum(L, k0[0], k0[1], qx, qy)
        sec = sectors.get((kx, ky))
        if sec is None:
            sec = Sector(L, h, nd, kx, ky)
            if keep_sectors:
                sectors[(kx, ky)] = sec
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        phi = s0.op_apply(sec, _lib.OP_SZQ, qx, qy, psi)
        norm2 = krylov.vec_norm(phi) ** 2 if sec.dim else 0.0
        t3 = time.perf_counter()
        if norm2 < 1e-16:
            spectra[(qx, qy)] = np.zeros(len(ws))
            coeffs[(qx, qy)] = (np.zeros(0), np.zeros(0), 0.0)
            t_basis += t2 - t1
            t_op += t3 - t2
            continue
        a, b = sec.lanczos(wa, phi, nsteps, precision=precision)
        t4 = time.perf_counter()
        if closure is not None:
            a, b = closure(a, b)
        spectra[(qx, qy)] = cf_eval(a, b, ws, e0, eta, -norm2 / np.pi)
        coeffs[(qx, qy)] = (a, b, norm2)
        t5 = time.perf_counter()
        t_basis += t2 - t1
        t_op += t3 - t2
        t_lz += t4 - t3
        t_cf += t5 - t4
        if not keep_sectors:
            sec.close()
    timings.update(final_basis=t_basis, operator=t_op, lanczos=t_lz, continued_fraction=t_cf,
                   total=time.perf_counter() - t0)
    return AllQResult(np.asarray(ws), distinct_qs(L, limited), spectra, coeffs, e0, timings)


# -- time to all-q S(q, w) at a given accuracy ----------------------------------

REFERENCE_EIG_TOL = 2e-7  # the reference's ARPACK tolerance (general/lanczos_diagonalization.py:89)
GS_STEPS = 100  # first-pass length of the fast forms' two-pass ground state (restarts cover slower systems)
LAST_FAST_TIMINGS = {}  # seconds per stage of the last allq_fast run
LAST_FAST_STEPS = {}  # Lanczos steps per q of the last allq_fast run


def _sync():
    import torch

    torch.cuda.synchronize()


def _gs(apply, dim, tol, seed=0):
    vals, vecs, _ = krylov.lowest_eigenpairs(apply, dim, 1, tol=tol, v0=krylov.device_start_vector(dim, seed))
    return float(vals[0]), vecs[0]


def allq_reference(L, h, nd, weights, ws, eta, nsteps=1000, k0=(0, 0), limited=True, log=None):
    """S_ref(q, w) for every distinct q: ground state to 1e-12, then nsteps FP64 Lanczos steps (stored H).
    ``log`` (optional) receives one progress line per q."""
    s0 = Sector(L, h, nd, k0[0], k0[1])
    hm0 = s0.store(weights)
    e0, psi = _gs(hm0.hv, s0.dim, 1e-12)
    hm0.close()
    if log is not None:
        log(f"  reference ground state E0 = {e0:.12f}")
    out = {}
    for qx, qy in distinct_qs(L, limited):
        kx, ky = final_momentum(L, k0[0], k0[1], qx, qy)
        sec = Sector(L, h, nd, kx, ky)
        phi = s0.op_apply(sec, _lib.OP_SZQ, qx, qy, psi)
        n2 = krylov.vec_norm(phi) ** 2 if sec.dim else 0.0
        if n2 < 1e-16:
            out[(qx, qy)] = np.zeros(len(ws))
        else:
            hm = sec.store(weights)
            a, b = hm.lanczos(phi, nsteps)
            out[(qx, qy)] = cf_eval(a, b, ws, e0, eta, -n2 / np.pi)
            hm.close()
        sec.close()
        if log is not None:
            log(f"  reference spectrum q=({qx},{qy}) done")
    s0.close()
    _sync()
    return e0, out


def allq_plain(L, h, nd, weights, ws, eta, nsteps, k0=(0, 0), limited=True, tol=REFERENCE_EIG_TOL):
    """The plain method on the GPU: FP64 vectors, matrix-free H·v, truncated continued fraction, one q
    after the other (tJ_spin_half_ladder/api.py:116-132 with get_spectrum per q).  Returns (seconds, spectra)."""
    from exactdiag_amd.sector import weights_array

    wa = weights_array(np.asarray(weights, np.complex128))
    _sync()
    t0 = time.perf_counter()
    s0 = Sector(L, h, nd, k0[0], k0[1])
    e0, psi = _gs(lambda v: s0.hv(wa, v), s0.dim, tol)
    out = {}
    for qx, qy in distinct_qs(L, limited):
        kx, ky = final_momentum(L, k0[0], k0[1], qx, qy)
        sec = Sector(L, h, nd, kx, ky)
        phi = s0.op_apply(sec, _lib.OP_SZQ, qx, qy, psi)
        n2 = krylov.vec_norm(phi) ** 2 if sec.dim else 0.0
        if n2 < 1e-16:
            out[(qx, qy)] = np.zeros(len(ws))
        else:
            a, b = sec.lanczos(wa, phi, nsteps)
            out[(qx, qy)] = cf_eval(a, b, ws, e0, eta, -n2 / np.pi)
        sec.close()
    s0.close()
    _sync()
    return time.perf_counter() - t0, out


def allq_fast(L, h, nd, weights, ws, eta, nsteps, k0=(0, 0), limited=True, tol=REFERENCE_EIG_TOL, precision=1,
              terminator=True, two_pass_gs=True, real=True, spectrum_tol=None):
    """The fast form: stored H (built once per sector on the GPU), complex64 Lanczos vectors with FP64
    coefficient reductions, fewer steps closed by the square-root terminator; the ground state by restarted
    two-pass Lanczos in complex128 (no re-orthogonalisation, the same residual criterion as the plain run).
    With real=True, sectors whose stored H is real (kx = 0, L/2) run in real vectors of the same precision
    (float64 for the ground state, float32 for the spectra) when the start vector is real up to a global
    phase.  With spectrum_tol, each q's run stops by the rule of :class:`StopRule` (its terminated spectrum
    changes by less than spectrum_tol, relative L1, over both 10 and 20 steps; checked every 5 steps from
    step 60) and
    nsteps is only the cap; the steps each q took are in LAST_FAST_STEPS.  With precision=0, terminator=False,
    two_pass_gs=False and real=False it is the plain method on the stored H.  Returns (seconds, spectra)."""
    from exactdiag_amd.closure import sqrt_terminator
    from exactdiag_amd.sector import StopRule, real_phase

    tm = dict.fromkeys(("ground_state", "bases", "stored_h", "operator", "lanczos", "continued_fraction"), 0.0)

    def lap(key, t):
        _sync()
        now = time.perf_counter()
        tm[key] += now - t
        return now

    _sync()
    t0 = t = time.perf_counter()
    s0 = Sector(L, h, nd, k0[0], k0[1])
    t = lap("bases", t)
    hm0 = s0.store(weights)
    t = lap("stored_h", t)
    if two_pass_gs:
        from exactdiag_amd.forms import two_pass_ground_state

        e0, psi, _, _ = two_pass_ground_state(hm0, precision=0, nsteps=GS_STEPS, tol=tol, real=real)
    else:
        e0, psi = _gs(hm0.hv, s0.dim, tol)
    hm0.close()
    t = lap("ground_state", t)
    out, steps = {}, {}
    for qx, qy in distinct_qs(L, limited):
        kx, ky = final_momentum(L, k0[0], k0[1], qx, qy)
        sec = Sector(L, h, nd, kx, ky)
        t = lap("bases", t)
        phi = s0.op_apply(sec, _lib.OP_SZQ, qx, qy, psi)
        n2 = krylov.vec_norm(phi) ** 2 if sec.dim else 0.0
        t = lap("operator", t)
        if n2 < 1e-16:
            out[(qx, qy)] = np.zeros(len(ws))
        else:
            hm = sec.store(weights)
            t = lap("stored_h", t)
            phr = real_phase(phi) if (real and hm.is_real) else None
            v0, prec = (phr, _lib.PREC_R64 if precision == 0 else _lib.PREC_R32) if phr is not None else (phi, precision)
            if spectrum_tol is not None:
                a, b, why = hm.lanczos_to_tol(v0, nsteps, StopRule(ws, e0, eta, spectrum_tol), precision=prec)
                exhausted = why == 2
            else:
                a, b = hm.lanczos(v0, nsteps, precision=prec)
                exhausted = len(a) < nsteps
            steps[(qx, qy)] = len(a)
            hm.close()
            t = lap("lanczos", t)
            at, bt, tail = sqrt_terminator(a, b, len(a)) if (terminator and not exhausted) else (a, b, None)
            out[(qx, qy)] = cf_eval(at, bt, ws, e0, eta, -n2 / np.pi, tail=tail)
            t = lap("continued_fraction", t)
        sec.close()
    s0.close()
    _sync()
    LAST_FAST_TIMINGS.clear()
    LAST_FAST_TIMINGS.update(tm)
    LAST_FAST_STEPS.clear()
    LAST_FAST_STEPS.update(steps)
    return time.perf_counter() - t0, out


def max_l1(spectra, ref, ws):
    return max(spectral_l1(spectra[q], ref[q], ws) for q in ref)


def allq_batched(L, h, nd, weights, ws, eta, nsteps, k0=(0, 0), limited=True, tol=REFERENCE_EIG_TOL,
                 terminator=True):
    """All q of the spectrum as one batch: every final sector's Lanczos run is an instance of one batched,
    graph-captured run over stored H (batch.spectra), FP64 vectors, closed by the square-root terminator as in
    the fast form.  Returns (seconds, spectra, stage timings)."""
    from exactdiag_amd.batch import Excitation, spectra
    from exactdiag_amd.closure import sqrt_terminator
    from exactdiag_amd.forms import two_pass_ground_state

    tm = {}
    _sync()
    t0 = time.perf_counter()
    s0 = Sector(L, h, nd, k0[0], k0[1])
    hm0 = s0.store(weights)
    e0, psi, _, _ = two_pass_ground_state(hm0, precision=0, nsteps=GS_STEPS, tol=tol, real=True)
    hm0.close()
    _sync()
    tm["ground_state"] = time.perf_counter() - t0
    qs = distinct_qs(L, limited)
    exc = [Excitation("Szq", _lib.OP_SZQ, qx, qy, Sector(L, h, nd, *final_momentum(L, k0[0], k0[1], qx, qy)))
           for qx, qy in qs]
    t1 = time.perf_counter()
    _, coeffs = spectra(exc, s0, psi, e0, weights, ws, eta, nsteps, stored=True)
    _sync()
    tm["batched_lanczos"] = time.perf_counter() - t1
    out = {}
    for (qx, qy), ex, (a, b) in zip(qs, exc, coeffs):
        if len(a) == 0:
            out[(qx, qy)] = np.zeros(len(ws))
            continue
        at, bt, tail = sqrt_terminator(a, b, len(a)) if (terminator and len(a) == nsteps) else (a, b, None)
        out[(qx, qy)] = cf_eval(at, bt, ws, e0, eta, -ex.norm2 / np.pi, tail=tail)
        ex.sector.close()
    s0.close()
    _sync()
    return time.perf_counter() - t0, out, tm


def allq_multik(L, h, nd, weights, ws, eta, nsteps, k0=(0, 0), limited=True, tol=REFERENCE_EIG_TOL, precision=1,
                terminator=True):
    """The fast form with every q in one run: the final sectors share one stored H over the union basis
    (multik.MultiK, up to 16 momenta per run), their Lanczos recurrences advance together (complex64 vectors
    with FP64 reductions by default), each closed by the square-root terminator.  The ground state is the fast
    form's.  Returns (seconds, spectra, stage timings)."""
    from exactdiag_amd.closure import sqrt_terminator
    from exactdiag_amd.forms import two_pass_ground_state
    from exactdiag_amd.multik import SLOTS, MultiK

    tm = dict.fromkeys(("ground_state", "bases", "stored_h", "operator", "lanczos", "continued_fraction"), 0.0)

    def lap(key, t):
        _sync()
        now = time.perf_counter()
        tm[key] += now - t
        return now

    _sync()
    t0 = t = time.perf_counter()
    s0 = Sector(L, h, nd, k0[0], k0[1])
    t = lap("bases", t)
    hm0 = s0.store(weights)
    t = lap("stored_h", t)
    e0, psi, _, _ = two_pass_ground_state(hm0, precision=0, nsteps=GS_STEPS, tol=tol, real=True)
    hm0.close()
    t = lap("ground_state", t)
    qs = distinct_qs(L, limited)
    u = Sector.union(L, h, nd)
    t = lap("bases", t)
    out = {}
    for lo in range(0, len(qs), SLOTS):
        group = qs[lo: lo + SLOTS]
        ks = [final_momentum(L, k0[0], k0[1], qx, qy) for qx, qy in group]
        mk = MultiK(u, weights, ks)
        t = lap("stored_h", t)
        v0 = mk.block()
        norm2 = []
        for slot, ((qx, qy), k) in enumerate(zip(group, ks)):
            sec = Sector(L, h, nd, *k)
            t = lap("bases", t)
            phi = s0.op_apply(sec, _lib.OP_SZQ, qx, qy, psi)
            n2 = krylov.vec_norm(phi) ** 2 if sec.dim else 0.0
            if n2 >= 1e-16:
                mk.attach(slot, sec, phi, v0)
            norm2.append(n2)
            del phi
            sec.close()
            t = lap("operator", t)
        coeffs = mk.lanczos(v0, nsteps, precision=precision)
        del v0
        mk.close()
        t = lap("lanczos", t)
        for (q, n2, (a, b)) in zip(group, norm2, coeffs):
            if n2 < 1e-16 or len(a) == 0:
                out[q] = np.zeros(len(ws))
                continue
            at, bt, tail = sqrt_terminator(a, b, len(a)) if (terminator and len(a) == nsteps) else (a, b, None)
            out[q] = cf_eval(at, bt, ws, e0, eta, -n2 / np.pi, tail=tail)
        t = lap("continued_fraction", t)
    u.close()
    s0.close()
    _sync()
    return time.perf_counter() - t0, out, tm
