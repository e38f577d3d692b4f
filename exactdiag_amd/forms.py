"""The batching and precision forms, measured end to end.

* :func:`batching_form` -- every momentum sector of a ladder for several
  couplings as uneven Lanczos instances of one batch (lowest eigenpair each,
  restarted two-pass Lanczos in HIP-graph launches), then the one-particle
  spectral functions A+(k, w) and A-(k, w) of each coupling's ground state
  for every k in one more batch; the same work done one sector at a time
  (the reference's loops, tJ_spin_half_ladder/api.py:45-59 and 135-161) for
  the timing and the agreement.
* :func:`precision_form` -- ground state and S(q, w) of one sector with
  complex64 Lanczos vectors and H·v (FP64 coefficient reductions,
  tridiagonal solve and continued fraction) against the same computation in
  complex128; every energy and spectrum error stated.
"""

from __future__ import annotations

import time

import numpy as np
from scipy.linalg import eigh_tridiagonal

from exactdiag_amd import _lib
from exactdiag_amd.allq import final_momentum, spectral_l1
from exactdiag_amd.general import krylov
from exactdiag_amd.sector import Sector, cf_eval, weights_array


def _sync():
    import torch

    torch.cuda.synchronize()


def _shortest_converged_prefix(a, b, tol, stride=8):
    """Lowest eigenvector y of the shortest leading k-step tridiagonal whose Ritz residual estimate
    b_k |y_{k-1}| is below tol / 4: prefixes every `stride` steps, then bisection inside the first stride that
    converges (the estimate falls steadily until orthogonality is lost, so the scan runs forward); the whole
    tridiagonal when no prefix converges."""
    n = len(a)

    def lowest(k):
        if k == 1:
            return np.array([a[0]]), np.ones((1, 1))
        return eigh_tridiagonal(a[:k], b[1:k], select="i", select_range=(0, 0))

    def converged(k):
        th, Y = lowest(k)
        return abs(b[k] * Y[-1, 0]) <= 0.25 * tol * max(1.0, abs(th[0]))

    prev = 0
    for k in range(stride, n, stride):  # prefixes k < n have b[k]
        if converged(k):
            lo, hi = prev + 1, k
            while lo < hi:
                mid = (lo + hi) // 2
                if converged(mid):
                    hi = mid
                else:
                    lo = mid + 1
            return lowest(hi)[1][:, 0]
        prev = k
    return lowest(n)[1][:, 0]


BASIS_BYTES = 16 << 30  # HBM for the pass-one Lanczos vectors of two_pass_ground_state (16 GiB of the 288 GB)


def two_pass_ground_state(hm, precision=0, nsteps=150, tol=1e-10, max_restarts=30, seed=0, real=False):
    """Lowest eigenpair of a stored H by restarted two-pass Lanczos in the given precision.

    Pass one gives the coefficients; the shortest leading tridiagonal whose Ritz residual estimate meets the
    tolerance gives the lowest eigenvector y; pass two reruns that many steps of the recurrence and accumulates
    x = sum_n y_n v_n in FP64; restart from x until ||H x - E x|| <= tol * max(1, |E|), E the FP64 Rayleigh
    quotient.  In FP64 (precision 0), when the pass-one vectors fit in BASIS_BYTES of HBM, pass one keeps them
    (ed_hmat_lanczos_run_basis) and x is their combination instead of a second pass.  With real=True and a
    real stored H (hm.is_real) the vectors are real (float64 for precision 0, float32 for 1) from a real start
    vector; x is still returned as complex128.
    Returns (E, x, restarts, residual).
    """
    import torch

    from exactdiag_amd._lib import PREC_R32, PREC_R64

    dim = hm.dim
    v = krylov.device_start_vector(dim, seed)
    real = bool(real and hm.is_real)
    if real:
        v = v.real.to(torch.complex128)
        precision = PREC_R64 if precision == 0 else PREC_R32
    n1 = min(nsteps, dim)
    Dp = dim + (dim & 1)  # real rows padded to an even length, read as complex pairs by vec_combine
    keep = precision in (0, PREC_R64) and n1 * (Dp * 8 if real else dim * 16) <= BASIS_BYTES
    res = np.inf
    for restart in range(max_restarts):
        if keep:
            a, b, V = hm.lanczos_basis(v, n1, precision=precision, ld=Dp if real else dim)
            y = _shortest_converged_prefix(a, b, tol)
            k = len(y)
            if real:
                Vc = torch.view_as_complex(V[:k].view(k, Dp // 2, 2))
                xc = torch.empty(Dp // 2, dtype=torch.complex128, device=V.device)
                krylov.vec_combine(Vc, k, y.astype(np.complex128), xc)  # real weights: the real combination
                x = torch.view_as_real(xc).reshape(-1)[:dim].to(torch.complex128)
            else:
                x = torch.empty(dim, dtype=torch.complex128, device=V.device)
                krylov.vec_combine(V, k, y, x)
            del V
        else:
            a, b = hm.lanczos(v, n1, precision=precision)
            y = _shortest_converged_prefix(a, b, tol)
            x = hm.lanczos_ritz(v, y, precision=precision)
        krylov.vec_scale(x, 1.0 / krylov.vec_norm(x))
        if real:
            xr = x.real.contiguous()
            hx, e = hm.hv(xr, with_dot=True)
            e = e.real
            res = float(torch.linalg.vector_norm(hx - e * xr))
        else:
            hx, e = hm.hv(x, with_dot=True)
            e = e.real
            r = hx.clone()
            krylov.vec_axpys(x.view(1, -1), 1, [e], r)
            res = krylov.vec_norm(r)
        if res <= tol * max(1.0, abs(e)) or len(a) < n1:
            return e, x, restart, res
        v = x
    raise RuntimeError(f"two-pass ground state: residual {res:.3e} above {tol:.1e} x max(1, |E|) after "
                       f"{max_restarts} restarts (E = {e:.12f})")


def precision_form(L=14, h=2, nd=13, k=(7, 0), q=(7, 1), weights=(1.0, 1.0, 0.4, 0.4), nsteps=300,
                   ws=np.linspace(0, 4, 2001), eta=0.05, log=print, flip=False, keep_spectra=False):
    """complex64 vs complex128: ground-state energy and S(q, w) errors of one sector (stored H).

    flip: the sector split by the spin flip (ed_sector_create_flip): the ground state is the lower of the two
    parity blocks' (found by a complex128 two-pass run in each), and Szq leads to the opposite parity block of
    the final momentum; half the states of the momentum sector per block."""
    w = list(weights)
    t0 = time.perf_counter()
    kx, ky = final_momentum(L, k[0], k[1], q[0], q[1])
    out = {}
    if flip:
        blocks = {z: Sector(L, h, nd, k[0], k[1], flip=z) for z in (1, -1)}
        lows = {}
        for z, sz in blocks.items():
            hm = sz.store(w)
            lows[z] = two_pass_ground_state(hm, precision=0, tol=1e-10)[0]
            hm.close()
        z0 = 1 if lows[1] <= lows[-1] else -1
        out["flip"] = {"ground_state_parity": z0, "block_dims": {str(z): blocks[z].dim for z in blocks},
                       "block_lowest": {str(z): lows[z] for z in lows}}
        s0 = blocks[z0]
        blocks[-z0].close()
        s1 = Sector(L, h, nd, kx, ky, flip=-z0)
    else:
        s0 = Sector(L, h, nd, k[0], k[1])
        s1 = Sector(L, h, nd, kx, ky)
    hm0 = s0.store(w)
    hm1 = s1.store(w)
    _sync()
    t_setup = time.perf_counter() - t0
    log(f"precision form{' (spin flip)' if flip else ''}: sectors {s0.dim} and {s1.dim} states, stored H "
        f"{(hm0.bytes + hm1.bytes) / 1e9:.1f} GB, set up in {t_setup:.1f}s")
    out.update({"dims": [s0.dim, s1.dim], "setup_s": t_setup, "stored_gb": (hm0.bytes + hm1.bytes) / 1e9})
    spectra = {}
    for name, prec, tol in (("c128", 0, 1e-10), ("c64", 1, 1e-5)):
        t1 = time.perf_counter()
        e0, psi, restarts, res = two_pass_ground_state(hm0, precision=prec, tol=tol)
        _sync()
        t_gs = time.perf_counter() - t1
        phi = s0.op_apply(s1, _lib.OP_SZQ, q[0], q[1], psi)
        n2 = krylov.vec_norm(phi) ** 2
        t2 = time.perf_counter()
        a, b = hm1.lanczos(phi, nsteps, precision=prec)
        _sync()
        t_lz = time.perf_counter() - t2
        spectra[name] = cf_eval(a, b, ws, e0, eta, -n2 / np.pi)
        out[name] = {"e0": e0, "gs_restarts": restarts, "gs_residual": res, "gs_seconds": t_gs,
                     "lanczos_seconds": t_lz, "steps": len(a), "norm2": n2}
        log(f"  {name}: E0 = {e0:.12f} ({restarts} restarts, residual {res:.1e}) in {t_gs:.1f}s; "
            f"{len(a)} Lanczos steps in {t_lz:.1f}s")
    out["e0_abs_error_c64"] = abs(out["c64"]["e0"] - out["c128"]["e0"])
    out["e0_rel_error_c64"] = out["e0_abs_error_c64"] / abs(out["c128"]["e0"])
    out["spectrum_l1_error_c64"] = spectral_l1(spectra["c64"], spectra["c128"], ws)
    out["spectrum_max_rel_error_c64"] = float(np.max(np.abs(spectra["c64"] - spectra["c128"]))
                                              / np.max(np.abs(spectra["c128"])))
    if keep_spectra:
        out["spectra"] = spectra
    for obj in (hm1, hm0, s1, s0):
        obj.close()
    return out


def _spectral_excitations(L, h, nd, k0):
    """(name, op, qx, qy, final sector params) for A+ and A- at every k (matrix_setup.pyx:281-330)."""
    N = 2 * L
    tsp = (N - h) - 2 * nd
    out = []
    for qx in range(L):
        for qy in range(2):
            if h > 0:
                ndf, op = (nd + 1, _lib.OP_APLUS_DOWN) if tsp > 0 else (nd, _lib.OP_APLUS_UP)
                kx, ky = final_momentum(L, k0[0], k0[1], qx, qy)
                out.append(("A+", op, qx, qy, (L, h - 1, ndf, kx, ky)))
            if h < N:
                ndf, op = (nd - 1, _lib.OP_AMINUS_DOWN) if tsp <= 0 else (nd, _lib.OP_AMINUS_UP)
                kx, ky = final_momentum(L, k0[0], k0[1], -qx, (-qy) % 2)
                if 0 <= ndf <= N - h - 1:
                    out.append(("A-", op, -qx, (-qy) % 2, (L, h + 1, ndf, kx, ky)))
    return out


def _groups(items, size=15):
    """Items in groups of at most `size`, as even as possible."""
    n = max(1, -(-len(items) // size))
    per = -(-len(items) // n)
    return [items[i: i + per] for i in range(0, len(items), per)]


def _batching_multik(L, h, nd, Js, inst, sectors, energies_batched, batched_spectra, nsteps, ws, eta, tol, log):
    """The batching form's two steps with the instances of one ladder in multi-momentum products (per-slot
    momentum and couplings, at most 16 slots per product): lowest eigenpairs of all (J, k) instances, then
    A(k, w) for every k from each coupling's ground state.  Compared with the batched engine's results."""
    import torch

    from exactdiag_amd.multik import MultiK

    out = {}
    _sync()
    t0 = time.perf_counter()
    u = Sector.union(L, h, nd)
    energies = np.empty(len(inst))
    psis = {}
    best = {J: min((i for i, (JJ, _) in enumerate(inst) if JJ == J), key=lambda i: energies_batched[i]) for J in Js}
    for grp in _groups(list(range(len(inst)))):
        mk = MultiK(u, [[1.0, 1.0, inst[i][0], inst[i][0]] for i in grp], [inst[i][1] for i in grp])
        v0 = mk.block()
        for slot, i in enumerate(grp):
            s = sectors[inst[i][1]]
            mk.attach(slot, s, krylov.device_start_vector(s.dim, i), v0)
        e, x, _, _ = mk.lowest(v0, tol=tol)
        for slot, i in enumerate(grp):
            energies[i] = e[slot]
            if i in best.values():
                psis[i] = mk.extract(slot, sectors[inst[i][1]], x).clone()
        del v0, x
        mk.close()
    _sync()
    out["gs_s"] = time.perf_counter() - t0
    out["gs_max_abs_diff_vs_batched"] = float(np.max(np.abs(energies - energies_batched)))
    log(f"  multi-momentum lowest eigenpairs: {out['gs_s']:.2f}s, max |dE| vs batched "
        f"{out['gs_max_abs_diff_vs_batched']:.1e}")
    # spectra: the excitations of all couplings, grouped by final ladder
    t0 = time.perf_counter()
    jobs = {}
    for J in Js:
        i0 = best[J]
        k0, e0, exc = batched_spectra[J][:3]
        assert inst[i0][1] == k0
        for n_ex, ex in enumerate(exc):
            jobs.setdefault((ex.sector.h, ex.sector.nd), []).append((J, n_ex, ex, i0))
    got = _multik_spectra(L, jobs, sectors, inst, psis, batched_spectra, nsteps, ws, eta)
    _sync()
    out["spectra_s"] = time.perf_counter() - t0
    diffs = [spectral_l1(got[(J, n)], g, ws) for J in Js for n, g in enumerate(batched_spectra[J][3])
             if np.abs(g).sum() > 0]
    out["spectra_max_l1_vs_batched"] = float(max(diffs))
    out["spectra_median_l1_vs_batched"] = float(np.median(diffs))
    out["n_spectra"] = len(got)
    # the same spectra from the batched engine's ground states: the difference of the two Lanczos paths alone
    same = _multik_spectra(L, jobs, sectors, inst, {best[J]: batched_spectra[J][4] for J in Js}, batched_spectra,
                           nsteps, ws, eta)
    diffs2 = [spectral_l1(same[(J, n)], g, ws) for J in Js for n, g in enumerate(batched_spectra[J][3])
              if np.abs(g).sum() > 0]
    out["spectra_same_gs_max_l1_vs_batched"] = float(max(diffs2))
    out["spectra_same_gs_median_l1_vs_batched"] = float(np.median(diffs2))
    log(f"  multi-momentum A(k,w): {out['spectra_s']:.2f}s, L1 vs batched median {out['spectra_median_l1_vs_batched']:.1e}"
        f" max {out['spectra_max_l1_vs_batched']:.1e}; from the batched ground states median "
        f"{out['spectra_same_gs_median_l1_vs_batched']:.1e} max {out['spectra_same_gs_max_l1_vs_batched']:.1e}")
    u.close()
    torch.cuda.empty_cache()
    return out


def _multik_spectra(L, jobs, sectors, inst, psis, batched_spectra, nsteps, ws, eta):
    """Spectra of the grouped excitations (jobs: final ladder -> [(J, index, excitation, ground-state
    instance)]) with multi-momentum products; returns {(J, index): S(w)}."""
    from exactdiag_amd.multik import MultiK

    got = {}
    for (hf, ndf), items in jobs.items():
        uf = Sector.union(L, hf, ndf)
        for grp in _groups(items):
            mk = MultiK(uf, [[1.0, 1.0, J, J] for J, *_ in grp], [(ex.sector.kx, ex.sector.ky) for _, _, ex, _ in grp])
            v0 = mk.block()
            norm2 = []
            for slot, (J, n_ex, ex, i0) in enumerate(grp):
                phi = sectors[inst[i0][1]].op_apply(ex.sector, ex.op, ex.qx, ex.qy, psis[i0])
                n2 = krylov.vec_norm(phi) ** 2 if ex.sector.dim else 0.0
                if n2 >= 1e-16:
                    mk.attach(slot, ex.sector, phi, v0)
                norm2.append(n2)
            coeffs = mk.lanczos(v0, nsteps)
            for (J, n_ex, ex, i0), n2, (a, b) in zip(grp, norm2, coeffs):
                e0 = batched_spectra[J][1]
                got[(J, n_ex)] = (np.zeros(len(ws)) if n2 < 1e-16 or len(a) == 0
                                  else cf_eval(a, b, ws, e0, eta, -n2 / np.pi))
            del v0
            mk.close()
        uf.close()
    return got


def batching_form(L=10, h=2, nd=9, Js=(0.3, 0.4, 0.5), nsteps=250, ws=np.linspace(-6, 6, 2001), eta=0.05,
                  tol=1e-9, serial=True, stored=True, log=print):
    """All k sectors x couplings as one batch, then A(k, w) for every k as another; optionally the serial runs.

    With ``stored`` both the batch and the one-sector-at-a-time runs use stored H (built inside the timed
    regions of the spectra; for the ground states the builds are timed separately, ``stored_build_s``)."""
    import torch

    from exactdiag_amd.batch import Batch, Excitation, lowest_eigenpairs, spectra

    ks = [(kx, ky) for kx in range(L) for ky in range(2)]
    t0 = time.perf_counter()
    sectors = {kk: Sector(L, h, nd, kk[0], kk[1]) for kk in ks}
    _sync()
    t_basis = time.perf_counter() - t0
    weights = [[1.0, 1.0, J, J] for J in Js]
    inst = [(J, kk) for J in Js for kk in ks]
    t0 = time.perf_counter()
    batch = Batch([sectors[kk] for _, kk in inst], [[1.0, 1.0, J, J] for J, _ in inst], stored=stored)
    _sync()
    t_build = time.perf_counter() - t0
    tiles, gather, ms = batch.tune()
    log(f"batching form: {len(inst)} instances, dims {int(batch.dims.min())}..{int(batch.dims.max())}, total "
        f"{batch.total}; tuned tiles={tiles} gather={gather} ({ms:.2f} ms per batched H·v)")
    _sync()
    t1 = time.perf_counter()
    energies, x, restarts = lowest_eigenpairs(batch, tol=tol)
    _sync()
    t_gs_batch = time.perf_counter() - t1
    log(f"  batched lowest eigenpairs: {t_gs_batch:.2f}s ({restarts} restarts)")
    out = {"instances": len(inst), "dims": [int(d) for d in batch.dims], "basis_s": t_basis,
           "tuned": {"tiles": tiles, "gather": gather, "ms_per_batched_hv": ms},
           "gs_batched_s": t_gs_batch, "gs_restarts": restarts, "stored": stored,
           "stored_build_s": t_build if stored else 0.0}
    if serial:
        ops = {}
        for J, kk in inst:
            wa = weights_array([1.0, 1.0, J, J])
            ops[(J, kk)] = (sectors[kk].store(wa).hv if stored
                            else (lambda v, s=sectors[kk], wa=wa: s.hv(wa, v)))
        _sync()
        t1 = time.perf_counter()
        e_serial = []
        for J, kk in inst:
            vals, _, _ = krylov.lowest_eigenpairs(ops[(J, kk)], sectors[kk].dim, 1, tol=tol)
            e_serial.append(float(vals[0]))
        _sync()
        out["gs_serial_s"] = time.perf_counter() - t1
        del ops
        out["gs_max_abs_diff"] = float(np.max(np.abs(np.array(e_serial) - energies)))
        log(f"  serial lowest eigenpairs: {out['gs_serial_s']:.2f}s, max |dE| {out['gs_max_abs_diff']:.2e}")
    # ground state per coupling: the lowest of its sectors
    batched_spectra = {}
    exc_params = {}
    spec_out = {}
    t_spec_b = t_spec_s = 0.0
    max_diff = 0.0
    for ji, J in enumerate(Js):
        idx = [i for i, (JJ, _) in enumerate(inst) if JJ == J]
        i0 = idx[int(np.argmin(energies[idx]))]
        k0 = inst[i0][1]
        e0 = float(energies[i0])
        psi = batch.part(x, i0).clone()
        exc = []
        for name, op, qx, qy, prm in _spectral_excitations(L, h, nd, k0):
            if prm not in exc_params:
                exc_params[prm] = Sector(*prm)
            exc.append(Excitation(name, op, qx, qy, exc_params[prm]))
        _sync()
        t1 = time.perf_counter()
        got, coeffs = spectra(exc, sectors[k0], psi, e0, weights[ji], ws, eta, nsteps, stored=stored)
        _sync()
        t_spec_b += time.perf_counter() - t1
        spec_out[J] = {"k0": k0, "e0": e0, "n_spectra": len(exc)}
        batched_spectra[J] = (k0, e0, exc, got, psi)
        if serial:
            t1 = time.perf_counter()
            wa = weights_array(weights[ji])
            for ex, g in zip(exc, got):
                phi = sectors[k0].op_apply(ex.sector, ex.op, ex.qx, ex.qy, psi)
                n2 = krylov.vec_norm(phi) ** 2 if ex.sector.dim else 0.0
                if n2 < 1e-16:
                    continue
                if stored:
                    m = ex.sector.store(wa)
                    a, b = m.lanczos(phi, nsteps)
                    m.close()
                else:
                    a, b = ex.sector.lanczos(wa, phi, nsteps)
                ref = cf_eval(a, b, ws, e0, eta, -n2 / np.pi)
                max_diff = max(max_diff, spectral_l1(g, ref, ws))
            _sync()
            t_spec_s += time.perf_counter() - t1
    out["spectra"] = {str(J): v for J, v in spec_out.items()}
    out["spectra_batched_s"] = t_spec_b
    if serial:
        out["spectra_serial_s"] = t_spec_s
        out["spectra_max_l1_batched_vs_serial"] = max_diff
    log(f"  A(k,w): batched {t_spec_b:.2f}s" + (f", serial {t_spec_s:.2f}s, max L1 diff {max_diff:.1e}"
                                                   if serial else ""))
    out["multik"] = _batching_multik(L, h, nd, Js, inst, sectors, energies, batched_spectra, nsteps, ws, eta, tol,
                                     log)
    batch.close()
    for s in list(exc_params.values()) + list(sectors.values()):
        s.close()
    del x
    torch.cuda.empty_cache()
    return out


def csr_form(L=12, h=2, nd=11, k=(0, 0), weights=(1.0, 1.0, 0.4, 0.4), reps=20, log=print):
    """The reference's stored matrices of one sector built on the GPU, and their summed product.

    calc_sparse_matrix (general/cython_sparse_matrices.pyx:185-241) for the four Hamiltonian terms, in the
    default layout (row-major Hermitian upper triangle), then the weighted sum of the four applied as
    Added_Sparse_Matrices.dot; the product is checked against the stored-H product.  Returns (results,
    matrices) so that a caller can compare the arrays.
    """
    import torch

    from exactdiag_amd.csr import CSRSum, StoredCSR

    t0 = time.perf_counter()
    s = Sector(L, h, nd, *k)
    _sync()
    out = {"dim": s.dim, "basis_s": time.perf_counter() - t0, "terms": []}
    mats = []
    for t in range(4):
        _sync()
        t1 = time.perf_counter()
        m = StoredCSR.from_sector(s, t)
        _sync()
        dt = time.perf_counter() - t1
        mats.append(m)
        out["terms"].append({"term": ("tl", "tr", "jl", "jr")[t], "nnz": m.nnz, "n_values": m.n_values,
                             "max_count": m.max_count, "build_s": dt})
        log(f"  term {t}: {m.nnz} elements, {m.n_values} unique values, built in {dt:.3f}s")
    out["build_s"] = sum(d["build_s"] for d in out["terms"])
    t1 = time.perf_counter()
    op = CSRSum(mats, list(weights))
    _sync()
    out["sum_build_s"] = time.perf_counter() - t1
    x = krylov.device_start_vector(s.dim, 1)
    y = op.hv(x)
    hm = s.store(weights_array(list(weights)))
    y_ref = hm.hv(x)
    out["rel_diff_vs_stored_h"] = float((y - y_ref).abs().max() / y_ref.abs().max())
    hm.close()
    for _ in range(3):
        op.hv(x, out=y)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        op.hv(x, out=y)
    e1.record(stream)
    _sync()
    ms = e0.elapsed_time(e1) / reps
    out.update({"sum_nnz": op.nnz, "sum_slots": op.slots, "hv_ms": ms, "hv_gstates_per_s": s.dim / ms / 1e6})
    log(f"  sum of the four: {op.nnz} elements ({op.slots} slots), H·v {ms:.3f} ms, "
        f"rel. diff vs stored H {out['rel_diff_vs_stored_h']:.1e}")
    op.close()
    return out, mats


def positions_form(reps=20, log=print):
    """Position correlations of ground states (tJ_spin_half_ladder/position_correlations.pyx).

    Hole correlations of the 2x12, 2-hole, k = (0,0) ground state (D = 8,112,468; J = 0.4, t = 1): one
    reduction over the basis for all 24 node distances.  Singlet-singlet correlations of the 2x10, 4-hole,
    k = (0,0) ground state (weights 0.45, 0.36, 0.15, 0.12, y = x = (1, 1)): one fused expectation value per
    relative position of the singlets (20).  Ground states by two-pass Lanczos on the stored H.  Returns
    (results, states) with the states on the host for a CPU comparison.
    """
    import torch

    from exactdiag_amd import _lib

    out, states = {}, {}
    stream = torch.cuda.current_stream()
    # untimed: one small ground state first, so that the timed ones do not pay for loading kernels
    _warm = Sector(6, 2, 5, 0, 0).store(weights_array([1.0, 1.0, 0.4, 0.4]))
    two_pass_ground_state(_warm, tol=1e-9, real=True)
    _warm.close()

    def timed(fn):
        fn()
        _sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(reps):
            r = fn()
        e1.record(stream)
        _sync()
        return r, e0.elapsed_time(e1) / reps, (time.perf_counter() - t0) / reps

    # hole correlations, 2x12
    s = Sector(12, 2, 11, 0, 0)
    hm = s.store(weights_array([1.0, 1.0, 0.4, 0.4]))
    t0 = time.perf_counter()
    e0, x, _, res = two_pass_ground_state(hm, tol=1e-9, real=True)
    _sync()
    t_gs = time.perf_counter() - t0
    corr, ms, wall = timed(lambda: s.position_correlations(_lib.POS_HOLES, np.zeros((0, 2), np.int32), x))
    out["holes_2x12"] = {"dim": s.dim, "e0": e0, "ground_state_s": t_gs, "residual": res, "kernel_ms": ms,
                         "call_ms": wall * 1e3, "correlations": (corr / corr[0]).tolist(),
                         "sum_rule": float(corr.sum()), "bytes_per_call": 24 * s.dim,
                         "gbs": 24 * s.dim / ms / 1e6}
    log(f"  holes 2x12: D={s.dim}, E0={e0:.10f}, kernel {ms:.3f} ms ({24 * s.dim / ms / 1e6:.0f} GB/s)")
    states["holes"] = (s, x.cpu().numpy())
    hm.close()
    # singlet-singlet, 2x10 with 4 holes
    s4 = Sector(10, 4, 8, 0, 0)
    hm = s4.store(weights_array([0.45, 0.36, 0.15, 0.12]))
    t0 = time.perf_counter()
    e0, x, _, res = two_pass_ground_state(hm, tol=1e-9, real=True)
    _sync()
    t_gs = time.perf_counter() - t0
    fixed = np.array([[1, 1], [1, 1]], np.int32)

    def all_shifts():
        return [s4.position_expectation(_lib.POS_SINGLET, np.vstack([fixed, [[d // 2, d % 2]]]), x).real
                for d in range(s4.N)]

    corr, ms, wall = timed(all_shifts)
    out["singlet_2x10_h4"] = {"dim": s4.dim, "e0": e0, "ground_state_s": t_gs, "residual": res,
                              "all_shifts_ms": ms, "per_shift_ms": ms / s4.N, "correlations": corr}
    log(f"  singlet 2x10 h4: D={s4.dim}, E0={e0:.10f}, {s4.N} expectation values in {ms:.2f} ms")
    states["singlet"] = (s4, x.cpu().numpy())
    hm.close()
    return out, states


def multik_form(L=12, h=2, nd=11, weights=(1.0, 1.0, 0.4, 0.4), reps=10, log=print):
    """The final sectors of all distinct q (14 momenta of the 2x12 ladder, 2 holes) in one stored product over
    the union basis, against the same 14 products with one stored H per sector (StoredH), complex64 and
    complex128.  Agreement: each slot against its sector's own product.  Returns a dict."""
    import torch

    from exactdiag_amd.allq import distinct_qs, final_momentum
    from exactdiag_amd.multik import MultiK

    stream = torch.cuda.current_stream()

    def timed(fn):
        for _ in range(2):
            fn()
        _sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        _sync()
        return e0.elapsed_time(e1) / reps

    ks = [final_momentum(L, 0, 0, qx, qy) for qx, qy in distinct_qs(L)]
    t0 = time.perf_counter()
    u = Sector.union(L, h, nd)
    _sync()
    t_union = time.perf_counter() - t0
    t0 = time.perf_counter()
    mk = MultiK(u, weights, ks)
    _sync()
    t_build = time.perf_counter() - t0
    secs = [Sector(L, h, nd, *k) for k in ks]
    mats = [s.store(weights) for s in secs]
    out = {"union_dim": u.dim, "sector_dims": [s.dim for s in secs], "slots": len(ks), "nnz": mk.nnz,
           "union_basis_s": t_union, "build_s": t_build, "bytes": mk.bytes}
    g = torch.Generator(device="cuda").manual_seed(7)
    xs = [torch.complex(torch.randn(s.dim, dtype=torch.float64, device="cuda", generator=g),
                        torch.randn(s.dim, dtype=torch.float64, device="cuda", generator=g)) for s in secs]
    block = mk.block()
    for k, (s, x) in enumerate(zip(secs, xs)):
        mk.attach(k, s, x, block)
    worst = 0.0
    y = mk.hv(block)
    for k, (s, m, x) in enumerate(zip(secs, mats, xs)):
        yk = m.hv(x.clone())
        worst = max(worst, float((mk.extract(k, s, y) - yk).abs().max() / yk.abs().max()))
    out["max_rel_diff_c128"] = worst
    del y
    for name, dt in (("c64", torch.complex64), ("c128", torch.complex128)):
        xb = block.to(dt)
        yb = torch.empty_like(xb)
        ms_mk = timed(lambda: mk.hv(xb, out=yb))
        xv = [x.to(dt) for x in xs]
        yv = [torch.empty_like(x) for x in xv]

        def serial():
            for m, a, b in zip(mats, xv, yv):
                m.hv(a, out=b)

        ms_serial = timed(serial)
        states = sum(out["sector_dims"])
        out[name] = {"multik_ms": ms_mk, "serial_ms": ms_serial, "speedup": ms_serial / ms_mk,
                     "gstates_per_s": states / ms_mk / 1e6}
        log(f"  {name}: one product over {len(ks)} momenta {ms_mk:.3f} ms, {len(ks)} stored products "
            f"{ms_serial:.3f} ms ({ms_serial / ms_mk:.2f}x)")
        del xb, yb, xv, yv
    for m in mats:
        m.close()
    mk.close()
    return out
