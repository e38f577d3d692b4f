"""Benchmark: t-J H·v throughput on one MI355X.

Workload: 2x12 ladder (24 sites, periodic along the legs), 2 holes, S^z = 0
(11 down spins), k = (0, 0) even-parity sector, t_leg = t_rung = 1,
J_leg = J_rung = 0.4, complex128 vectors; D = 8,112,468 states.  One step is
one H·v of the whole sector with the fused <x, Hx> reduction the Lanczos
recurrence uses (the hot path of every Lanczos step).  Inputs are seeded
synthetic vectors already resident in HBM.

The timed H·v is the stored form (sliced-ELL H built once on the GPU before
the timed region, as the reference builds its CSR matrices once before its
Lanczos loop); the matrix-free kernel is timed too and reported under
``matrix_free``.  Prints one JSON line (metric, value, unit, steps, timing and config fields) plus
``roofline`` (the H·v kernel against the 8 TB/s HBM peak, timed with HIP
events on the launch stream, and under ``read_rate`` the lines it moves per
second against the device's measured rate for scattered 16-byte loads) and ``cpu_baseline`` (the reference's own
sparse-dot kernels, from oracle/_ref, timed on the host's cores over the
whole sector's stored matrix, and one q of the all-q run end to end).
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

L, HOLES, NDOWN, KX, KY = 12, 2, 11, 0, 0
WEIGHTS = [1.0, 1.0, 0.4, 0.4]
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per state per H·v (complex128): read x_f, write y_f,
# 8-byte representative key and the orbit size read as 4 bytes
ALG_BYTES_PER_STATE = 2 * 16 + 8 + 4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel, name="hv_2x12_h2_traffic.md"):
    """HBM bytes per launch of ``kernel`` from a committed rocprofv3 PMC summary, if any."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\[" + re.escape(kernel) + r"\]:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def host_cpu():
    """The host's processor model, its logical CPU count, the CPUs this process may use and the threads used."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    # every host core this process may use, capped by OMP_NUM_THREADS where the machine sets its CPU share there
    threads = min(affinity, int(omp)) if omp.isdigit() and int(omp) > 0 else affinity
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": omp or None,
            "threads": threads}


def _ref_hv(csr, x, y, threads):
    """y = H x with the reference's compiled sparse-dot kernels over the four stored terms, as
    Added_Sparse_Matrices.dot drives them (general/sparse_matrices.py:233-237)."""
    from oracle import oracle

    for t in range(4):
        values, vinds, minor, major = csr[t]
        oracle.ref_csr_dot(values, vinds, minor, major, x, y, WEIGHTS[t], t == 0, threads=threads)


CPU_Q = (6, 1)  # q = (pi, pi) of the per-q run
CPU_Q_STEPS = 24  # Lanczos steps timed per q (each step costs the same; the plain method runs 256)


def cpu_baseline(x, y_gpu, threads):
    """The reference CPU path on the headline sector, measured on this host.

    * H·v: the reference's compiled sparse-dot kernels (oracle/_ref, general/cpp_sparse_dot.h) over the whole
      sector's 4-term upper-triangle CSR (built by the oracle's restatement of calc_sparse_matrix), on the same
      x as the GPU, its output checked against the oracle's own product and the GPU's;
    * one q of the all-q run end to end: the final sector's basis, the Szq operator's CSR, the final sector's
      4-term CSR and Lanczos steps (reference dot + numpy vector updates, as
      cython_lanczos_diagonalization.pyx:50-77)."""
    import numpy as np

    # the compiled reference kernels run under LLVM's OpenMP runtime, whose idle threads spin for 200 ms after
    # each parallel region by default and then starve numpy's own BLAS threads in the Lanczos loop
    # (a 10x slower step on a 2x8 sector); a zero block time puts them to sleep at once
    os.environ.setdefault("KMP_BLOCKTIME", "0")
    from oracle import oracle

    if not oracle.ref_available():
        return {"value": None, "error": "oracle/_ref not built (no reference checkout where it was built)"}
    host = host_cpu()
    log(f"cpu baseline: {host}")
    t0 = time.perf_counter()
    o = oracle.OracleSector(L, HOLES, NDOWN, KX, KY, threads=threads)
    t_basis = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.build_csr((0, 1, 2, 3), threads=threads)
    t_csr = time.perf_counter() - t0
    csr = [o.term_csr(t) for t in range(4)]
    D = o.dim
    log(f"cpu baseline: basis {t_basis:.1f}s, 4-term CSR of all {D} rows {t_csr:.1f}s "
        f"({sum(len(c[2]) for c in csr)} stored elements)")
    y = np.empty(D, np.complex128)
    _ref_hv(csr, x, y, threads)
    y_or = o.hv(WEIGHTS, x, threads=threads)
    same_oracle = bool(np.array_equal(y, y_or))
    rel_gpu = float(np.max(np.abs(y - y_gpu)) / np.max(np.abs(y_gpu)))
    log(f"cpu baseline: reference product equals the oracle's: {same_oracle}; max rel diff to the GPU {rel_gpu:.1e}")

    def timed(nthreads, budget, max_reps):
        reps, total = 0, 0.0
        while (total < budget and reps < max_reps) or reps == 0:
            t1 = time.perf_counter()
            _ref_hv(csr, x, y, nthreads)
            total += time.perf_counter() - t1
            reps += 1
        return reps, total

    reps, total = timed(threads, 8.0, 100)
    reps1, total1 = timed(1, 3.0, 5)
    rate, rate1 = D * reps / total / 1e9, D * reps1 / total1 / 1e9
    log(f"cpu baseline: {rate:.4f} Gstates/s on {threads} threads, {rate1:.4f} on one")
    del csr
    o = None
    # one q end to end
    from exactdiag_amd.allq import final_momentum

    rng = np.random.default_rng(1)
    kf = final_momentum(L, KX, KY, *CPU_Q)
    t0 = time.perf_counter()
    oi = oracle.OracleSector(L, HOLES, NDOWN, KX, KY, threads=threads)
    of = oracle.OracleSector(L, HOLES, NDOWN, *kf, threads=threads)
    t_q_basis = time.perf_counter() - t0  # the final sector's basis (the initial one is the ground state's)
    t0 = time.perf_counter()
    oi.op_csr(of, 0, *CPU_Q, threads=threads)
    t_q_op = time.perf_counter() - t0
    oi = None
    t0 = time.perf_counter()
    of.build_csr((0, 1, 2, 3), threads=threads)
    t_q_csr = time.perf_counter() - t0
    fcsr = [of.term_csr(t) for t in range(4)]
    Df = of.dim
    of = None
    v = rng.normal(size=Df) + 1j * rng.normal(size=Df)  # stands in for O|gs>: the step cost does not depend on it
    v /= np.linalg.norm(v)
    vm = np.zeros(Df, np.complex128)
    bn = 0.0
    t0 = time.perf_counter()
    for _ in range(CPU_Q_STEPS):
        # _lanczos_get_anbn / _lanczos_vec_iteration (cython_lanczos_diagonalization.pyx:50-77): a fresh product
        # vector, numpy updates with their temporaries, then the two vector copies
        hv = np.empty(Df, np.complex128)
        _ref_hv(fcsr, v, hv, threads)
        an = np.real(np.vdot(v, hv))
        vp = hv - an * v - bn * vm
        bp = np.linalg.norm(vp)
        vp /= bp
        vm[:] = v
        v[:] = vp
        bn = bp
    t_step = (time.perf_counter() - t0) / CPU_Q_STEPS
    del fcsr
    per_q = t_q_basis + t_q_op + t_q_csr + ALLQ_STEPS_PLAIN * t_step
    log(f"cpu baseline: one q: basis {t_q_basis:.1f}s, operator CSR {t_q_op:.1f}s, 4-term CSR {t_q_csr:.1f}s, "
        f"{t_step:.3f}s per Lanczos step -> {per_q:.1f}s with {ALLQ_STEPS_PLAIN} steps")
    return {
        "value": rate,
        "unit": "Gstates/s",
        "cores": threads,
        "kind": "reference",
        "host": host,
        "single_core": {"value": rate1, "unit": "Gstates/s", "cores": 1, "reps": reps1, "seconds": total1},
        "product_equals_oracle": same_oracle,
        "product_max_rel_diff_to_gpu": rel_gpu,
        "ground_state_sector": {"basis_s": t_basis, "csr_build_s": t_csr, "hv_s": total / reps},
        "one_q": {"q": list(CPU_Q), "final_basis_s": t_q_basis, "operator_csr_s": t_q_op, "csr_build_s": t_q_csr,
                  "lanczos_step_s": t_step, "steps_timed": CPU_Q_STEPS,
                  "seconds_with_plain_steps": per_q},
        "sample": (f"reference sparse-dot kernels (general/cpp_sparse_dot.h, compiled from the reference's headers) "
                   f"over all {D} rows of the 2x12 h2 k=(0,0) sector's 4-term upper-triangle CSR, explicit + "
                   f"Hermitian passes, {reps} products in {total:.1f}s on {threads} threads; the CSR from the "
                   f"oracle's restatement of calc_sparse_matrix ({t_csr:.1f}s)"),
    }


ALLQ_WS = (0.0, 4.0, 2001)
ALLQ_ETA = 0.05
ALLQ_STEPS_REF = 1000
ALLQ_STEPS_PLAIN = 256
ALLQ_STEPS_FAST = 170  # fixed count: 150 meets 1e-3 only by a few per cent (its FP64 max is 9.0e-4); 170: 5.8e-4
ALLQ_STEPS_CAP = 300  # the fast form's cap when each q stops by the spectrum-change rule
ALLQ_TARGET = 1e-3


def steps_summary(steps):
    v = list(steps.values())
    return {"rule": "per q: stop when the terminated spectrum changes by < 1e-3 (relative L1) over both 10 and 20 steps, "
                    f"checked every 5 steps from step 60, at most {ALLQ_STEPS_CAP}",
            "mean": float(sum(v) / max(len(v), 1)), "max": max(v, default=0), "total": sum(v),
            "per_q": {f"{q[0]},{q[1]}": n for q, n in steps.items()}}


def allq_larger(rungs):
    """The fast and multi-momentum all-q forms on a 2 x rungs ladder with 2 holes (the headline's settings),
    against 1000-step FP64 spectra: the forms at a larger size than the bench line's."""
    import numpy as np

    from exactdiag_amd import allq

    ws = np.linspace(*ALLQ_WS)
    nd = rungs - 1
    for run in (allq.allq_fast, allq.allq_multik):  # kernel loading, untimed
        run(6, HOLES, 5, WEIGHTS, ws, ALLQ_ETA, 40)
    t0 = time.perf_counter()
    log(f"2x{rungs}: all-q reference ({ALLQ_STEPS_REF} steps) ...")
    e0, ref = allq.allq_reference(rungs, HOLES, nd, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_REF, log=log)
    t_ref = time.perf_counter() - t0
    log(f"2x{rungs}: all-q reference ({ALLQ_STEPS_REF} steps) in {t_ref:.1f}s, E0 = {e0:.12f}")
    t_fast, s_fast = allq.allq_fast(rungs, HOLES, nd, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_CAP, spectrum_tol=ALLQ_TARGET)
    fast_breakdown, fast_steps = dict(allq.LAST_FAST_TIMINGS), dict(allq.LAST_FAST_STEPS)
    log(f"2x{rungs}: all-q fast in {t_fast:.2f}s: {fast_breakdown}, steps {fast_steps}")
    t_m, s_m, tm_m = allq.allq_multik(rungs, HOLES, nd, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_FAST)
    log(f"2x{rungs}: all-q over one multi-momentum H in {t_m:.2f}s: {tm_m}")
    return {
        "workload": f"S(q,w) for the {len(ref)} distinct q of the 2x{rungs} ladder, 2 holes, Sz=0, k=(0,0), t=1, "
                    "J=0.4; eta=0.05, w in [0,4] on 2001 points; L1 vs 1000-step FP64",
        "reference_seconds": t_ref,
        "fast": {"seconds": t_fast, "steps": steps_summary(fast_steps), "breakdown_s": fast_breakdown,
                 "max_l1": max(allq.spectral_l1(s_fast[q], ref[q], ws) for q in ref)},
        "multik": {"seconds": t_m, "steps": ALLQ_STEPS_FAST, "breakdown_s": tm_m,
                   "max_l1": max(allq.spectral_l1(s_m[q], ref[q], ws) for q in ref)},
        "best": {"form": "multik" if t_m < t_fast else "fast", "seconds": min(t_m, t_fast)},
    }


def allq_section():
    """All-q S(q, w) of the same ladder (14 distinct q) to <= 1e-3 spectral L1 error: plain FP64 matrix-free
    Lanczos with truncated fraction vs the fast form (stored H, complex64 vectors, square-root terminator);
    both from a fresh start (bases, ground state at the reference's tolerance, every q), against the
    1000-step FP64 spectra."""
    import numpy as np

    from exactdiag_amd import allq

    ws = np.linspace(*ALLQ_WS)
    # every form once on the 2x6 ladder first, so that no timed run pays for loading kernels
    for run in (allq.allq_plain, allq.allq_fast, allq.allq_batched, allq.allq_multik):
        run(6, HOLES, 5, WEIGHTS, ws, ALLQ_ETA, 40)
    t0 = time.perf_counter()
    e0, ref = allq.allq_reference(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_REF)
    log(f"all-q reference ({ALLQ_STEPS_REF} steps) in {time.perf_counter() - t0:.1f}s, E0 = {e0:.12f}")
    t_plain, s_plain = allq.allq_plain(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_PLAIN)
    log(f"all-q plain in {t_plain:.2f}s")
    t_fast, s_fast = allq.allq_fast(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_CAP, spectrum_tol=ALLQ_TARGET)
    fast_breakdown, fast_steps = dict(allq.LAST_FAST_TIMINGS), dict(allq.LAST_FAST_STEPS)
    log(f"all-q fast in {t_fast:.2f}s: {fast_breakdown}, steps {fast_steps}")
    t_ff, s_ff = allq.allq_fast(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_FAST)
    log(f"all-q fast with a fixed {ALLQ_STEPS_FAST} steps in {t_ff:.2f}s")
    t_ps, s_ps = allq.allq_fast(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_PLAIN, precision=0,
                                terminator=False, two_pass_gs=False, real=False)
    log(f"all-q plain method on the stored H in {t_ps:.2f}s")
    t_b, s_b, tm_b = allq.allq_batched(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_FAST)
    log(f"all-q as one batch in {t_b:.2f}s: {tm_b}")
    t_m, s_m, tm_m = allq.allq_multik(L, HOLES, NDOWN, WEIGHTS, ws, ALLQ_ETA, ALLQ_STEPS_FAST)
    log(f"all-q over one multi-momentum H in {t_m:.2f}s: {tm_m}")
    err_plain = {f"{q[0]},{q[1]}": allq.spectral_l1(s_plain[q], ref[q], ws) for q in ref}
    err_fast = {f"{q[0]},{q[1]}": allq.spectral_l1(s_fast[q], ref[q], ws) for q in ref}
    return {
        "workload": "S(q,w) for the 14 distinct q of the 2x12 ladder, 2 holes, Sz=0, k=(0,0), t=1, J=0.4; "
                    "eta=0.05, w in [0,4] on 2001 points; L1 = sum|S-S_ref|dw / sum|S_ref|dw vs 1000-step FP64",
        "target_l1": ALLQ_TARGET,
        "plain": {"seconds": t_plain, "steps": ALLQ_STEPS_PLAIN, "max_l1": max(err_plain.values()),
                  "method": "complex128, matrix-free H·v, ground state by thick-restart Lanczos at tol 2e-7, "
                            "truncated continued fraction, q one after the other"},
        "fast": {"seconds": t_fast, "steps": steps_summary(fast_steps), "max_l1": max(err_fast.values()),
                 "breakdown_s": fast_breakdown,
                 "method": "stored sliced-ELL H per sector, complex64 vectors with FP64 coefficient reductions "
                           "(float32 in the sectors whose H is real, kx = 0 and L/2), square-root terminator; ground "
                           "state by restarted two-pass Lanczos (FP64: float64 vectors in the real k = (0,0) "
                           "sector; no re-orthogonalisation) to the same tolerance; each q's run stops when its "
                           "terminated spectrum has converged (the steps rule), checked on the host while the GPU "
                           "runs the next 5 steps"},
        "fast_fixed_steps": {"seconds": t_ff, "steps": ALLQ_STEPS_FAST,
                             "max_l1": max(allq.spectral_l1(s_ff[q], ref[q], ws) for q in ref),
                             "method": f"the fast form with a fixed {ALLQ_STEPS_FAST} steps for every q"},
        "plain_stored": {"seconds": t_ps, "steps": ALLQ_STEPS_PLAIN,
                         "max_l1": max(allq.spectral_l1(s_ps[q], ref[q], ws) for q in ref),
                         "method": "the plain method (complex128, truncated fraction) on the stored H: the H·v "
                                   "kernel's share of the speed-up"},
        "batched": {"seconds": t_b, "steps": ALLQ_STEPS_FAST,
                    "max_l1": max(allq.spectral_l1(s_b[q], ref[q], ws) for q in ref), "breakdown_s": tm_b,
                    "method": "every q's Lanczos run as one instance of one batched, graph-captured run over stored "
                              "H (complex128 vectors), square-root terminator: the batching form applied to all q"},
        "multik": {"seconds": t_m, "steps": ALLQ_STEPS_FAST,
                   "max_l1": max(allq.spectral_l1(s_m[q], ref[q], ws) for q in ref), "breakdown_s": tm_m,
                   "method": "the fast form with all final sectors in one stored H over the union basis: one "
                             "pass over the matrix serves the 14 momenta, each element gathering one contiguous "
                             "line of the 14 complex64 vectors; Lanczos recurrences advance together"},
        "best": {"form": "multik" if t_m < t_fast else "fast", "seconds": min(t_m, t_fast)},
        "speedup_vs_plain": t_plain / min(t_m, t_fast),
        "l1_plain": err_plain,
        "l1_fast": err_fast,
    }


PUBLISHED_SECONDS = 3 * 3600.0  # "almost three hours to calculate on my 14 cores"


def published_workload():
    """The reference's one published timing, run through the drop-in API on its own configuration: all-q
    S(q,w) (`get_Szq_spectra`, limited_qs) of the 2x12 ladder with 2 holes, k=(0,0), weights tl=0.45,
    tr=0.36, jl=0.15, jr=0.12, 5 eigenpairs, 250 Lanczos vectors, eta=0.05, 1000 w on [0, 0.7]
    (tests/tJ_spin_half_ladder/manual/paulina_spin_susceptibility_visual.py:117-121, config :162-178).
    Cache-cold like the quoted figure (a fresh calculations/ folder: bases, eigenpairs and spectra are
    computed and written), then once more from the disk cache the first run wrote."""
    import shutil
    import tempfile

    import numpy as np
    import torch

    from exactdiag_amd.tJ_spin_half_ladder import api, configs

    def config():
        return configs.Full_Spectrum_Config(
            hamiltonian={"num_rungs": 12, "num_holes": 2, "weights": {"tl": 0.45, "tr": 0.36, "jl": 0.15, "jr": 0.12},
                         "symmetry_qs": {"leg": 0, "rung": 0}, "num_threads": 14},
            spectrum={"name": "Szq", "operator_symmetry_qs": None, "broadening": 5e-2, "omega_max": 0.7,
                      "num_threads": 14},
            eigenpair={"num_eigenpairs": 5, "num_threads": 14})

    cwd = os.getcwd()
    tmp = tempfile.mkdtemp(prefix="ed_published_")
    try:
        os.chdir(tmp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ws, spectra, _ = api.get_Szq_spectra(config(), limited_qs=True)
        torch.cuda.synchronize()
        t_cold = time.perf_counter() - t0
        t0 = time.perf_counter()
        ws2, spectra2, _ = api.get_Szq_spectra(config(), limited_qs=True)
        t_warm = time.perf_counter() - t0
        same = bool(np.array_equal(spectra, spectra2))
    finally:
        os.chdir(cwd)
        shutil.rmtree(tmp, ignore_errors=True)
    log(f"published workload: {t_cold:.2f}s cache-cold, {t_warm:.2f}s from the cache; spectra "
        f"{spectra.shape}, min {spectra.min():.2e}, max {spectra.max():.3f}")
    return {"workload": "api.get_Szq_spectra(limited_qs=True) on the reference's published configuration: 2x12 "
                        "ladder, 2 holes, k=(0,0), tl=0.45 tr=0.36 jl=0.15 jr=0.12, 5 eigenpairs, 250 Lanczos "
                        "vectors, eta=0.05, 1000 w on [0,0.7]; fresh calculations/ folder",
            "seconds_cold": t_cold, "seconds_from_cache": t_warm, "cache_reproduces": same,
            "reference": {"seconds": PUBLISHED_SECONDS, "cores": 14,
                          "source": "'almost three hours to calculate on my 14 cores' "
                                    "(tests/tJ_spin_half_ladder/manual/paulina_spin_susceptibility_visual.py:117-121)"},
            "speedup_vs_published": PUBLISHED_SECONDS / t_cold,
            "spectra_finite_nonnegative": bool(np.all(np.isfinite(spectra)) and spectra.min() > -1e-12)}


BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA


def gemm_section():
    """The closure network's BF16 MFMA GEMM (ed_gemm_bf16): its products at the training batch (256 rows) and a
    large square product, timed with HIP events on the launch stream, against the dense BF16 peak; and one
    training step of the network (forward, backward, Adam) at batch 256."""
    import ctypes as ct

    import numpy as np
    import torch

    from exactdiag_amd import closure
    from exactdiag_amd._lib import check, lib

    stream = torch.cuda.current_stream()
    sp = ct.c_void_p(stream.cuda_stream)
    out = {}
    for name, M, N, K, reps in (("layer 256x256, batch 256", 256, 256, 256, 200),
                                ("output layer 256x384, batch 256", 256, 384, 256, 200),
                                ("square 8192", 8192, 8192, 8192, 10)):
        A = torch.randint(0x3c00, 0x3f80, (M, K), dtype=torch.int16, device="cuda")  # bf16 values in [2^-7, 1)
        B = torch.randint(0x3c00, 0x3f80, (K, N), dtype=torch.int16, device="cuda")
        C = torch.empty((M, N), dtype=torch.float32, device="cuda")
        args = (M, N, K, ct.c_void_p(A.data_ptr()), K, 1, ct.c_void_p(B.data_ptr()), N, 1, ct.c_void_p(C.data_ptr()),
                N, sp)
        for _ in range(3):
            check(lib().ed_gemm_bf16(*args))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            check(lib().ed_gemm_bf16(*args))
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tflops = 2.0 * M * N * K / (ms * 1e-3) / 1e12
        out[name] = {"M": M, "N": N, "K": K, "ms": ms, "tflops": tflops, "frac": tflops / BF16_PEAK_TFLOPS}
        log(f"gemm {name}: {ms * 1e3:.1f} us, {tflops:.1f} TFLOP/s ({tflops / BF16_PEAK_TFLOPS:.3f} of peak)")
        del A, B, C
    net = closure.GpuMLP(closure.DIMS, seed=0)
    rng = np.random.default_rng(0)
    x = torch.as_tensor(rng.normal(size=(256, closure.DIMS[0])).astype(np.float32), device="cuda")
    t = torch.as_tensor(rng.normal(size=(256, closure.DIMS[-1])).astype(np.float32), device="cuda")
    for _ in range(3):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(50):
        net.train_step(x, t, 1e-4)
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / 50 * 1e3
    big = out["square 8192"]
    return {"kernel": "gemm_bf16_tiled_kernel (256x256 tiles of eight wavefronts for the square product, 128x128 "
                      "of four for the network's layers; two LDS buffers, 16-byte loads, v_mfma_f32_32x32x16_bf16)",
            "products": out, "train_step_ms_batch256": step_ms,
            "roofline": {"bound": "mfma", "achieved": big["tflops"], "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": big["frac"], "traffic": None, "shape": "8192^3"}}


def forms_section():
    """Compact results of the three added forms at their benchmark configurations: batching (2x10, 60
    instances), the learned closure (2x12, all q, 64 levels) with its GEMM rate, and precision (2x14)."""
    import numpy as np

    from exactdiag_amd import closure, forms

    out = {}
    t0 = time.perf_counter()
    b = forms.batching_form(log=log)
    mk = b.get("multik", {})
    out["batching"] = {
        "workload": "2x10 ladder, 2 holes, Sz=0, t=1, J in {0.3,0.4,0.5}, all 20 k sectors (60 instances of 461,890 "
                    "states): lowest eigenpairs, then A+(k,w) and A-(k,w) for all 20 k, 250 steps, FP64, stored H",
        "lowest_eigenpairs_s": {"batched": b["gs_batched_s"], "serial": b.get("gs_serial_s"),
                                "multik": mk.get("gs_s")},
        "spectra_s": {"batched": b["spectra_batched_s"], "serial": b.get("spectra_serial_s"),
                      "multik": mk.get("spectra_s")},
        "max_abs_energy_diff_batched_vs_serial": b.get("gs_max_abs_diff"),
        "tuned": b["tuned"], "seconds_total": time.perf_counter() - t0}
    if closure.MODEL_PATH.exists():
        t0 = time.perf_counter()
        per_q, t_net = closure.evaluate_allq(closure.Closure.load(), L, HOLES, NDOWN, WEIGHTS,
                                             np.linspace(*ALLQ_WS), ALLQ_ETA)
        errs = np.array(list(per_q.values()))
        out["closure"] = {
            "workload": "2x12 ladder, 2 holes, k=(0,0), t=1, J=0.4: S(q,w) of the 13 distinct q with weight at 64 "
                        "Lanczos levels closed by the network vs by the square-root terminator; L1 vs 1000-step FP64",
            "mean_l1": {"terminator": float(errs[:, 0].mean()), "learned": float(errs[:, 1].mean())},
            "max_l1": {"terminator": float(errs[:, 0].max()), "learned": float(errs[:, 1].max())},
            "geometric_gain": float(np.exp(np.mean(np.log(errs[:, 0] / errs[:, 1])))),
            "network_seconds": t_net, "seconds_total": time.perf_counter() - t0}
        log(f"closure on 2x12: {out['closure']['mean_l1']}, gain {out['closure']['geometric_gain']:.2f}")
    out["closure_gemm"] = gemm_section()
    t0 = time.perf_counter()
    p = forms.precision_form(log=log, keep_spectra=True)
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    pf = forms.precision_form(log=log, flip=True, keep_spectra=True)
    t_flip = time.perf_counter() - t0
    from exactdiag_amd.allq import spectral_l1

    ws = np.linspace(0, 4, 2001)
    out["precision_spin_flip"] = {
        "workload": "the same 2x14 sector split by the optional spin-flip symmetry (no counterpart in the "
                    "reference): ground state in the lower parity block, S(q=(pi,pi),w) in the opposite block",
        "dims": pf["dims"], "block_dims": pf["flip"]["block_dims"], "ground_state_parity": pf["flip"]["ground_state_parity"],
        "e0_abs_diff_vs_momentum_sector": abs(pf["c128"]["e0"] - p["c128"]["e0"]),
        "spectrum_l1_vs_momentum_sector_c128": spectral_l1(pf["spectra"]["c128"], p["spectra"]["c128"], ws),
        "e0_rel_error_c64": pf["e0_rel_error_c64"], "spectrum_l1_error_c64": pf["spectrum_l1_error_c64"],
        "seconds": {"c128": pf["c128"]["gs_seconds"] + pf["c128"]["lanczos_seconds"],
                    "c64": pf["c64"]["gs_seconds"] + pf["c64"]["lanczos_seconds"]},
        "stored_gb": pf["stored_gb"], "seconds_total": t_flip}
    log(f"precision with the spin flip: dims {pf['dims']}, |dE0| vs the momentum sector "
        f"{out['precision_spin_flip']['e0_abs_diff_vs_momentum_sector']:.1e}, S L1 "
        f"{out['precision_spin_flip']['spectrum_l1_vs_momentum_sector_c128']:.1e}")
    out["precision"] = {
        "workload": "2x14 ladder, 2 holes, Sz=0, k=(pi,0), t=1, J=0.4 (D=140,408,100); S(q=(pi,pi),w), 300 steps",
        "e0": {"c128": p["c128"]["e0"], "c64": p["c64"]["e0"]}, "e0_rel_error_c64": p["e0_rel_error_c64"],
        "spectrum_l1_error_c64": p["spectrum_l1_error_c64"],
        "spectrum_max_rel_error_c64": p["spectrum_max_rel_error_c64"],
        "seconds": {"c128": p["c128"]["gs_seconds"] + p["c128"]["lanczos_seconds"],
                    "c64": p["c64"]["gs_seconds"] + p["c64"]["lanczos_seconds"]},
        "stored_gb": p["stored_gb"], "seconds_total": t_plain}
    return out


def cpu_allq(cpu, section):
    """Reference CPU path for the same all-q run from the measured pieces: the ground-state sector's basis and
    CSR, ~200 products for ARPACK's ground state, then per distinct q the measured final-sector basis, operator
    CSR, 4-term CSR and the plain method's Lanczos steps at the measured step time (one q timed; the other
    sectors have the same size)."""
    if not cpu.get("value"):
        return None
    gs = cpu["ground_state_sector"]
    n_q = len(section["l1_plain"])
    gs_s = gs["basis_s"] + gs["csr_build_s"] + 200 * gs["hv_s"]
    total = gs_s + n_q * cpu["one_q"]["seconds_with_plain_steps"]
    return {"seconds": total, "ground_state_s": gs_s, "per_q_s": cpu["one_q"]["seconds_with_plain_steps"],
            "n_q": n_q, "cores": cpu["cores"],
            "method": "measured per stage on this host: one q run end to end with the reference's compiled dot "
                      f"({CPU_Q_STEPS} Lanczos steps timed, the plain method's {ALLQ_STEPS_PLAIN} at that step "
                      "time), ARPACK's ~200 products at the measured product time; disk I/O and the reference's "
                      "Python overhead not counted",
            "speedup_fast_vs_cpu": total / section["fast"]["seconds"],
            "speedup_best_vs_cpu": total / section["best"]["seconds"],
            "speedup_plain_gpu_vs_cpu": total / section["plain"]["seconds"]}


CSR_CPU_SAMPLE_ROWS = None  # None: every row (about 15 s on 16 cores for the 2x12 sector)


def cpu_csr_build(mats, threads):
    """The oracle's restatement of calc_sparse_matrix over rows [0, CSR_CPU_SAMPLE_ROWS) of each term (all
    rows by default): its time (scaled to the whole sector when sampled) and whether the GPU's arrays
    reproduce it (the whole arrays, or their prefix for a sample)."""
    import numpy as np

    from oracle import oracle

    o = oracle.OracleSector(L, HOLES, NDOWN, KX, KY, threads=threads)
    rows = o.dim if CSR_CPU_SAMPLE_ROWS is None else min(CSR_CPU_SAMPLE_ROWS, o.dim)
    t_terms, same = [], True
    for t in range(4):
        t1 = time.perf_counter()
        o.build_csr_rows(t, 0, rows, threads=threads)
        t_terms.append(time.perf_counter() - t1)
        vo, io, mo, po = o.term_csr(t)
        vg, ig, mg, pg = mats[t].arrays()
        if po[-1] == 1 and not po[:-1].any() and len(vo) == 1 and vo[0] == 0:
            same &= bool(pg[rows] == 0)  # no elements in the sample rows (the oracle's empty-matrix marker)
        else:
            n = int(po[-1])
            same &= bool(np.array_equal(pg[: rows + 1], po) and np.array_equal(mg[:n], mo)
                         and np.array_equal(ig[:n], io) and np.array_equal(vg[: len(vo)], vo))
        log(f"cpu csr build: term {t} rows [0, {rows}) in {t_terms[-1]:.1f}s; GPU prefix equal: {same}")
    est = sum(t_terms) * o.dim / rows
    scaled = "" if rows == o.dim else f", scaled by D / {rows}"
    return {"value": est, "unit": "s", "cores": threads, "kind": "port",
            "sample": (f"oracle restatement of calc_sparse_matrix (columns, unique-value walk, CSR) for rows "
                       f"[0, {rows}) of each of the four terms (D = {o.dim}), {sum(t_terms):.1f}s{scaled}")}, same


def cpu_positions(states, res, threads):
    """The oracle's restatement of calculate_hole_spin_projection_correlations on the same 2x12 ground state
    (single-threaded, as the reference's loop), and of the singlet operator's columns for one shift."""
    import numpy as np

    from oracle import oracle

    s, x = states["holes"]
    o = oracle.OracleSector(12, 2, 11, 0, 0, threads=threads)
    t0 = time.perf_counter()
    corr = o.position_correlations(0, [], x)
    t_holes = time.perf_counter() - t0
    got = np.asarray(res["holes_2x12"]["correlations"])
    same = bool(np.max(np.abs(corr / corr[0] - got)) < 1e-10)
    log(f"cpu positions: holes 2x12 in {t_holes:.1f}s; equal to the GPU to 1e-10: {same}")
    s4, x4 = states["singlet"]
    o4 = oracle.OracleSector(10, 4, 8, 0, 0, threads=threads)
    t0 = time.perf_counter()
    ptr, rows, vals = o4.position_columns(2, [(1, 1), (1, 1), (0, 0)], threads=threads)
    cols = np.repeat(np.arange(o4.dim), np.diff(ptr))
    e = np.sum(np.conj(x4[rows]) * vals * x4[cols]).real
    t_singlet = time.perf_counter() - t0
    same_s = bool(abs(e - res["singlet_2x10_h4"]["correlations"][0]) < 1e-10)
    log(f"cpu positions: singlet 2x10 h4 shift 0 in {t_singlet:.1f}s; equal: {same_s}")
    return ({"value": t_holes * 1e3, "unit": "ms", "cores": 1, "kind": "port",
             "sample": (f"oracle calculate_hole_spin_projection_correlations on the same 2x12 ground state, all 24 "
                        f"distances, single-threaded ({t_holes:.1f}s); singlet-singlet of the 2x10 h4 state, one "
                        f"shift: {t_singlet:.1f}s on {threads} threads (GPU: "
                        f"{res['singlet_2x10_h4']['per_shift_ms']:.2f} ms)")}, same and same_s)


def run_form(form, args=None):
    import torch

    from exactdiag_amd import forms

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    if form == "positions":
        res, states = forms.positions_form(log=log)
        h = res["holes_2x12"]
        line = {"metric": "hole correlations of the 2x12 ground state, all 24 distances (ms per call)",
                "value": h["kernel_ms"], "unit": "ms", "n_gpus": 1, "higher_is_better": False,
                "dtype": "complex128", "data": "synthetic (model built from its definition)",
                "config": {"workload": "2x12 ladder, 2 holes, Sz=0, k=(0,0), t=1, J=0.4 ground state (D=8112468); "
                                       "singlet-singlet of the 2x10, 4-hole ground state"},
                "roofline": {"bound": "hbm", "achieved": h["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": h["gbs"] / HBM_PEAK_GBS, "traffic": None},
                "form": res}
        if args is None or not args.no_cpu_baseline:
            threads = host_cpu()["threads"]
            line["cpu_baseline"], line["matches_oracle"] = cpu_positions(states, res, threads)
    elif form == "csr":
        res, mats = forms.csr_form(L, HOLES, NDOWN, (KX, KY), WEIGHTS, log=log)
        line = {"metric": "GPU build of the reference's stored CSR matrices (four Hamiltonian terms), seconds",
                "value": res["build_s"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
                "dtype": "complex128", "data": "synthetic (model built from its definition)",
                "config": {"workload": "2x12 ladder, 2 holes, Sz=0, k=(0,0), t=1, J=0.4: tl, tr, jl, jr as "
                                       "row-major Hermitian upper triangles with unique-value tables"},
                "form": res}
        if args is None or not args.no_cpu_baseline:
            threads = host_cpu()["threads"]
            cpu, same = cpu_csr_build(mats, threads)
            line["cpu_baseline"] = cpu
            line["arrays_match_oracle"] = same
            line["speedup_vs_cpu"] = cpu["value"] / res["build_s"]
        for m in mats:
            m.close()
    elif form == "allq" and args is not None and args.allq_rungs and args.allq_rungs != L:
        res = allq_larger(args.allq_rungs)
        line = {"metric": "time to all-q S(q,w) at <= 1e-3 spectral L1 (seconds, best form)",
                "value": res["best"]["seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
                "dtype": "complex64", "data": "synthetic (model built from its definition)",
                "config": {"workload": res["workload"]}, "allq": res}
    elif form == "allq":
        res = allq_section()
        line = {"metric": "time to all-q S(q,w) at <= 1e-3 spectral L1 (seconds, best form)",
                "value": res["best"]["seconds"], "unit": "s", "n_gpus": 1, "higher_is_better": False,
                "dtype": "complex64", "data": "synthetic (model built from its definition)",
                "config": {"workload": res["workload"]}, "allq": res}
    elif form == "multik":
        res = forms.multik_form(log=log)
        c64 = res["c64"]
        states = sum(res["sector_dims"])
        achieved = 28 * states / c64["multik_ms"] / 1e6
        line = {"metric": "one stored H·v over the 14 final sectors of all-q S(q,w) (ms per product, complex64)",
                "value": c64["multik_ms"], "unit": "ms", "n_gpus": 1, "higher_is_better": False,
                "dtype": "complex64", "data": "synthetic (model built from its definition)",
                "config": {"workload": "2x12 ladder, 2 holes, Sz=0, t=1, J=0.4: the 14 momenta k = q of the "
                                       "distinct q, random vectors, one product for all vs one stored H per sector"},
                "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS,
                             "traffic": traffic_from_profiles("mk_spmv", "multik_2x12_h2.md"),
                             "kernel": "mk_spmv_kernel, per launch (HIP events)", "alg_bytes_per_state": 28},
                "form": res}
    elif form == "gemm":
        res = gemm_section()
        big = res["products"]["square 8192"]
        line = {"metric": "closure BF16 MFMA GEMM, 8192^3 (TFLOP/s)", "value": big["tflops"], "unit": "TFLOP/s",
                "n_gpus": 1, "higher_is_better": True, "dtype": "bf16 (fp32 accumulation)",
                "data": "synthetic (random BF16 values in [2^-7, 1))",
                "config": {"workload": "C[8192][8192] = A[8192][8192] B[8192][8192], A k-contiguous, B n-contiguous; "
                                       "and the closure network's 256-row layers and one training step"},
                "roofline": res["roofline"], "form": res}
    elif form == "batching":
        res = forms.batching_form(log=log)
        mk = res.get("multik", {})
        line = {"metric": "batched all-k lowest eigenpairs + all-k A(k,w) (seconds, batched vs one sector at a time)",
                "value": res["gs_batched_s"] + res["spectra_batched_s"], "unit": "s", "n_gpus": 1,
                "higher_is_better": False, "dtype": "complex128", "data": "synthetic (model built from its definition)",
                "config": {"workload": "2x10 ladder, 2 holes, Sz=0, t=1, J in {0.3,0.4,0.5}, all 20 k sectors "
                                       "(60 instances); A+(k,w), A-(k,w) for all 20 k, 250 steps, FP64; stored H on "
                                       "both sides"},
                "value_multik": mk.get("gs_s", 0.0) + mk.get("spectra_s", 0.0),
                "form": res}
    else:
        res = forms.precision_form(log=log)
        line = {"metric": "complex64 Lanczos: ground-state energy and S(q,w) error vs complex128",
                "value": res["spectrum_l1_error_c64"], "unit": "spectral L1", "n_gpus": 1, "higher_is_better": False,
                "dtype": "complex64", "data": "synthetic (model built from its definition)",
                "config": {"workload": "2x14 ladder, 2 holes, Sz=0, k=(pi,0), t=1, J=0.4; S(q=(pi,pi),w), "
                                       "300 steps, eta=0.05"},
                "form": res}
    line["seconds_total"] = time.perf_counter() - t0
    print(json.dumps(line), flush=True)


def read_rate(traffic_bytes, ms_per_launch):
    """The device's own read ceiling for the stored product (ed_hmat_line_probe): 128-byte lines per second for
    16-byte loads from a 128 MiB buffer (the size of the 2x12 sector's complex128 x), scattered one line per lane
    as the far gathers are, and 1 KiB in a row per wavefront; against it the lines the product moves per launch
    (its calibrated L2-miss traffic / 128 B) over its event time."""
    import ctypes as ct

    import torch

    from exactdiag_amd._lib import check, lib

    sp = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, scattered in (("scattered", 1), ("contiguous", 0)):
        r = ct.c_double(0.0)
        check(lib().ed_hmat_line_probe(scattered, 128 << 20, ct.byref(r), sp))
        out[f"probe_{name}_lines_per_s"] = r.value
        out[f"probe_{name}_GBs"] = r.value * 128 / 1e9
    if traffic_bytes:
        lines = traffic_bytes / 128.0
        out["kernel_lines_per_launch"] = lines
        out["kernel_lines_per_s"] = lines / (ms_per_launch * 1e-3)
        out["frac_of_scattered_probe"] = out["kernel_lines_per_s"] / out["probe_scattered_lines_per_s"]
    log(f"read rate: probe scattered {out['probe_scattered_GBs']:.0f} GB/s, contiguous "
        f"{out['probe_contiguous_GBs']:.0f} GB/s; product {out.get('frac_of_scattered_probe', float('nan')):.3f} "
        "of the scattered rate")
    return out


def time_steps(step, steps, warmup, stream):
    """W untimed steps, then exactly K steps between synchronisations; (wall s, HIP-event ms) of the K."""
    import torch

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    start.record(stream)
    for _ in range(steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    return time.perf_counter() - t1, start.elapsed_time(end)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--no-allq", action="store_true", help="skip the all-q time-to-accuracy section")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the reference's published all-q workload run through the drop-in API")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batching, closure (with its GEMM rate) and precision results")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy one vector (known bytes) for PMC calibration")
    ap.add_argument("--allq-rungs", type=int, default=0,
                    help="with --form allq: the ladder's rungs (default: the headline's 12); others run the fast "
                         "and multi-momentum forms only")
    ap.add_argument("--form", choices=["hv", "batching", "precision", "csr", "positions", "multik", "allq", "gemm"],
                    default="hv",
                    help="hv: the headline H·v + all-q line; allq: the all-q section alone; batching / precision / csr / "
                         "positions / multik / gemm: their own line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); several GPUs are not supported: "
              f"independent q values or sectors would be distributed over GPUs, one sector per GPU", file=sys.stderr)
        sys.exit(2)
    if args.form != "hv":
        run_form(args.form, args)
        return

    import ctypes as ct

    import numpy as np
    import torch

    from exactdiag_amd._lib import check, lib
    from exactdiag_amd.sector import Sector, weights_array

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    sec = Sector(L, HOLES, NDOWN, KX, KY)
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - t0
    D = sec.dim
    log(f"basis D={D} built in {t_basis:.3f}s")
    t0 = time.perf_counter()
    hm = sec.store(WEIGHTS)
    torch.cuda.synchronize()
    t_store = time.perf_counter() - t0
    log(f"stored H: {hm.nnz} elements ({hm.nnz / D:.2f} per row), {hm.slots} slots, {hm.n_values} unique values, "
        f"{hm.bytes / 1e9:.3f} GB, built in {t_store:.3f}s")

    rng = np.random.default_rng(0)
    dt = torch.complex128 if args.precision == "c128" else torch.complex64
    x = torch.as_tensor(rng.normal(size=D) + 1j * rng.normal(size=D), device="cuda").to(dt)
    y = torch.empty_like(x)
    dot = torch.zeros(1, dtype=torch.complex128, device="cuda")
    w = weights_array(WEIGHTS)
    wp = w.ctypes.data_as(ct.POINTER(ct.c_double))
    prec = 0 if dt == torch.complex128 else 1
    stream = torch.cuda.current_stream()
    sp = ct.c_void_p(stream.cuda_stream)
    xp, yp, dp = ct.c_void_p(x.data_ptr()), ct.c_void_p(y.data_ptr()), ct.c_void_p(dot.data_ptr())

    def step_stored():
        check(lib().ed_hmat_hv(hm._h, xp, yp, dp, prec, sp))

    def step_matrix_free():
        check(lib().ed_hv(sec.handle, wp, xp, yp, dp, prec, sp))

    wall_mf, ev_mf = time_steps(step_matrix_free, args.steps, args.warmup, stream)
    y_mf = y.clone()
    wall, ev = time_steps(step_stored, args.steps, args.warmup, stream)
    agree = float((y - y_mf).abs().max() / y_mf.abs().max())
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(x.to(torch.complex128).cpu().numpy(), y.to(torch.complex128).cpu().numpy(),
                               host_cpu()["threads"])
        except Exception as e:  # the GPU figure stands on its own
            cpu = {"value": None, "error": str(e)}
    ms_per_step = ev / args.steps
    value = D * args.steps / wall / 1e9
    s_bytes = 16 if prec == 0 else 8
    alg = (2 * s_bytes + 12) * D
    achieved = alg / (ms_per_step * 1e-3) / 1e9
    # compulsory bytes of the stored form: x, y, the stored elements (4 or 6 bytes), count + diagonal index + orbit
    alg_stored = (2 * s_bytes + 4) * D + hm.element_bytes * hm.nnz
    line = {
        "metric": "t-J H·v Gstates/s (% HBM peak); time to all-q S(q,ω) at 1e-3 vs plain",
        "value": value,
        "unit": "Gstates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128" if prec == 0 else "complex64",
        "data": "synthetic (seeded complex normal vectors, model built from its definition)",
        "config": {"workload": "t-J ladder 2x12, 2 holes, Sz=0, k=(0,0), t=1, J=0.4, one H·v per step "
                               "(stored sliced-ELL H, built once before the timed region)",
                   "dim": D, "basis_build_s": round(t_basis, 4), "stored_build_s": round(t_store, 4),
                   "stored_bytes": hm.bytes, "stored_elements_per_row": round(hm.nnz / D, 3),
                   "stored_element_bytes": hm.element_bytes,
                   "stored_config": {"tiles_per_workgroup": 1, "gathers_in_flight": 4,
                                     "nontemporal_stream": prec == 0,
                                     "note": "the autotuner's choice for this sector (ed_hmat_tune), built in as "
                                             "the default"}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profiles("hm_spmv") if prec == 0 else None,
                     "kernel": "hm_spmv_kernel (+ reduce_partials_kernel), per launch (HIP events)",
                     "alg_bytes_per_state": 2 * s_bytes + 12, "event_ms_per_launch": ms_per_step,
                     "achieved_incl_matrix": alg_stored / (ms_per_step * 1e-3) / 1e9,
                     "frac_incl_matrix": alg_stored / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "read_rate": read_rate(traffic_from_profiles("hm_spmv") if prec == 0 else None, ms_per_step)},
        "matrix_free": {"value": D * args.steps / wall_mf / 1e9, "unit": "Gstates/s",
                        "event_ms_per_launch": ev_mf / args.steps,
                        "frac": alg / (ev_mf / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "traffic": traffic_from_profiles("hv_lds") if prec == 0 else None,
                        "kernel": "hv_lds_kernel (+ reduce_partials_kernel)",
                        "max_rel_diff_vs_stored": agree},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu
    if args.copy_calibration:
        for _ in range(3):
            y.copy_(x)  # reads and writes D * itemsize bytes
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = D * x.element_size()
    if not args.no_allq and prec == 0:
        del hm, x, y, y_mf
        torch.cuda.empty_cache()
        line["allq"] = allq_section()
    if "cpu_baseline" in line and line.get("allq"):
        line["allq"]["cpu"] = cpu_allq(line["cpu_baseline"], line["allq"])
    if not args.no_forms and prec == 0:
        torch.cuda.empty_cache()
        line["forms"] = forms_section()
    if not args.no_published and prec == 0:
        torch.cuda.empty_cache()
        line["published"] = published_workload()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
