"""Benchmark: matrix-free t-J H·v throughput on one MI355X.

Workload: 2x12 ladder (24 sites, periodic along the legs), 2 holes, S^z = 0
(11 down spins), k = (0, 0) even-parity sector, t_leg = t_rung = 1,
J_leg = J_rung = 0.4, complex128 vectors; D = 8,112,468 states.  One step is
one H·v of the whole sector with the fused <x, Hx> reduction the Lanczos
recurrence uses (the hot path of every Lanczos step).  Inputs are seeded
synthetic vectors already resident in HBM.

Prints one JSON line with the driver's fields plus ``roofline`` (the H·v
kernel against the 8 TB/s HBM peak, timed with HIP events on the launch
stream) and ``cpu_baseline`` (the reference's own sparse-dot kernels, from
oracle/_ref, timed on the host's cores over a bounded row sample of the same
sector's stored matrix).
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
CPU_SAMPLE_ROWS = 1_000_000


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per H·v launch from the committed rocprofv3 PMC summary, if any."""
    p = ROOT / "profiles" / "hv_2x12_h2_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(threads):
    """Reference kernels (oracle/_ref) over rows [0, CPU_SAMPLE_ROWS) of the stored 4-term CSR."""
    import numpy as np

    from oracle import oracle

    if not oracle.ref_available():
        kind = "port"
    else:
        kind = "reference"
    t0 = time.perf_counter()
    o = oracle.OracleSector(L, HOLES, NDOWN, KX, KY, threads=threads)
    D = o.dim
    rows = min(CPU_SAMPLE_ROWS, D)
    csr = []
    for t in range(4):
        o.build_csr_rows(t, 0, rows, threads=threads)
        csr.append(o.term_csr(t))
    log(f"cpu baseline: built sample CSR in {time.perf_counter() - t0:.1f}s")
    rng = np.random.default_rng(0)
    x = rng.normal(size=D) + 1j * rng.normal(size=D)
    y = np.zeros(D, np.complex128)

    def once():
        for t in range(4):
            values, vinds, minor, major = csr[t]
            if kind == "reference":
                oracle.ref_csr_dot_rows(D, 0, values, vinds, minor, major, x, y, WEIGHTS[t], threads=threads)
            else:
                raise RuntimeError("oracle/_ref missing: no CPU sparse-dot kernel to time")

    once()
    reps, t_total = 0, 0.0
    while t_total < 10.0 and reps < 200:
        t1 = time.perf_counter()
        once()
        t_total += time.perf_counter() - t1
        reps += 1
    rate = rows * reps / t_total / 1e9
    return {
        "value": rate,
        "unit": "Gstates/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"reference sparse-dot kernels (general/cpp_sparse_dot.h) over rows [0, {rows}) of the "
                   f"2x12 h2 k=(0,0) sector's stored 4-term upper-triangle CSR (explicit + Hermitian passes), "
                   f"{reps} repetitions, {t_total:.1f}s"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy one vector (known bytes) for PMC calibration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part of "
              f"this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from exactdiag_amd.sector import Sector

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    sec = Sector(L, HOLES, NDOWN, KX, KY)
    torch.cuda.synchronize()
    t_basis = time.perf_counter() - t0
    D = sec.dim
    log(f"basis D={D} built in {t_basis:.3f}s")
    rng = np.random.default_rng(0)
    dt = torch.complex128 if args.precision == "c128" else torch.complex64
    x = torch.as_tensor(rng.normal(size=D) + 1j * rng.normal(size=D), device="cuda").to(dt)
    y = torch.empty_like(x)
    dot = torch.zeros(1, dtype=torch.complex128, device="cuda")
    import ctypes as ct

    from exactdiag_amd._lib import check, lib
    from exactdiag_amd.sector import weights_array

    w = weights_array(WEIGHTS)
    wp = w.ctypes.data_as(ct.POINTER(ct.c_double))
    prec = 0 if dt == torch.complex128 else 1
    stream = torch.cuda.current_stream()
    sp = ct.c_void_p(stream.cuda_stream)

    def step():
        check(lib().ed_hv(sec.handle, wp, ct.c_void_p(x.data_ptr()), ct.c_void_p(y.data_ptr()),
                          ct.c_void_p(dot.data_ptr()), prec, sp))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    ms_event = start.elapsed_time(end)
    ms_per_step = ms_event / args.steps
    value = D * args.steps / (wall) / 1e9
    s_bytes = 16 if prec == 0 else 8
    alg = (2 * s_bytes + 12) * D
    achieved = alg / (ms_per_step * 1e-3) / 1e9
    traffic = traffic_from_profiles() if prec == 0 else None
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
        "config": {"workload": "t-J ladder 2x12, 2 holes, Sz=0, k=(0,0), t=1, J=0.4, one H·v per step",
                   "dim": D, "basis_build_s": round(t_basis, 4)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "hv_kernel + reduce_partials_kernel, per launch (HIP events)",
                     "alg_bytes_per_state": 2 * s_bytes + 12, "event_ms_per_launch": ms_per_step},
    }
    if args.copy_calibration:
        for _ in range(3):
            y.copy_(x)  # reads and writes D * itemsize bytes
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = D * x.element_size()
    if not args.no_cpu_baseline:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
            line["cpu_baseline"] = cpu_baseline(threads)
        except Exception as e:  # the GPU figure stands on its own
            line["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
