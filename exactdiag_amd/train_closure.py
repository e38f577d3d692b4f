"""Generate the closure's training data on the GPU, train it, evaluate it.

    python -m exactdiag_amd.train_closure --data out/seqs.npz          # generate 1000-step sequences
    python -m exactdiag_amd.train_closure --data out/seqs.npz --train  # train, save the model
    python -m exactdiag_amd.train_closure --eval out/eval.npz           # 2x12, 2 holes, all q (J = 0.4)

Training systems: 2xL ladders (L = 6, 8, 10, 12), 0-4 holes (0-2 at L = 10
and 12), S^z = 0 or 1/2, J drawn uniformly from [0.2, 1.0] (t = 1), the
k = (0, 0) ground state, every q (qx <= L/2) of S^z(q), N(q) and A+/A-
(S^z(q) only for 2x12 with 2 holes).
"""

from __future__ import annotations

import argparse
import json
import pathlib
import sys
import time

import numpy as np

from exactdiag_amd import closure
from exactdiag_amd.allq import spectral_l1


def systems(seed=0, n_j=3):
    """(L, h, nd, J, operators) of the training set."""
    rng = np.random.default_rng(seed)
    out = []
    for L, hmax, nj in ((6, 4, n_j), (8, 4, n_j), (10, 2, n_j), (12, 2, 2)):
        for h in range(hmax + 1):
            ne = 2 * L - h
            nd = ne // 2
            ops = ("Szq",) if (L == 12 and h == 2) else ("Szq", "Nq", "A+", "A-")
            for J in rng.uniform(0.2, 1.0, size=nj):
                out.append((L, h, nd, float(J), ops))
    return out


def generate(path, nsteps=1000, seed=0):
    """Every system of ``systems``; the file is rewritten after each system, a failing system is reported and
    skipped."""
    seqs = []
    t0 = time.perf_counter()
    for L, h, nd, J, ops in systems(seed):
        w = [1.0, 1.0, J, J]
        try:
            s = closure.lanczos_sequences(L, h, nd, w, nsteps=nsteps, ops=ops)
        except (RuntimeError, ValueError) as exc:
            print(f"L={L} h={h} J={J:.3f}: skipped ({exc})", flush=True)
            continue
        seqs.extend(s)
        save_sequences(path, seqs)
        print(f"L={L} h={h} J={J:.3f}: {len(s)} sequences, total {len(seqs)}, {time.perf_counter() - t0:.1f}s",
              flush=True)
    return seqs


def generate_eval(path, nsteps=1000):
    """The held-out evaluation set: 2x12 ladder, 2 holes, S^z = 0, t = 1, J = 0.4, S^z(q) for every distinct q."""
    t0 = time.perf_counter()
    seqs = closure.lanczos_sequences(12, 2, 11, [1.0, 1.0, 0.4, 0.4], nsteps=nsteps, ops=("Szq",))
    save_sequences(path, seqs)
    print(f"eval set: {len(seqs)} sequences in {time.perf_counter() - t0:.1f}s", flush=True)
    return seqs


def save_sequences(path, seqs):
    n = max(len(s["a"]) for s in seqs)
    A = np.zeros((len(seqs), n))
    B = np.zeros((len(seqs), n))
    lens = np.array([len(s["a"]) for s in seqs])
    for i, s in enumerate(seqs):
        A[i, : len(s["a"])] = s["a"]
        B[i, : len(s["b"])] = s["b"]
    meta = np.array([(s["L"], s["h"], s["nd"], s["J"], ["Szq", "Nq", "A+", "A-"].index(s["op"]), s["qx"], s["qy"],
                      s["e0"], s["norm2"], s["dim"]) for s in seqs])
    np.savez(path, a=A, b=B, lens=lens, meta=meta)


def load_sequences(path):
    d = np.load(path, allow_pickle=False)
    out = []
    for i in range(len(d["lens"])):
        n = int(d["lens"][i])
        m = d["meta"][i]
        out.append(dict(a=d["a"][i, :n], b=d["b"][i, :n], L=int(m[0]), h=int(m[1]), J=m[3], op=int(m[4]),
                        qx=int(m[5]), qy=int(m[6]), e0=m[7], norm2=m[8], dim=int(m[9])))
    return out


def gain(err_without, err_with):
    """Geometric mean of err_without / err_with over the spectra where both are nonzero."""
    e0, e1 = np.asarray(err_without), np.asarray(err_with)
    ok = (e0 > 0) & (e1 > 0) & np.isfinite(e0) & np.isfinite(e1)
    return float(np.exp(np.mean(np.log(e0[ok] / e1[ok])))) if ok.any() else float("nan")


def spectra_errors(seqs, model, ws, eta, n_exact=closure.N_IN):
    """L1 errors of the closed 64-step spectra against the full sequences' spectra."""
    from exactdiag_amd.sector import cf_eval

    rows = []
    for s in seqs:
        a, b = s["a"], s["b"]
        f = -s["norm2"] / np.pi
        ref = cf_eval(a, b, ws, s["e0"], eta, f)
        trunc = cf_eval(a[:n_exact], b[:n_exact], ws, s["e0"], eta, f)
        at, bt, tail = closure.sqrt_terminator(a, b, n_exact)
        term = cf_eval(at, bt, ws, s["e0"], eta, f, tail=tail)
        ae, be, tail2 = model.extend(a, b)
        learned = cf_eval(ae, be, ws, s["e0"], eta, f, tail=tail2)
        long = cf_eval(a[:256], b[:256], ws, s["e0"], eta, f)
        rows.append([spectral_l1(trunc, ref, ws), spectral_l1(term, ref, ws), spectral_l1(learned, ref, ws),
                     spectral_l1(long, ref, ws)])
    return np.array(rows)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", type=pathlib.Path)
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--smooth", type=int, default=1, help="train on block means over this many levels")
    ap.add_argument("--sweep", action="store_true", help="train over smoothing x steps, keep the best on validation")
    ap.add_argument("--eval", type=pathlib.Path, help="npz of held-out sequences (all-q run) to evaluate on")
    ap.add_argument("--model", type=pathlib.Path, default=closure.MODEL_PATH)
    ap.add_argument("--gen-eval", type=pathlib.Path, help="write the 2x12, 2-hole evaluation sequences here")
    args = ap.parse_args(argv)
    ws = np.linspace(0, 4, 2001)
    if args.gen_eval:
        generate_eval(args.gen_eval)
    if args.data and not args.train:
        generate(args.data)
    if args.train:
        seqs = load_sequences(args.data)
        usable = [s for s in seqs if closure.usable(s["a"], s["b"])]
        rng = np.random.default_rng(1)
        perm = rng.permutation(len(usable))
        n_val = max(1, len(usable) // 10)
        val, tr = perm[:n_val], perm[n_val:]
        configs = [(args.smooth, args.steps)] if not args.sweep else [
            (sm, st) for sm in (1, 4, 8, 16, 32) for st in (500, 1500, 4000)]
        best = None
        for smooth, steps in configs:
            X, T = closure.training_arrays(usable, smooth)
            t0 = time.perf_counter()
            model, hist = closure.train(X[tr], T[tr], steps=steps)
            yv = model.net.forward(X[val])
            errs = spectra_errors([usable[i] for i in val], model, ws, 0.05)
            gm = gain(errs[:, 1], errs[:, 2])
            print(f"smooth={smooth} steps={steps}: {time.perf_counter() - t0:.1f}s, final loss {hist[-1][1]:.3e}, "
                  f"validation MSE {float(np.mean((yv - T[val]) ** 2)):.3e} (mean-tail baseline "
                  f"{float(np.mean((T[val] - T[tr].mean(0)) ** 2)):.3e}); validation L1 mean truncated, terminator, "
                  f"learned, 256 steps: {json.dumps(errs.mean(0).tolist())}; terminator/learned geo-mean {gm:.3f}",
                  flush=True)
            if best is None or gm > best[0]:
                best = (gm, smooth, steps, model)
        print(f"selected smooth={best[1]} steps={best[2]} (validation gain {best[0]:.3f})", flush=True)
        best[3].save(args.model)
    if args.eval:
        model = closure.Closure.load(args.model)
        seqs = load_sequences(args.eval)
        errs = spectra_errors(seqs, model, ws, 0.05)
        print("eval L1 per q: truncated, sqrt terminator, learned, 256 steps")
        for s, e in zip(seqs, errs):
            print(s["qx"], s["qy"], " ".join(f"{v:.3e}" for v in e))
        print("eval mean", json.dumps(errs.mean(0).tolist()), "max", json.dumps(errs.max(0).tolist()), flush=True)
        print(f"eval terminator/learned geo-mean {gain(errs[:, 1], errs[:, 2]):.3f}, "
              f"truncated/learned {gain(errs[:, 0], errs[:, 2]):.3f}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
