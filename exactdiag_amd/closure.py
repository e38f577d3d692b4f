"""Learned continued-fraction closure.

No counterpart in the reference.  A small MLP predicts the tail of the
Lanczos coefficient sequence from its first 64 (a_n, b_n) pairs, so a
spectrum needs 64 Lanczos steps instead of hundreds.  The network is trained
on long (1000-step, FP64) Lanczos sequences that this package generates, and
it runs forward and backward in the library's BF16 MFMA kernels
(``ed_mlp_*`` in include/exactdiag_hip.h).

Normalisation by bandwidth: with theta_min, theta_max the extreme Ritz
values of the 64-step tridiagonal matrix, c = (theta_max + theta_min) / 2 and
w = (theta_max - theta_min) / 2, the network sees a' = (a - c) / w and
b' = b / w (a band [c - w, c + w] has a_n -> 0, b_n -> 1/2 in these units).

Input: a'_0..a'_63, b'_1..b'_64 (128 values).  Output: a'_64..a'_255,
b'_65..b'_256 (384 values = 192 pairs).  Beyond the predicted levels the
fraction is closed with the square-root terminator of the last predicted
pair's running mean.
"""

from __future__ import annotations

import ctypes as ct
import dataclasses
import pathlib

import numpy as np
from scipy.linalg import eigvalsh_tridiagonal

from exactdiag_amd._lib import check, lib

N_IN = 64
N_OUT = 192
DIMS = [2 * N_IN, 256, 256, 256, 2 * N_OUT]
MODEL_PATH = pathlib.Path(__file__).resolve().parent / "data" / "closure_mlp.npz"


def _torch():
    import torch

    return torch


def _stream():
    return ct.c_void_p(_torch().cuda.current_stream().cuda_stream)


class GpuMLP:
    """The closure network on the device (FP32 master weights, BF16 MFMA products)."""

    def __init__(self, dims=DIMS, seed=0):
        self.dims = list(dims)
        arr = (ct.c_int * len(self.dims))(*self.dims)
        h = ct.c_void_p()
        check(lib().ed_mlp_create(arr, len(self.dims), int(seed), ct.byref(h)))
        self._h = h
        self.n_params = int(lib().ed_mlp_num_params(h))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().ed_mlp_destroy(h)
            self._h = None

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        check(lib().ed_mlp_get_params(self._h, p.ctypes.data_as(ct.POINTER(ct.c_float)), self.n_params))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(lib().ed_mlp_set_params(self._h, p.ctypes.data_as(ct.POINTER(ct.c_float)), self.n_params))

    def _dev(self, a):
        torch = _torch()
        return torch.as_tensor(np.ascontiguousarray(a, np.float32), device="cuda")

    def forward(self, x):
        torch = _torch()
        xd = self._dev(x)
        y = torch.empty((xd.shape[0], self.dims[-1]), dtype=torch.float32, device="cuda")
        check(lib().ed_mlp_forward(self._h, ct.c_void_p(xd.data_ptr()), xd.shape[0], ct.c_void_p(y.data_ptr()),
                                   _stream()))
        return y.cpu().numpy()

    def gradients(self, x, t):
        xd, td = self._dev(x), self._dev(t)
        g = np.empty(self.n_params, np.float32)
        loss = ct.c_double(0.0)
        check(lib().ed_mlp_gradients(self._h, ct.c_void_p(xd.data_ptr()), ct.c_void_p(td.data_ptr()), xd.shape[0],
                                     g.ctypes.data_as(ct.POINTER(ct.c_float)), ct.byref(loss), _stream()))
        return loss.value, g

    def train_step(self, x, t, lr):
        xd, td = (x, t) if not isinstance(x, np.ndarray) else (self._dev(x), self._dev(t))
        loss = ct.c_double(0.0)
        check(lib().ed_mlp_train_step(self._h, ct.c_void_p(xd.data_ptr()), ct.c_void_p(td.data_ptr()), xd.shape[0],
                                      ct.c_float(lr), ct.byref(loss), _stream()))
        return loss.value


# -- features ----------------------------------------------------------------

def bandwidth(a, b, n=N_IN):
    """(centre, half-width) from the extreme Ritz values of the n-step tridiagonal."""
    n = min(n, len(a))
    th = eigvalsh_tridiagonal(np.asarray(a[:n], np.float64), np.asarray(b[1:n], np.float64))
    c = 0.5 * (th[-1] + th[0])
    w = max(0.5 * (th[-1] - th[0]), 1e-12)
    return c, w


def features(a, b):
    c, w = bandwidth(a, b)
    x = np.concatenate([(np.asarray(a[:N_IN]) - c) / w, np.asarray(b[1:N_IN + 1]) / w])
    return x.astype(np.float32), c, w


def _block_mean(v, k):
    """Mean over consecutive blocks of k levels, repeated back to the original length."""
    if k <= 1:
        return v
    return np.asarray(v).reshape(-1, k).mean(1).repeat(k)


def targets(a, b, c, w, smooth=1):
    """The tail a'_64..a'_255, b'_65..b'_256 in bandwidth units; with smooth > 1, block means over smooth levels
    (the level-by-level fluctuations carry the discrete spectrum and cannot be inferred from 64 levels)."""
    ta = _block_mean((np.asarray(a[N_IN:N_IN + N_OUT]) - c) / w, smooth)
    tb = _block_mean(np.asarray(b[N_IN + 1:N_IN + N_OUT + 1]) / w, smooth)
    return np.concatenate([ta, tb]).astype(np.float32)


def usable(a, b):
    """Sequences long enough to provide input and target (the Krylov space was not exhausted)."""
    return len(a) >= N_IN + N_OUT + 1 and len(b) >= N_IN + N_OUT + 1


# -- closing a spectrum ---------------------------------------------------------

@dataclasses.dataclass
class Closure:
    net: GpuMLP

    @classmethod
    def load(cls, path=MODEL_PATH):
        d = np.load(path, allow_pickle=False)
        net = GpuMLP([int(v) for v in d["dims"]])
        net.set_params(d["params"])
        return cls(net)

    def save(self, path=MODEL_PATH):
        pathlib.Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, dims=np.asarray(self.net.dims), params=self.net.get_params())

    def extend(self, a, b):
        """64 measured pairs -> 64 + 192 pairs and the terminator (a_inf, b_inf)."""
        if len(a) < N_IN + 1:
            return np.asarray(a), np.asarray(b), None  # Krylov space exhausted: the fraction is exact
        x, c, w = features(a, b)
        y = self.net.forward(x[None, :])[0]
        a_ext = np.concatenate([np.asarray(a[:N_IN]), c + w * y[:N_OUT]])
        b_ext = np.concatenate([np.asarray(b[:N_IN + 1]), w * np.maximum(y[N_OUT:], 0.0)])
        a_inf = float(np.mean(a_ext[-32:]))
        b_inf = float(np.mean(b_ext[-32:]))
        return a_ext, b_ext[: len(a_ext)], (a_inf, b_inf)


def sqrt_terminator(a, b, n=N_IN):
    """The method without the network: n measured pairs closed by the square-root terminator
    at the bandwidth of the n-step tridiagonal (a_inf = c, b_inf = w / 2)."""
    if len(a) < n:
        return np.asarray(a), np.asarray(b), None  # Krylov space exhausted: the fraction is exact
    c, w = bandwidth(a, b, n)
    return np.asarray(a[:n]), np.asarray(b[:n]), (c, 0.5 * w)


def evaluate_allq(model, L=12, h=2, nd=11, weights=(1.0, 1.0, 0.4, 0.4), ws=None, eta=0.05, nsteps_ref=1000):
    """S(q, w) of every distinct q of a ladder closed at 64 levels, with the network and with the square-root
    terminator, each against the 1000-step FP64 spectrum (stored H, ground state to 1e-12).  Returns per-q
    spectral L1 errors and the seconds the network's forward passes took."""
    import time

    from exactdiag_amd import _lib
    from exactdiag_amd.allq import _gs, distinct_qs, final_momentum, spectral_l1
    from exactdiag_amd.general.krylov import vec_norm
    from exactdiag_amd.sector import Sector, cf_eval

    ws = np.linspace(0, 4, 2001) if ws is None else np.asarray(ws)
    w = list(weights)
    s0 = Sector(L, h, nd, 0, 0)
    hm0 = s0.store(w)
    e0, psi = _gs(hm0.hv, s0.dim, 1e-12)
    hm0.close()
    per_q, t_net = {}, 0.0
    for qx, qy in distinct_qs(L):
        s1 = Sector(L, h, nd, *final_momentum(L, 0, 0, qx, qy))
        phi = s0.op_apply(s1, _lib.OP_SZQ, qx, qy, psi)
        n2 = vec_norm(phi) ** 2
        if n2 < 1e-14:
            s1.close()
            continue
        hm = s1.store(w)
        a, b = hm.lanczos(phi, nsteps_ref)
        hm.close()
        s1.close()
        f = -n2 / np.pi
        ref = cf_eval(a, b, ws, e0, eta, f)
        at, bt, tail = sqrt_terminator(a, b)
        t1 = time.perf_counter()
        ae, be, tail2 = model.extend(a, b)  # reads the first 64 levels only
        t_net += time.perf_counter() - t1
        per_q[(qx, qy)] = (spectral_l1(cf_eval(at, bt, ws, e0, eta, f, tail=tail), ref, ws),
                           spectral_l1(cf_eval(ae, be, ws, e0, eta, f, tail=tail2), ref, ws))
    s0.close()
    return per_q, t_net


# -- training data ----------------------------------------------------------------

def lanczos_sequences(L, h, nd, weights, nsteps=1000, ops=("Szq", "Nq", "A+", "A-"), limited=True, seed=0):
    """Long FP64 Lanczos runs of one system for every q and operator: list of dicts."""
    import torch

    from exactdiag_amd import _lib
    from exactdiag_amd.allq import distinct_qs, final_momentum, ground_state
    from exactdiag_amd.general.krylov import vec_norm
    from exactdiag_amd.sector import Sector, weights_array

    out = []
    s0, e0, psi = ground_state(L, h, nd, 0, 0, weights, tol=1e-10, seed=seed)
    N = 2 * L
    ne = N - h
    tsp = ne - 2 * nd
    for op in ops:
        for qx, qy in distinct_qs(L, limited):
            if op in ("Szq", "Nq"):
                hf, ndf, opid = h, nd, _lib.OP_SZQ if op == "Szq" else _lib.OP_NQ
                qxx, qyy = qx, qy
            elif op == "A+":
                if h == 0:
                    continue
                hf = h - 1
                ndf, opid = (nd + 1, _lib.OP_APLUS_DOWN) if tsp > 0 else (nd, _lib.OP_APLUS_UP)
                qxx, qyy = qx, qy
            else:
                if h == N:
                    continue
                hf = h + 1
                ndf, opid = (nd - 1, _lib.OP_AMINUS_DOWN) if tsp <= 0 else (nd, _lib.OP_AMINUS_UP)
                qxx, qyy = -qx, (-qy) % 2
                if ndf < 0 or ndf > N - hf:
                    continue
            kx, ky = final_momentum(L, 0, 0, qxx, qyy)
            sec = Sector(L, hf, ndf, kx, ky)
            if sec.dim == 0:
                continue
            phi = s0.op_apply(sec, opid, qxx, qyy, psi)
            n2 = vec_norm(phi) ** 2
            if n2 < 1e-14:
                sec.close()
                continue
            a, b = sec.lanczos(weights_array(np.asarray(weights, np.complex128)), phi, nsteps)
            out.append(dict(L=L, h=h, nd=nd, J=float(weights[2]), op=op, qx=qx, qy=qy, e0=e0, norm2=n2, a=a, b=b,
                            dim=sec.dim))
            sec.close()
            torch.cuda.synchronize()
    s0.close()
    return out


def training_arrays(seqs, smooth=1):
    X, T = [], []
    for s in seqs:
        if not usable(s["a"], s["b"]):
            continue
        x, c, w = features(s["a"], s["b"])
        X.append(x)
        T.append(targets(s["a"], s["b"], c, w, smooth))
    return np.stack(X), np.stack(T)


def train(X, T, steps=3000, batch=256, lr=1e-3, seed=0, log_every=500):
    """Adam on the GPU; returns (Closure, loss history)."""
    torch = _torch()
    net = GpuMLP(DIMS, seed=seed)
    rng = np.random.default_rng(seed)
    Xd = torch.as_tensor(X, device="cuda")
    Td = torch.as_tensor(T, device="cuda")
    hist = []
    for it in range(steps):
        idx = torch.as_tensor(rng.integers(0, len(X), size=min(batch, len(X))), device="cuda")
        xb = Xd.index_select(0, idx).contiguous()
        tb = Td.index_select(0, idx).contiguous()
        lr_t = lr * 0.5 * (1 + np.cos(np.pi * it / steps))
        loss = net.train_step(xb, tb, lr_t)
        if it % log_every == 0 or it == steps - 1:
            hist.append((it, loss))
    return Closure(net), hist
