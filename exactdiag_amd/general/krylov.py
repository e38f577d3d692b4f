"""Thick-restart Lanczos eigensolver on the GPU (replaces ARPACK in
general/lanczos_diagonalization.py:81-121).

The reference calls scipy's ARPACK (implicitly restarted Arnoldi, 'SR',
ncv = nev + 4, tol 2e-7).  This module keeps the basis on the device and
runs thick-restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22, 2000)
with full re-orthogonalisation (classical Gram-Schmidt applied twice) in the
library's vector kernels.  Only the small projected matrix (at most a few
dozen rows) is diagonalised on the host with numpy, as the reference's ARPACK
does with its own small dense problem.
"""

from __future__ import annotations

import ctypes as ct
import logging

import numpy as np

from exactdiag_amd._lib import check, lib

_logger = logging.getLogger(__name__)


def _torch():
    import torch

    return torch


def _stream():
    return ct.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _dptr(arr):
    return arr.ctypes.data_as(ct.POINTER(ct.c_double))


def vec_dots(V, nv, w):
    """<V_k, w> for the first nv rows of V (device, complex128)."""
    D = w.numel()
    out = np.zeros(2 * max(nv, 1), np.float64)
    if nv:
        check(lib().ed_vec_dots(ct.c_void_p(V.data_ptr()), nv, V.stride(0), ct.c_void_p(w.data_ptr()), D,
                                _dptr(out), _stream()))
    return out.view(np.complex128)[:nv].copy()


def vec_axpys(V, nv, c, w):
    """w -= sum_k c_k V_k."""
    c = np.ascontiguousarray(np.asarray(c, np.complex128)).view(np.float64)
    if nv:
        check(lib().ed_vec_axpys(ct.c_void_p(V.data_ptr()), nv, V.stride(0), _dptr(c), ct.c_void_p(w.data_ptr()),
                                 w.numel(), _stream()))


def vec_combine(V, nv, c, out):
    """out = sum_k c_k V_k."""
    c = np.ascontiguousarray(np.asarray(c, np.complex128)).view(np.float64)
    check(lib().ed_vec_combine(ct.c_void_p(V.data_ptr()), nv, V.stride(0), _dptr(c), ct.c_void_p(out.data_ptr()),
                               out.numel(), _stream()))


def vec_norm(w):
    out = np.zeros(1, np.float64)
    check(lib().ed_vec_norm2(ct.c_void_p(w.data_ptr()), w.numel(), _dptr(out), _stream()))
    return float(np.sqrt(out[0]))


def vec_scale(w, c):
    check(lib().ed_vec_scale(ct.c_void_p(w.data_ptr()), w.numel(), float(np.real(c)), float(np.imag(c)),
                             _stream()))


def start_vector(D, seed=0):
    """Seeded complex normal start vector (host numpy, like the reference's random ARPACK start)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=D) + 1j * rng.normal(size=D)
    return v / np.linalg.norm(v)


def device_start_vector(D, seed=0):
    """Seeded complex normal start vector drawn on the device (no host round trip for large sectors)."""
    torch = _torch()
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    v = torch.randn(D, dtype=torch.complex128, device="cuda", generator=g)
    vec_scale(v, 1.0 / vec_norm(v))
    return v


def plain_lanczos(apply, v0, nsteps):
    """(a_n, b_n) of the plain recurrence for any device operator (cython_lanczos_diagonalization.pyx:50-77).

    v_0 = v0 / |v0|; a_n = Re<v_n, H v_n>; w = H v_n - a_n v_n - b_n v_{n-1}; b_{n+1} = |w|; stop after
    the step whose b_{n+1} < 1e-8.  Returns a[:n], b[:n] with b[0] = 0, as the reference's
    _lanczos_get_anbn.  One host round trip per coefficient: for operators without a fused driver.
    """
    torch = _torch()
    D = v0.numel()
    V = torch.zeros((2, D), dtype=torch.complex128, device=v0.device)
    V[0].copy_(v0)
    vec_scale(V[0], 1.0 / vec_norm(V[0]))
    a = np.zeros(nsteps)
    b = np.zeros(nsteps + 1)
    cur, n = 0, 0
    for i in range(nsteps):
        w = apply(V[cur])
        a[i] = float(np.real(vec_dots(V[cur : cur + 1], 1, w)[0]))
        coef = [a[i], b[i]] if cur == 0 else [b[i], a[i]]  # rows in memory order
        vec_axpys(V, 2, coef, w)
        b[i + 1] = vec_norm(w)
        n += 1
        if b[i + 1] < 1e-8:
            break
        V[1 - cur].copy_(w)
        vec_scale(V[1 - cur], 1.0 / b[i + 1])
        cur = 1 - cur
    return a[:n], b[:n]


def lowest_eigenpairs(apply, D, nev, tol=1e-10, ncv=None, v0=None, max_restarts=500, real=False):
    """Lowest ``nev`` eigenpairs of the Hermitian operator ``apply`` (device tensor in, out).

    Returns (eigenvalues ascending, device tensor (nev, D) of eigenvectors, restarts).
    Convergence: residual norm of every wanted Ritz pair below tol * max(1, |theta|).
    real=True: ``apply`` is a real symmetric operator on float64 vectors; the basis is kept in float64 (half the
    bytes), v0 rotated by its global phase starts it (a seeded real vector when v0 stays complex), and the
    eigenvectors come back as complex128 with zero imaginary part.
    The vector helpers then see each real vector as complex pairs (length padded to even): the real part of a
    pair-view dot is the real dot, and real coefficients give the real combinations.
    """
    torch = _torch()
    m = ncv or max(2 * nev + 24, 40)
    m = min(m, D)
    if isinstance(v0, torch.Tensor):
        v = v0.to("cuda", torch.complex128)
    elif v0 is None:
        v = device_start_vector(D)
    else:
        v = torch.as_tensor(np.asarray(v0, np.complex128), device="cuda")
    if real:
        # a complex guess is rotated by its global phase first (its real part alone could vanish); a guess that
        # stays complex gives way to a seeded real start vector
        from exactdiag_amd.sector import real_phase

        rotated = real_phase(v) if v0 is not None else v
        if rotated is None:
            rotated = torch.as_tensor(np.random.default_rng(0).normal(size=D), device="cuda").to(torch.complex128)
        v = rotated
        if float(torch.linalg.vector_norm(v.real)) == 0.0:
            raise ValueError("the start vector of the real Krylov basis is zero")
        Dp = D + (D & 1)
        Vr = torch.zeros((m + 1, Dp), dtype=torch.float64, device="cuda")
        V = torch.view_as_complex(Vr.view(m + 1, Dp // 2, 2))
        Vr[0, :D].copy_(v.real)
        wr = torch.zeros(Dp, dtype=torch.float64, device="cuda")
        wbuf = torch.view_as_complex(wr.view(Dp // 2, 2))

        def step(j):
            wr[:D].copy_(apply(Vr[j, :D]))
            return wbuf

        def dots(nv, w):
            return vec_dots(V, nv, w).real.astype(np.complex128)
    else:
        V = torch.empty((m + 1, D), dtype=torch.complex128, device="cuda")
        V[0].copy_(v)

        def step(j):
            return apply(V[j])

        def dots(nv, w):
            return vec_dots(V, nv, w)
    vec_scale(V[0], 1.0 / vec_norm(V[0]))
    T = np.zeros((m, m))
    k = 0
    beta = 0.0
    for restart in range(max_restarts):
        m_eff = m
        for j in range(k, m):
            w = step(j)
            c = dots(j + 1, w)
            vec_axpys(V, j + 1, c, w)
            c2 = dots(j + 1, w)
            vec_axpys(V, j + 1, c2, w)
            T[j, j] = (c[j] + c2[j]).real
            beta = vec_norm(w)
            scale = max(1.0, abs(T[j, j]))
            if beta < 1e-13 * scale:
                m_eff = j + 1
                beta = 0.0
                break
            V[j + 1].copy_(w)
            vec_scale(V[j + 1], 1.0 / beta)
            if j + 1 < m:
                T[j, j + 1] = T[j + 1, j] = beta
        theta, Y = np.linalg.eigh(T[:m_eff, :m_eff])
        res = np.abs(beta * Y[m_eff - 1, :])
        want = min(nev, m_eff)
        conv = res[:want] < tol * np.maximum(1.0, np.abs(theta[:want]))
        if np.all(conv) or m_eff == D or beta == 0.0:
            X = torch.empty((want, V.shape[1]), dtype=torch.complex128, device="cuda")
            for i in range(want):
                vec_combine(V, m_eff, Y[:, i], X[i])
                vec_scale(X[i], 1.0 / vec_norm(X[i]))
            if real:
                X = torch.view_as_real(X).reshape(want, -1)[:, :D].to(torch.complex128)
            _logger.info(f"thick-restart Lanczos converged after {restart} restarts, residuals {res[:want]}")
            return theta[:want], X, restart
        k = min(m_eff - 1, nev + max(nev, (m_eff - nev) // 2))
        tmp = torch.empty((k, V.shape[1]), dtype=torch.complex128, device="cuda")
        for i in range(k):
            vec_combine(V, m_eff, Y[:, i], tmp[i])
        V[:k].copy_(tmp)
        V[k].copy_(V[m_eff])
        del tmp
        T = np.zeros((m, m))
        T[np.arange(k), np.arange(k)] = theta[:k]
        s = beta * Y[m_eff - 1, :k]
        T[:k, k] = s
        T[k, :k] = s
    raise RuntimeError("thick-restart Lanczos did not converge")
