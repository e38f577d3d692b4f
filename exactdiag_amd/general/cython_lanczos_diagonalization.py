"""Lanczos dynamical response on the GPU (reference: general/cython_lanczos_diagonalization.pyx).

``lanczos_dynamical_response`` keeps the reference's signature: apply the
excitation operator to the state, normalise, run the plain Lanczos
recurrence (no re-orthogonalisation, stop when b < 1e-8) for
``num_lanczos_vecs`` steps and evaluate the continued fraction at every z.
Recurrence and continued fraction run in the HIP library; only the
coefficient arrays (a few hundred doubles) come back to the host.
"""

from __future__ import annotations

import numpy as np

from exactdiag_amd.general.krylov import vec_norm
from exactdiag_amd.sector import cf_eval


def _as_device(v):
    import torch

    if isinstance(v, torch.Tensor):
        return v.to("cuda", torch.complex128).contiguous()
    return torch.as_tensor(np.ascontiguousarray(v, np.complex128), device="cuda")


def _lanczos_get_anbn(v0, num_lanczos_vecs, hamiltonian, precision=0):
    """(a_n, b_n) with b_0 = 0 (cython_lanczos_diagonalization.pyx:50-66)."""
    return hamiltonian.lanczos(_as_device(v0), num_lanczos_vecs, precision=precision)


def _spectral_value_at_z(ans, bns, z):
    """Im(1/det(z)) of the continued fraction (cython_lanczos_diagonalization.pyx:42-48)."""
    return float(cf_eval(ans, bns, [np.real(z)], 0.0, np.imag(z), 1.0)[0])


def continued_fraction(ans, bns, zs, factor):
    zs = np.asarray(zs, np.complex128)
    etas = np.unique(zs.imag)
    if len(etas) == 1:
        return cf_eval(ans, bns, zs.real, 0.0, float(etas[0]), factor)
    return np.array([factor * _spectral_value_at_z(ans, bns, z) for z in zs])


def lanczos_dynamical_response(initial_state, excitation_operator, excited_state_hamiltonian, zs, num_lanczos_vecs,
                               num_threads=None, precision=0):
    """Spectrum of an operator applied to a state (cython_lanczos_diagonalization.pyx:18-40)."""
    x = _as_device(initial_state)
    excited = excitation_operator.dot(x)
    norm = vec_norm(excited) if excited.numel() else 0.0
    spectrum = np.zeros(len(zs), dtype=np.double)
    if norm < 1e-8:
        return spectrum
    ans, bns = _lanczos_get_anbn(excited, num_lanczos_vecs, excited_state_hamiltonian, precision)
    factor = -1.0 / np.pi * norm**2
    return continued_fraction(ans, bns, zs, factor)
