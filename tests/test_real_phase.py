"""Host logic of the real-vector path: the global phase rotation that makes a start vector real.

Lanczos coefficients depend only on the direction of the start vector up to a global phase
(general/cython_lanczos_diagonalization.pyx:18-40 normalises it and uses |<v|v>| only), so a vector that is
real up to a phase can run in real vectors once that phase is rotated away.  CPU tensors; no GPU needed.
"""

import numpy as np
import pytest


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


def test_rotates_real_up_to_phase(torch):
    from exactdiag_amd.sector import real_phase

    rng = np.random.default_rng(0)
    r = rng.normal(size=500)
    for phase in (0.0, 0.3, -2.1, np.pi):
        v = torch.as_tensor(r * np.exp(1j * phase))
        out = real_phase(v)
        assert out is not None
        assert float(out.imag.abs().max()) <= 1e-14
        i = int(np.argmax(np.abs(r)))
        assert float(out[i].real) > 0
        # the same vector up to a sign: |<r, out>| = ||r||^2
        assert abs(abs(float((out.real.numpy() * r).sum())) - float((r * r).sum())) <= 1e-10 * float((r * r).sum())


def test_refuses_complex_and_zero_vectors(torch):
    from exactdiag_amd.sector import real_phase

    rng = np.random.default_rng(1)
    z = torch.as_tensor(rng.normal(size=300) + 1j * rng.normal(size=300))
    assert real_phase(z) is None
    # a small imaginary admixture above the tolerance is refused, one below it accepted
    r = rng.normal(size=300)
    w = r + 1j * 1e-9 * rng.normal(size=300)
    assert real_phase(torch.as_tensor(w)) is None
    assert real_phase(torch.as_tensor(r + 1j * 1e-14 * rng.normal(size=300))) is not None
    assert real_phase(torch.zeros(10, dtype=torch.complex128)) is None
    assert real_phase(torch.zeros(0, dtype=torch.complex128)) is None
