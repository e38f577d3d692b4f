"""Lanczos with the basis kept in HBM (ed_hmat_lanczos_run_basis) and the ground state that uses it."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = [1.0, 1.0, 0.4, 0.4]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.mark.parametrize("kx,precision", [(1, 0), (0, 2)])
def test_kept_basis_is_the_lanczos_basis(torch, kx, precision):
    """Rows are orthonormal and satisfy H v_n = b_n v_{n-1} + a_n v_n + b_{n+1} v_{n+1}; the coefficients are the
    plain run's (to rounding: the update's norm sums run in another order)."""
    from exactdiag_amd.sector import Sector

    s = Sector(6, 2, 5, kx, 0)
    hm = s.store(W)
    rng = np.random.default_rng(7)
    v0 = rng.normal(size=s.dim) + (1j * rng.normal(size=s.dim) if precision == 0 else 0.0)
    v0 = torch.as_tensor(np.asarray(v0, np.complex128), device="cuda")
    n = 30
    print("2x6 k =", kx, "dim", s.dim, "precision", precision, flush=True)
    a, b, V = hm.lanczos_basis(v0, n, precision=precision, ld=s.dim + 3)
    af, bf = hm.lanczos(v0, n, precision=precision)
    assert len(a) == n and np.allclose(a, af, rtol=0, atol=1e-12) and np.allclose(b, bf, rtol=0, atol=1e-12)
    Vn = V[:, : s.dim].cpu().numpy().astype(np.complex128)
    assert np.all(V[:, s.dim:].cpu().numpy() == 0)  # columns past dim untouched
    G = Vn.conj() @ Vn.T
    assert np.max(np.abs(G - np.eye(n))) < 1e-10
    v0n = v0.cpu().numpy() / np.linalg.norm(v0.cpu().numpy())
    assert np.max(np.abs(Vn[0] - v0n)) < 1e-13
    for k in range(1, n - 1):
        x = V[k, : s.dim].to(torch.complex128) if precision == 0 else V[k, : s.dim].contiguous()
        hx = hm.hv(x).cpu().numpy().astype(np.complex128)
        expect = b[k] * Vn[k - 1] + a[k] * Vn[k] + b[k + 1] * Vn[k + 1]
        assert np.max(np.abs(hx - expect)) < 1e-9
    hm.close()
    s.close()


@pytest.mark.parametrize("real", [False, True])
def test_ground_state_from_the_kept_basis(torch, real):
    """The ground state combined from the kept basis equals the two-pass one."""
    from exactdiag_amd import forms
    from exactdiag_amd.sector import Sector

    s = Sector(8, 2, 7, 0, 0)
    hm = s.store(W)
    e1, x1, r1, res1 = forms.two_pass_ground_state(hm, nsteps=100, tol=1e-10, real=real)
    keep = forms.BASIS_BYTES
    try:
        forms.BASIS_BYTES = 0
        e2, x2, r2, res2 = forms.two_pass_ground_state(hm, nsteps=100, tol=1e-10, real=real)
    finally:
        forms.BASIS_BYTES = keep
    print("real", real, "E", e1, e2, "restarts", r1, r2, "residuals", res1, res2, flush=True)
    assert abs(e1 - e2) < 1e-11 * abs(e1)
    assert res1 <= 1e-10 * abs(e1)
    ov = abs(np.vdot(x1.cpu().numpy(), x2.cpu().numpy()))
    assert abs(ov - 1.0) < 1e-9
    hm.close()
    s.close()
