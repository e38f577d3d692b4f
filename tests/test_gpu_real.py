"""Real vectors on stored matrices whose values are all real (momenta kx = 0 and kx = L/2).

The reference computes every sector in complex128 (general/cython_lanczos_diagonalization.pyx:18-77 over
general/cpp_sparse_dot.h:59-138).  For kx = 0 and kx = L/2 its characters are +-1, so H is a real symmetric
matrix in the sector's basis and a real start vector keeps every Lanczos vector real.  These tests pin the real
path to the complex one (same elements, same recurrence) and to the oracle's dense product, and check that the
library refuses sectors or start vectors that are not real.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = [1.0, 1.0, 0.4, 0.4]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _real(torch, n, seed):
    return torch.as_tensor(np.random.default_rng(seed).normal(size=n), device="cuda", dtype=torch.float64)


@pytest.mark.parametrize("L,h,nd", [(6, 2, 5), (5, 1, 4), (4, 0, 4), (7, 2, 6)])
def test_real_sectors_are_those_with_kx_0_or_half(torch, L, h, nd):
    from exactdiag_amd.sector import Sector

    for kx in range(L):
        for ky in (0, 1):
            s = Sector(L, h, nd, kx, ky)
            if s.dim == 0:
                continue
            m = s.store(WEIGHTS)
            expect = kx == 0 or 2 * kx == L
            print("sector", (L, h, nd, kx, ky), "dim", s.dim, "real", m.is_real, flush=True)
            assert m.is_real == expect, (L, h, nd, kx, ky)
            m.close()
    # complex term weights make every sector complex
    s = Sector(L, h, nd, 0, 0)
    if s.dim:
        assert not s.store([1.0 + 0.3j, 0.7, 0.4, 0.25]).is_real


@pytest.mark.parametrize("p", [(6, 2, 5, 0, 0), (6, 2, 5, 3, 1), (8, 2, 7, 4, 0), (6, 1, 5, 3, 0)])
def test_real_product_matches_complex_and_oracle(torch, oracle_mod, p):
    from exactdiag_amd.sector import Sector

    s = Sector(*p)
    m = s.store(WEIGHTS)
    assert m.is_real
    x = _real(torch, s.dim, 3)
    y, d = m.hv(x, with_dot=True)
    assert y.dtype == torch.float64
    yc, dc = m.hv(x.to(torch.complex128), with_dot=True)
    scale = float(yc.abs().max())
    assert float(yc.imag.abs().max()) <= 1e-14 * scale
    assert float((y - yc.real).abs().max()) <= 1e-13 * scale
    assert abs(d - dc) <= 1e-11 * abs(dc) and d.imag == 0.0
    if s.dim <= 20000:
        ref = oracle_mod.OracleSector(*p).hv(WEIGHTS, x.cpu().numpy().astype(np.complex128))
        np.testing.assert_allclose(y.cpu().numpy(), ref.real, rtol=0, atol=1e-12 * np.abs(ref).max())
    # float32 against complex64
    y32 = m.hv(x.to(torch.float32))
    y64 = m.hv(x.to(torch.complex64))
    assert y32.dtype == torch.float32
    assert float((y32.double() - y64.real.double()).abs().max()) <= 1e-6 * scale


def test_real_lanczos_matches_complex(torch):
    from exactdiag_amd._lib import PREC_C64, PREC_C128, PREC_R32, PREC_R64
    from exactdiag_amd.sector import Sector

    s = Sector(8, 2, 7, 4, 1)
    m = s.store(WEIGHTS)
    assert m.is_real
    v0 = _real(torch, s.dim, 7).to(torch.complex128)
    a1, b1 = m.lanczos(v0, 60, precision=PREC_C128)
    a2, b2 = m.lanczos(v0, 60, precision=PREC_R64)
    assert len(a1) == len(a2) == 60
    np.testing.assert_allclose(a2, a1, rtol=0, atol=1e-10)
    np.testing.assert_allclose(b2, b1, rtol=0, atol=1e-10)
    a3, b3 = m.lanczos(v0, 30, precision=PREC_C64)
    a4, b4 = m.lanczos(v0, 30, precision=PREC_R32)
    np.testing.assert_allclose(a4, a1[:30], rtol=0, atol=1e-5)
    np.testing.assert_allclose(b4, b1[:30], rtol=0, atol=1e-5)
    np.testing.assert_allclose(a4, a3, rtol=0, atol=1e-5)
    # second pass: the Ritz combination is the same complex128 vector with a zero imaginary part
    y = np.random.default_rng(2).normal(size=40)
    xc = m.lanczos_ritz(v0, y, precision=PREC_C128)
    xr = m.lanczos_ritz(v0, y, precision=PREC_R64)
    assert xr.dtype == torch.complex128
    assert float(xr.imag.abs().max()) == 0.0
    assert float((xr - xc).abs().max()) < 1e-10


def test_real_precision_refused_where_not_real(torch):
    from exactdiag_amd._lib import PREC_R32, PREC_R64
    from exactdiag_amd.sector import Sector, real_phase

    s = Sector(6, 2, 5, 1, 0)
    m = s.store(WEIGHTS)
    assert not m.is_real
    v0 = _real(torch, s.dim, 1).to(torch.complex128)
    with pytest.raises(ValueError, match="real stored matrix"):
        m.lanczos(v0, 10, precision=PREC_R64)
    with pytest.raises(ValueError, match="real stored matrix"):
        m.hv(v0.real.contiguous())
    # a real sector with a start vector that carries a global phase: refused, and accepted once rotated
    s = Sector(6, 2, 5, 3, 0)
    m = s.store(WEIGHTS)
    v = _real(torch, s.dim, 4).to(torch.complex128) * complex(np.exp(0.7j))
    with pytest.raises(ValueError, match="not real"):
        m.lanczos(v, 10, precision=PREC_R32)
    r = real_phase(v)
    assert r is not None
    a1, b1 = m.lanczos(v, 40)
    a2, b2 = m.lanczos(r, 40, precision=PREC_R64)
    np.testing.assert_allclose(a2, a1, rtol=0, atol=1e-10)
    np.testing.assert_allclose(b2, b1, rtol=0, atol=1e-10)
    # a genuinely complex vector has no real phase
    z = torch.as_tensor(np.random.default_rng(5).normal(size=(s.dim, 2)) @ [1, 1j], device="cuda")
    assert real_phase(z) is None


def test_real_ground_state_and_fast_allq(torch):
    """Two-pass ground state in float64 vectors gives the complex128 energy; the fast all-q form with real
    sectors in float32 gives the spectra of the complex64 run."""
    from exactdiag_amd import allq
    from exactdiag_amd.forms import two_pass_ground_state
    from exactdiag_amd.sector import Sector

    s = Sector(10, 2, 9, 0, 0)
    m = s.store(WEIGHTS)
    e1, x1, _, r1 = two_pass_ground_state(m, tol=1e-10)
    e2, x2, _, r2 = two_pass_ground_state(m, tol=1e-10, real=True)
    print("E0 complex", e1, "real", e2, "residuals", r1, r2, flush=True)
    assert abs(e2 - e1) <= 1e-10 * abs(e1)
    assert r2 <= 1e-10 * max(1.0, abs(e2))
    assert float(x2.imag.abs().max()) == 0.0
    hx = m.hv(x2)
    assert float(torch.linalg.vector_norm(hx - e2 * x2)) <= 1e-10 * max(1.0, abs(e2))
    # the fast form with and without real vectors, FP64 and reduced precision, against 1000-step FP64 spectra
    # (complex128): every run meets the 1e-3 spectral L1 target.  The runs are not compared with each other:
    # once the recurrence loses orthogonality (after ~60 steps here) its coefficients depend on rounding, and
    # two unconverged 170-step spectra differ by about 1e-4 even in FP64.
    ws = np.linspace(0, 4, 401)
    _, ref = allq.allq_reference(10, 2, 9, WEIGHTS, ws, 0.05, 1000)
    runs = {}
    for prec in (0, 1):
        for real in (False, True):
            runs[prec, real] = allq.allq_fast(10, 2, 9, WEIGHTS, ws, 0.05, 170, tol=1e-10, precision=prec,
                                              real=real)[1]
    for q in ref:
        if not np.abs(ref[q]).sum():
            continue
        errs = {k: allq.spectral_l1(runs[k][q], ref[q], ws) for k in runs}
        print("q", q, "L1 vs 1000 steps:", errs, flush=True)
        assert max(errs.values()) <= 1e-3, q


def test_api_lanczos_takes_real_vectors_on_real_sectors(torch):
    """The reference-API Hamiltonian (Hamiltonian_Operator over the stored H) runs its Lanczos in real vectors
    on a real sector: the coefficients are those of the explicit float64 run, and equal the complex128 ones to
    rounding."""
    from exactdiag_amd._lib import PREC_R64
    from exactdiag_amd.general.sparse_matrices import Hamiltonian_Operator
    from exactdiag_amd.sector import Sector

    s = Sector(10, 2, 9, 5, 1)
    ham = Hamiltonian_Operator(s, {"tl": 1.0, "tr": 1.0, "jl": 0.4, "jr": 0.4})
    assert ham.stored is not None and ham.stored.is_real
    v = _real(torch, s.dim, 8).to(torch.complex128) * complex(np.exp(-1.3j))
    a1, b1 = ham.lanczos(v, 60)
    a2, b2 = ham.stored.lanczos(v * complex(np.exp(1.3j)), 60, precision=PREC_R64)
    np.testing.assert_allclose(a1, a2, rtol=0, atol=1e-12)
    np.testing.assert_allclose(b1, b2, rtol=0, atol=1e-12)
    ham.use_real_vectors = False
    a3, b3 = ham.lanczos(v, 60)
    np.testing.assert_allclose(a1, a3, rtol=0, atol=1e-10)
    np.testing.assert_allclose(b1, b3, rtol=0, atol=1e-10)


def test_thick_restart_in_real_vectors(torch, oracle_mod):
    """The eigenpair solver in float64 vectors on a real sector: the energies of the complex128 run and of the
    oracle's dense spectrum, eigenvectors real with small residuals, also for an odd dimension (padded pair
    view)."""
    from exactdiag_amd.general import krylov
    from exactdiag_amd.sector import Sector

    for p in [(10, 2, 9, 0, 0), (6, 2, 5, 3, 1), (7, 1, 6, 0, 1)]:
        s = Sector(*p)
        m = s.store(WEIGHTS)
        assert m.is_real
        print("sector", p, "dim", s.dim, flush=True)
        ec, Xc, _ = krylov.lowest_eigenpairs(m.hv, s.dim, 3, tol=1e-10)
        er, Xr, _ = krylov.lowest_eigenpairs(m.hv, s.dim, 3, tol=1e-10, real=True)
        np.testing.assert_allclose(er, ec, rtol=0, atol=1e-9 * max(1.0, abs(ec[0])))
        assert Xr.dtype == torch.complex128 and Xr.shape == (3, s.dim)
        assert float(Xr.imag.abs().max()) == 0.0
        for i in range(3):
            r = m.hv(Xr[i]) - er[i] * Xr[i]
            assert float(torch.linalg.vector_norm(r)) <= 1e-8 * max(1.0, abs(er[i]))
        if s.dim <= 3000:
            ev = np.linalg.eigvalsh(oracle_mod.OracleSector(*p).toarray(WEIGHTS))
            np.testing.assert_allclose(er, ev[:3], rtol=0, atol=1e-9 * max(1.0, abs(ev[0])))


def test_degenerate_ground_states_orthonormal_in_real_vectors(torch, oracle_mod):
    """A real sector whose ground state is two-fold degenerate (Heisenberg exchange with static holes, 2x6 with 3
    holes): the float64 Krylov basis returns orthonormal eigenvectors, the same eigenvalues as the complex128
    basis and the dense spectrum, and -- where both find the pair -- the same degenerate subspace, so the
    degenerate averages that calculate_spectrum forms agree."""
    from exactdiag_amd.general import krylov
    from exactdiag_amd.sector import Sector

    p, w = (6, 3, 4, 0, 0), [0.0, 0.0, 1.0, 1.0]
    s = Sector(*p)
    m = s.store(w)
    assert m.is_real
    ev = np.linalg.eigvalsh(np.conj(oracle_mod.OracleSector(*p).toarray(w)))
    assert ev[1] - ev[0] < 1e-9 < ev[2] - ev[1]
    out = {}
    for real in (True, False):
        vals, X, _ = krylov.lowest_eigenpairs(m.hv, s.dim, 3, tol=1e-11, real=real)
        G = (X.conj() @ X.T).cpu().numpy()
        assert np.max(np.abs(G - np.eye(len(vals)))) < 1e-10, real
        for e in vals:
            assert np.min(np.abs(ev - e)) < 1e-9
        assert abs(vals[0] - ev[0]) < 1e-10
        deg = X[np.abs(vals - ev[0]) < 1e-6]
        out[real] = (vals, deg)
        print("real" if real else "complex", vals, "degenerate copies found:", deg.shape[0], flush=True)
    if out[True][1].shape[0] == out[False][1].shape[0] == 2:
        P = [(d.T @ d.conj()).cpu().numpy() for d in (out[True][1], out[False][1])]
        assert np.max(np.abs(P[0] - P[1])) < 1e-8
