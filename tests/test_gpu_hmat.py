"""Stored Hamiltonian (sliced ELL on the device) against the matrix-free product and the CPU oracle.

The stored form is the GPU counterpart of the reference's CSR matrices
(general/cython_sparse_matrices.pyx:58-241) and their dot
(general/cpp_sparse_dot.h:59-138): its product must agree with the
matrix-free H·v (same elements, rounding of the per-row factor apart) and
with the oracle's dense matrix, for every term weight and momentum.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SECTORS = [
    (2, 0, 2, 0, 0), (3, 1, 2, 0, 0), (3, 2, 2, 0, 0), (4, 2, 3, 1, 1), (4, 3, 2, 2, 0), (4, 0, 4, 2, 1),
    (5, 2, 4, 2, 0), (6, 2, 5, 0, 0), (6, 1, 5, 3, 1), (6, 0, 6, 3, 0), (6, 4, 4, 0, 1), (7, 2, 6, 3, 0),
    (4, 2, 3, -1, 0), (8, 2, 7, 1, 0),
]
WEIGHTS = [
    [1.0, 1.0, 0.4, 0.4],
    [0.45, 0.36, 0.15, 0.12],
    [0.0, 0.0, 1.0, 1.0],
    [1.0 + 0.3j, 0.7 - 0.2j, 0.4j, 0.25],
]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _rand(torch, n, seed, dtype=None):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=n) + 1j * rng.normal(size=n)
    return torch.as_tensor(v, device="cuda", dtype=dtype or torch.complex128)


@pytest.fixture(params=[4, 6], ids=["4-byte", "6-byte"])
def element_bytes(request, monkeypatch):
    """Both stored element forms: packed 32-bit words and int32 column + uint16 code."""
    if request.param == 6:
        monkeypatch.setenv("ED_HMAT_ELEMENT_BYTES", "6")
    else:
        monkeypatch.delenv("ED_HMAT_ELEMENT_BYTES", raising=False)
    return request.param


def test_stored_matches_matrix_free(torch, element_bytes):
    from exactdiag_amd.sector import Sector

    for p in SECTORS:
        s = Sector(*p)
        for wi, w in enumerate(WEIGHTS):
            print("sector", p, "weights", wi, "dim", s.dim, flush=True)
            m = s.store(w)
            assert m.nnz <= m.slots and m.bytes > 0
            assert m.element_bytes == element_bytes
            if s.dim == 0:
                continue
            x = _rand(torch, s.dim, wi)
            y_mf, d_mf = s.hv(w, x, with_dot=True)
            y_st, d_st = m.hv(x, with_dot=True)
            scale = max(1.0, float(y_mf.abs().max()))
            assert float((y_st - y_mf).abs().max()) <= 1e-13 * scale, (p, wi)
            assert abs(d_st - d_mf) <= 1e-11 * max(1.0, abs(d_mf)), (p, wi)
            m.close()


@pytest.mark.parametrize("p", [(10, 2, 9, 0, 0), (10, 2, 9, 3, 1), (10, 2, 9, 5, 0), (9, 4, 7, 2, 1)])
def test_matrix_free_matches_stored_on_large_blocks(torch, p):
    """Sectors whose hole blocks span many wavefronts: the matrix-free product walks them with one hole pattern
    per wavefront (exchange targets by closed-form rank changes, hop targets by chunked ranks), the stored form
    is filled by the per-lane walk; both vector precisions, every weight set."""
    from exactdiag_amd.sector import Sector

    s = Sector(*p)
    print("sector", p, "dim", s.dim, flush=True)
    for wi, w in enumerate(WEIGHTS):
        m = s.store(w)
        for dt, tol in ((torch.complex128, 1e-13), (torch.complex64, 1e-5)):
            x = _rand(torch, s.dim, wi, dt)
            y_mf, d_mf = s.hv(w, x, with_dot=True)
            y_st, d_st = m.hv(x, with_dot=True)
            scale = max(1.0, float(y_mf.abs().max()))
            assert float((y_st - y_mf).abs().max()) <= tol * scale, (p, wi, dt)
            assert abs(d_st - d_mf) <= 1e3 * tol * max(1.0, abs(d_mf)), (p, wi, dt)
        m.close()


def test_stored_matches_oracle(torch, oracle_mod, element_bytes):
    from exactdiag_amd.sector import Sector

    for p in SECTORS[:12]:
        s = Sector(*p)
        if s.dim == 0:
            continue
        o = oracle_mod.OracleSector(*p)
        for wi, w in enumerate(WEIGHTS):
            # the oracle's product through the reference's CSR dot (explicit + Hermitian passes); for complex
            # weights this is not conj(toarray), which conjugates the term matrices but not their weights
            x = np.random.default_rng(wi).normal(size=s.dim) + 1j * np.random.default_rng(wi + 9).normal(size=s.dim)
            ref = o.hv(w, x)
            y = s.store(w).hv(torch.as_tensor(x, device="cuda")).cpu().numpy()
            np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_stored_c64(torch, element_bytes):
    from exactdiag_amd.sector import Sector

    s = Sector(8, 2, 7, 1, 0)
    w = WEIGHTS[0]
    m = s.store(w)
    x = _rand(torch, s.dim, 3)
    y128 = m.hv(x)
    y64 = m.hv(x.to(torch.complex64))
    assert y64.dtype == torch.complex64
    rel = float((y64.to(torch.complex128) - y128).abs().max() / y128.abs().max())
    assert rel < 1e-6


def test_stored_lanczos_and_ritz(torch):
    from exactdiag_amd.sector import Sector

    s = Sector(7, 2, 6, 3, 0)
    w = WEIGHTS[1]
    m = s.store(w)
    v0 = _rand(torch, s.dim, 5)
    a1, b1 = s.lanczos(w, v0, 80)
    a2, b2 = m.lanczos(v0, 80)
    assert len(a1) == len(a2)
    np.testing.assert_allclose(a2, a1, rtol=0, atol=1e-10)
    np.testing.assert_allclose(b2, b1, rtol=0, atol=1e-10)
    # second pass: sum_n y_n v_n equals the same combination of the stored Lanczos vectors
    _, _, V = s.lanczos(w, v0, 40, keep_basis=True)
    y = np.random.default_rng(1).normal(size=len(V))
    ref = (torch.as_tensor(y, dtype=torch.complex128, device="cuda")[:, None] * V).sum(0)
    for x in (s.lanczos_ritz(w, v0, y), m.lanczos_ritz(v0, y)):
        assert float((x - ref).abs().max()) < 1e-10
    # c64 recurrence with FP64 accumulation of the Ritz vector
    x64 = m.lanczos_ritz(v0, y, precision=1)
    assert x64.dtype == torch.complex128
    assert float((x64 - ref).abs().max()) < 1e-5


def test_two_pass_ground_state(torch, oracle_mod):
    """Ground state from two-pass Lanczos on the stored matrix matches the oracle's dense spectrum."""
    from scipy.linalg import eigh_tridiagonal

    from exactdiag_amd.sector import Sector

    p = (6, 2, 5, 0, 0)
    w = WEIGHTS[0]
    s = Sector(*p)
    m = s.store(w)
    v0 = _rand(torch, s.dim, 11)
    a, b = m.lanczos(v0, 150)
    th, Y = eigh_tridiagonal(a, b[1:], select="i", select_range=(0, 0))
    x = m.lanczos_ritz(v0, Y[:, 0])
    x = x / x.norm()
    ev = np.linalg.eigvalsh(oracle_mod.OracleSector(*p).toarray(w))
    assert abs(th[0] - ev[0]) < 1e-10
    r = m.hv(x) - th[0] * x
    assert float(r.norm()) < 1e-8


def test_stored_tuner_keeps_results(torch):
    from exactdiag_amd.sector import Sector

    s = Sector(8, 2, 7, 1, 0)
    m = s.store(WEIGHTS[3])
    x = _rand(torch, s.dim, 8)
    y0, d0 = m.hv(x, with_dot=True)
    t, u, nt, ms = m.tune()
    assert t in (1, 2, 4, 8, 16, 32) and u in (4, 8) and nt in (0, 1) and ms > 0
    y1, d1 = m.hv(x, with_dot=True)
    assert torch.equal(y0, y1)  # rows are computed the same way whatever the tiling
    assert abs(d1 - d0) <= 1e-12 * max(1.0, abs(d0))
    for cfg in ((1, 8, True), (16, 4, True), (3, 4, False), (0, 8, False)):
        m.set_config(*cfg)
        y2 = m.hv(x)
        assert torch.equal(y0, y2), cfg
    with pytest.raises(ValueError):
        m.set_config(1, 5, False)


def test_momentum_sectors_of_the_2x12_ladder_pack(torch):
    """Only the orbit sizes present in the basis get values (2 holes on 2x12: every orbit has 24 states), so a
    sector with complex phases (k = (1,0)) keeps 4-byte elements; its product still matches the matrix-free one."""
    from exactdiag_amd.sector import Sector

    s = Sector(12, 2, 11, 1, 0)
    m = s.store([1.0, 1.0, 0.4, 0.4])
    print("2x12 k=(1,0): values", m.n_values, "element bytes", m.element_bytes, flush=True)
    assert m.element_bytes == 4
    x = _rand(torch, s.dim, 5)
    y_mf = s.hv([1.0, 1.0, 0.4, 0.4], x)
    y_st = m.hv(x)
    assert float((y_st - y_mf).abs().max()) <= 1e-13 * float(y_mf.abs().max())
    m.close()
