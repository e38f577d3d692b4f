"""GPU operators, Lanczos, eigenpairs and spectra against the oracle and exact sums."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _op_matrix(ptr, rows, vals, nf, ni):
    M = np.zeros((nf, ni), np.complex128)
    for i in range(ni):
        for k in range(ptr[i], ptr[i + 1]):
            M[rows[k], i] += vals[k]
    return M


OP_CASES = [
    ((4, 2, 3, 0, 0), 0, (2, 1), (4, 2, 3, 2, 1)),
    ((6, 2, 5, 0, 0), 0, (3, 1), (6, 2, 5, 3, 1)),
    ((6, 2, 5, 1, 0), 1, (2, 0), (6, 2, 5, 3, 0)),
    ((5, 1, 2, 0, 0), 2, (1, 0), (5, 0, 3, 1, 0)),
    ((5, 2, 4, 1, 1), 3, (2, 1), (5, 1, 4, -2, 0)),
    ((5, 1, 4, 0, 0), 4, (-1, 1), (5, 2, 3, -1, 1)),
    ((4, 1, 3, 0, 1), 5, (1, 0), (4, 2, 3, 1, 1)),
    ((5, 2, 4, 0, 0), 6, (1, 1), (5, 2, 4, 1, 1)),
    ((5, 2, 4, 0, 0), 7, (2, 0), (5, 2, 4, 2, 0)),
]


@pytest.mark.parametrize("case", OP_CASES)
def test_operator_columns_bit_exact(oracle_mod, torch, case):
    from exactdiag_amd.sector import Sector

    pi, op, (qx, qy), pf = case
    print("operator case", case, flush=True)
    oi, of = oracle_mod.OracleSector(*pi), oracle_mod.OracleSector(*pf)
    gi, gf = Sector(*pi), Sector(*pf)
    po, ro, vo = oi.op_columns(of, op, qx, qy)
    pg, rg, vg = gi.op_columns(gf, op, qx, qy)
    assert np.array_equal(po, pg) and np.array_equal(ro, rg) and np.array_equal(vo, vg)
    rng = np.random.default_rng(5)
    x = rng.normal(size=oi.dim) + 1j * rng.normal(size=oi.dim)
    y = gi.op_apply(gf, op, qx, qy, torch.as_tensor(x, device="cuda")).cpu().numpy()
    M = _op_matrix(po, ro, vo, of.dim, oi.dim)
    assert np.max(np.abs(y - M @ x)) < 1e-13 * max(1, np.max(np.abs(M @ x)))


def test_lanczos_coefficients_match_oracle(oracle_mod, torch):
    from exactdiag_amd.sector import Sector

    p = (6, 2, 5, 0, 0)
    o, g = oracle_mod.OracleSector(*p), Sector(*p)
    w = [1, 1, 0.4, 0.4]
    rng = np.random.default_rng(11)
    v0 = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
    # every coefficient of 60 steps (measured difference 3e-14); later steps of this 1,386-state sector part
    # ways by design of plain Lanczos: once orthogonality is lost, copies of converged Ritz values appear at steps
    # that depend on the rounding of each implementation (the two runs differ by 0.6 from about step 60 on)
    a, b = g.lanczos(w, torch.as_tensor(v0, device="cuda"), 60)
    ao, bo = o.lanczos(w, v0, 60)
    assert len(a) == len(ao) == 60
    assert np.max(np.abs(a - ao)) < 1e-12
    assert np.max(np.abs(b - bo)) < 1e-12
    assert b[0] == 0.0


def test_lanczos_keeps_basis_and_stops_on_invariant_subspace(oracle_mod, torch):
    from exactdiag_amd.sector import Sector

    p = (3, 1, 2, 0, 0)  # D = 10
    g = Sector(*p)
    w = [1, 1, 0.4, 0.4]
    rng = np.random.default_rng(4)
    v0 = torch.as_tensor(rng.normal(size=g.dim) + 1j * rng.normal(size=g.dim), device="cuda")
    a, b, V = g.lanczos(w, v0, 30, keep_basis=True)
    assert 4 <= len(a) <= g.dim  # the Krylov space is exhausted: stopped when b < 1e-8
    G = (V[:4].conj() @ V[:4].T).cpu().numpy()
    assert np.max(np.abs(G - np.eye(4))) < 1e-12
    # an eigenvector as start vector stops after one step, as the reference's loop does
    H = np.stack([g.hv(w, torch.eye(g.dim, dtype=torch.complex128, device="cuda")[:, k].contiguous()).cpu().numpy()
                  for k in range(g.dim)], axis=1)
    vals, vecs = np.linalg.eigh(H)
    a1, b1 = g.lanczos(w, torch.as_tensor(vecs[:, 0], device="cuda"), 30)
    assert len(a1) == 1 and abs(a1[0] - vals[0]) < 1e-12


def _lehmann(H, gs, O, e0, ws, eta):
    vals, vecs = np.linalg.eigh(H)
    phi = O @ gs
    wts = np.abs(vecs.conj().T @ phi) ** 2
    return np.array([np.sum(wts * eta / np.pi / ((w + e0 - vals) ** 2 + eta**2)) for w in ws])


def test_spectrum_equals_lehmann_sum(oracle_mod, torch):
    """Full-length Lanczos + continued fraction reproduce the exact spectral sum."""
    from exactdiag_amd.general.cython_lanczos_diagonalization import lanczos_dynamical_response
    from exactdiag_amd.general.sparse_matrices import Excitation_Operator, Hamiltonian_Operator
    from exactdiag_amd.sector import Sector

    pi, pf, (qx, qy) = (4, 2, 3, 0, 0), (4, 2, 3, 2, 1), (2, 1)
    w = [1, 1, 0.4, 0.4]
    oi, of = oracle_mod.OracleSector(*pi), oracle_mod.OracleSector(*pf)
    Hi, Hf = np.conj(oi.toarray(w)), np.conj(of.toarray(w))
    ev, evec = np.linalg.eigh(Hi)
    gs = evec[:, 0]
    O = _op_matrix(*oi.op_columns(of, 0, qx, qy), of.dim, oi.dim)
    ws = np.linspace(0, 4, 401)
    eta = 0.05
    exact = _lehmann(Hf, gs, O, ev[0], ws, eta)
    gi, gf = Sector(*pi), Sector(*pf)
    spec = lanczos_dynamical_response(gs, Excitation_Operator(gi, gf, 0, qx, qy), Hamiltonian_Operator(gf, w),
                                      ws + ev[0] + 1j * eta, of.dim + 5)
    assert np.max(np.abs(spec - exact)) < 1e-8 * max(1.0, np.max(exact))
    assert np.trapz(spec, ws) > 0


def test_cf_kernel_matches_oracle(oracle_mod):
    from exactdiag_amd.sector import cf_eval

    rng = np.random.default_rng(3)
    a = rng.normal(size=200)
    b = np.abs(rng.normal(size=200)) + 0.1
    b[0] = 0
    ws = np.linspace(-2, 6, 2001)
    zs = ws + 0.3 + 1j * 0.05
    ref = oracle_mod.cf_imag(a, b, zs)
    got = cf_eval(a, b, ws, 0.3, 0.05, 1.0)
    assert np.max(np.abs(got - ref)) < 1e-10 * max(1, np.max(np.abs(ref)))


@pytest.mark.parametrize("p", [(6, 2, 5, 0, 0), (6, 1, 5, 2, 1), (7, 2, 6, 1, 0)])
def test_thick_restart_eigenpairs(oracle_mod, torch, p):
    from exactdiag_amd.general import krylov
    from exactdiag_amd.sector import Sector

    w = [0.45, 0.36, 0.15, 0.12]
    o, g = oracle_mod.OracleSector(*p), Sector(*p)
    exact = np.linalg.eigvalsh(np.conj(o.toarray(w)))[:5]
    vals, X, _ = krylov.lowest_eigenpairs(lambda v: g.hv(w, v), g.dim, 5, tol=1e-11)
    assert np.max(np.abs(vals - exact)) < 1e-10
    for i in range(5):
        r = g.hv(w, X[i]) - vals[i] * X[i]
        assert krylov.vec_norm(r) < 1e-9


def test_precision_c64_hv_error(oracle_mod, torch):
    from exactdiag_amd.sector import Sector

    p = (6, 2, 5, 0, 0)
    o, g = oracle_mod.OracleSector(*p), Sector(*p)
    w = [1, 1, 0.4, 0.4]
    rng = np.random.default_rng(9)
    x = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
    y64 = g.hv(w, torch.as_tensor(x, device="cuda").to(torch.complex64)).cpu().numpy()
    ref = o.hv(w, x.astype(np.complex64).astype(np.complex128))
    rel = np.linalg.norm(y64 - ref) / np.linalg.norm(ref)
    assert rel < 1e-6, rel
