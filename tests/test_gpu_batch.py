"""Batched Lanczos engine: against the single-sector path, the oracle and the reference's fixtures.

The batch runs sectors of very different sizes (from a handful of states to
~1e4) in one pool: every instance must reproduce what the single-sector
kernels give for it alone (H·v bit for bit, Lanczos coefficients to rounding
of the reduction order), the HIP-graph replay must equal the eager launches,
and the batched lowest eigenpairs must meet the reference's eigenvalue
fixtures (tests/tJ_spin_half_ladder/test_eigvals.py:9-22, tolerance 1e-8).
"""

import glob
import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).parent / "golden" / "precalculated_basis_size_eigenvalues"
MIXED = [(6, 2, 5, 0, 0), (6, 2, 5, 3, 1), (5, 2, 4, 2, 0), (4, 3, 2, 2, 0), (7, 2, 6, 3, 0), (3, 1, 2, 0, 0),
         (2, 0, 2, 0, 0), (8, 2, 7, 1, 0), (4, 0, 4, 2, 1)]
W = [1.0, 1.0, 0.4, 0.4]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def mixed(torch):
    from exactdiag_amd.sector import Sector

    return [Sector(*p) for p in MIXED]


def test_batch_hv_matches_single(torch, mixed):
    from exactdiag_amd.batch import Batch

    b = Batch(mixed, W)
    x = b.random(3)
    y, dots = b.hv(x, with_dots=True)
    for i, s in enumerate(mixed):
        print("instance", MIXED[i], "dim", s.dim, flush=True)
        yi, di = s.hv(W, b.part(x, i).clone(), with_dot=True)
        err = float((b.part(y, i) - yi).abs().max()) if s.dim else 0.0
        assert err <= 1e-14 * max(1.0, float(yi.abs().max())), (MIXED[i], err)
        assert abs(dots[i] - di) <= 1e-11 * max(1.0, abs(di)), MIXED[i]


def test_batch_weights_per_instance(torch, mixed):
    from exactdiag_amd.batch import Batch

    s = mixed[0]
    ws = [[1.0, 1.0, 0.4, 0.4], [0.45, 0.36, 0.15, 0.12], [0.0, 0.0, 1.0, 1.0], [1.0, 0.0, 0.0, 0.3]]
    b = Batch([s] * len(ws), ws)
    x = b.random(4)
    y = b.hv(x)
    for i, w in enumerate(ws):
        yi = s.hv(w, b.part(x, i).clone())
        assert float((b.part(y, i) - yi).abs().max()) <= 1e-14 * max(1.0, float(yi.abs().max())), w


def test_batch_lanczos_matches_single(torch, mixed):
    from exactdiag_amd.batch import Batch

    b = Batch(mixed, W)
    v0 = b.random(5)
    a, bb, _ = b.lanczos(v0, 60)
    for i, s in enumerate(mixed):
        ai, bi = s.lanczos(W, b.part(v0, i).clone(), 60)
        print("instance", MIXED[i], "dim", s.dim, "steps", len(ai), len(a[i]), flush=True)
        assert len(ai) == len(a[i]), MIXED[i]
        # the two engines sum the elements of a row in different orders; once a run spans a fair part of a
        # small space its coefficients amplify rounding, so the recurrences are compared over the first
        # quarter of the space and the converged lowest Ritz value over the whole run
        n = min(len(ai), max(1, s.dim // 4))
        np.testing.assert_allclose(a[i][:n], ai[:n], rtol=0, atol=1e-10)
        np.testing.assert_allclose(bb[i][:n], bi[:n], rtol=0, atol=1e-10)
        lo = [np.linalg.eigvalsh(np.diag(aa) + np.diag(bv[1 : len(aa)], 1) + np.diag(bv[1 : len(aa)], -1))[0]
              for aa, bv in ((np.asarray(a[i]), np.asarray(bb[i])), (np.asarray(ai), np.asarray(bi)))]
        assert abs(lo[0] - lo[1]) < 1e-9, (MIXED[i], lo)


def test_graph_replay_equals_eager(torch, mixed):
    from exactdiag_amd.batch import Batch

    b = Batch(mixed, W)
    v0 = b.random(6)
    steps = np.array([40, 13, 0, 25, 40, 7, 3, 40, 1], np.int32)
    eager = b.lanczos(v0, steps, graph=False)
    for rep in range(3):
        torch.cuda.synchronize()
        got = b.lanczos(v0, steps, graph=True)
        for i in range(len(mixed)):
            assert np.array_equal(got[0][i], eager[0][i]), (rep, i)
            assert np.array_equal(got[1][i], eager[1][i]), (rep, i)
    for i, n in enumerate(steps):
        assert len(eager[0][i]) <= n


def test_two_pass_ritz_vector(torch, mixed):
    """acc = sum_n y_n v_n from the second pass equals the same combination of the stored basis."""
    from exactdiag_amd.batch import Batch

    b = Batch(mixed[:3], W)
    v0 = b.random(7)
    a, bb, _ = b.lanczos(v0, 30)
    rng = np.random.default_rng(0)
    ritz = [rng.normal(size=len(ai)) for ai in a]
    _, _, acc = b.lanczos(v0, [len(ai) for ai in a], ritz=ritz)
    for i, s in enumerate(mixed[:3]):
        _, _, V = s.lanczos(W, b.part(v0, i).clone(), len(a[i]), keep_basis=True)
        ref = (torch.as_tensor(ritz[i], dtype=torch.complex128, device=V.device)[:, None] * V).sum(0)
        assert float((b.part(acc, i) - ref).abs().max()) < 1e-10


def _fixtures():
    out = []
    for f in sorted(glob.glob(str(GOLD / "*.txt"))):
        d = json.load(open(f, encoding="utf-8"))
        if "eigvals" in d:
            out.append(d)
    return out


def test_batched_lowest_eigenpairs_fixtures(torch, oracle_mod):
    """All 51 eigenvalue fixtures as one batch (each with its own weights)."""
    from exactdiag_amd.batch import Batch, lowest_eigenpairs
    from exactdiag_amd.sector import Sector

    fx = _fixtures()
    assert len(fx) == 51
    sectors, weights = [], []
    for d in fx:
        k = d["symmetry_qs"]
        sectors.append(Sector(d["num_rungs"], d["num_holes"], d["num_down_spins"], k[0], k[1]))
        w = d["weights"]
        weights.append([w["tl"], w["tr"], w["jl"], w["jr"]])
    b = Batch(sectors, weights)
    e, x, restarts = lowest_eigenpairs(b, tol=1e-10)
    print("restarts", restarts, flush=True)
    for i, d in enumerate(fx):
        assert abs(e[i] - d["eigvals"][0]) < 1e-8, (d["num_rungs"], d["num_holes"], d["symmetry_qs"], e[i])
        s = sectors[i]
        if s.dim <= 1500:
            o = oracle_mod.OracleSector(s.L, s.h, s.nd, s.kx, s.ky)
            ev = np.linalg.eigvalsh(o.toarray(weights[i]))
            assert abs(e[i] - ev[0]) < 1e-10, (s.L, s.h, s.kx, s.ky)
        y = s.hv(weights[i], b.part(x, i).clone())
        r = float((y - e[i] * b.part(x, i)).abs().max())
        assert r < 1e-8, (s.L, s.h, s.kx, s.ky, r)


def test_batched_spectra_match_serial(torch):
    from exactdiag_amd import _lib
    from exactdiag_amd.allq import final_momentum, ground_state
    from exactdiag_amd.batch import Excitation, spectra
    from exactdiag_amd.sector import Sector, cf_eval, weights_array
    from exactdiag_amd.general.krylov import vec_norm

    L, h, nd = 6, 2, 5
    s0, e0, psi = ground_state(L, h, nd, 0, 0, W, tol=1e-12)
    exc = []
    for qx in range(L):
        for qy in range(2):
            kx, ky = final_momentum(L, 0, 0, qx, qy)
            exc.append(Excitation("Szq", _lib.OP_SZQ, qx, qy, Sector(L, h, nd, kx, ky)))
            kx, ky = final_momentum(L, 0, 0, qx, qy)
            exc.append(Excitation("A+", _lib.OP_APLUS_DOWN, qx, qy, Sector(L, h - 1, nd + 1, kx, ky)))
            kx, ky = final_momentum(L, 0, 0, -qx, (-qy) % 2)
            exc.append(Excitation("A-", _lib.OP_AMINUS_DOWN, -qx, (-qy) % 2, Sector(L, h + 1, nd - 1, kx, ky)))
    ws = np.linspace(-2, 6, 801)
    # 30 steps: before the smallest sectors (a few hundred states) lose orthogonality, where the two
    # reduction orders would start to drift apart through ghost eigenvalues
    nsteps = 30
    got, coeffs = spectra(exc, s0, psi, e0, W, ws, 0.05, nsteps)
    for ex, g, (a, b) in zip(exc, got, coeffs):
        phi = s0.op_apply(ex.sector, ex.op, ex.qx, ex.qy, psi)
        n2 = vec_norm(phi) ** 2 if ex.sector.dim else 0.0
        if n2 < 1e-16:
            assert not np.any(g)
            continue
        ar, br = ex.sector.lanczos(weights_array(W), phi, nsteps)
        ref = cf_eval(ar, br, ws, e0, 0.05, -n2 / np.pi)
        err = np.max(np.abs(g - ref)) / np.max(np.abs(ref))
        assert err < 1e-8, (ex.name, ex.qx, ex.qy, err)


def test_autotuner_keeps_results(torch, mixed):
    from exactdiag_amd.batch import Batch

    b = Batch(mixed, W)
    x = b.random(8)
    y0 = b.hv(x).clone()
    tiles, gather, ms = b.tune()
    assert tiles in (1, 2, 4, 8, 16, 32) and gather in (2, 4, 8) and ms > 0
    assert b.config()[:2] == (tiles, gather)
    y1 = b.hv(x)
    assert float((y1 - y0).abs().max()) <= 1e-14 * float(y0.abs().max())
    for t, g in ((1, 2), (32, 8), (3, 4)):
        b.set_config(t, g)
        assert float((b.hv(x) - y0).abs().max()) <= 1e-14 * float(y0.abs().max())
    with pytest.raises(ValueError):
        b.set_config(4, 3)
    s = mixed[7]
    xs = b.part(x, 7).clone()
    ys0 = s.hv(W, xs)
    t, g, ms = s.tune(W)
    assert t in (1, 2, 4, 8, 16, 32) and g in (2, 4, 8)
    assert float((s.hv(W, xs) - ys0).abs().max()) <= 1e-14 * float(ys0.abs().max())


def test_stored_batch_matches_matrix_free_batch(torch, mixed):
    """Instances reading their stored H give the matrix-free batch's products and Lanczos coefficients."""
    from exactdiag_amd.batch import Batch

    ws = [[1.0, 1.0, 0.4, 0.4], [0.45, 0.36, 0.15, 0.12], [1.0 + 0.3j, 0.7, 0.4j, 0.25]]
    inst = [(s, ws[i % 3]) for i, s in enumerate(mixed)] + [(mixed[0], ws[1]), (mixed[0], ws[1])]
    b0 = Batch([s for s, _ in inst], [w for _, w in inst])
    b1 = Batch([s for s, _ in inst], [w for _, w in inst], stored=True)
    x = b0.random(9)
    y0, d0 = b0.hv(x, with_dots=True)
    y1, d1 = b1.hv(x, with_dots=True)
    assert float((y1 - y0).abs().max()) <= 1e-13 * float(y0.abs().max())
    np.testing.assert_allclose(d1, d0, rtol=1e-12, atol=1e-12)
    a0, bb0, _ = b0.lanczos(x, 30)
    a1, bb1, _ = b1.lanczos(x, 30)
    for i in range(len(inst)):
        assert len(a0[i]) == len(a1[i])
        np.testing.assert_allclose(a1[i], a0[i], rtol=0, atol=1e-10)
    tiles, gather, ms = b1.tune()
    assert float((b1.hv(x) - y1).abs().max()) <= 1e-14 * float(y1.abs().max())
    b1.close()
