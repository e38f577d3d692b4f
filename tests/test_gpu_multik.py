"""Multi-momentum stored product (ed_mk_*): every slot against its own sector.

One stored structure over the union basis carries up to 16 momentum sectors
of a ladder.  Each slot must reproduce, for its sector, what the
single-sector stored H gives (itself pinned against the oracle and the
reference's eigenvalue fixtures): the product to rounding (the element values
are formed in another order, so not bit for bit), the fused <x, Hx>, and the
Lanczos coefficients of general/cython_lanczos_diagonalization.pyx:18-77,
including where a slot stops at b < 1e-8.  Rows outside a slot's sector stay
zero.  The union basis holds every sector's states with the same orbit sizes
(the attach kernel checks both) and nothing but lowest images.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = [1.0, 1.0, 0.4, 0.4]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _momenta(L):
    from exactdiag_amd.allq import distinct_qs, final_momentum

    return [final_momentum(L, 0, 0, qx, qy) for qx, qy in distinct_qs(L)]


def _setup(L, h, nd):
    from exactdiag_amd.multik import MultiK
    from exactdiag_amd.sector import Sector

    u = Sector.union(L, h, nd)
    ks = _momenta(L)
    secs = [Sector(L, h, nd, kx, ky) for kx, ky in ks]
    return u, secs, MultiK(u, W, ks)


def _rand(torch, n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    re = torch.randn(max(n, 1), dtype=torch.float64, device="cuda", generator=g)
    im = torch.randn(max(n, 1), dtype=torch.float64, device="cuda", generator=g)
    return torch.complex(re, im)[:n].contiguous()


@pytest.fixture(scope="module")
def ladder6(torch):
    return _setup(6, 2, 5)


@pytest.mark.parametrize("L,h,nd", [(4, 2, 3), (5, 1, 4), (6, 2, 5)])
def test_union_holds_every_sector_and_only_lowest_images(torch, L, h, nd):
    from exactdiag_amd.sector import Sector

    print("union", L, h, nd, flush=True)
    u = Sector.union(L, h, nd)
    us, uo = u.basis()
    assert np.all(np.diff(us) > 0)
    all_k = [Sector(L, h, nd, kx, ky) for kx in range(L) for ky in (0, 1)]
    covered = np.zeros(len(us), bool)
    for s in all_k:
        ss, so = s.basis()
        at = np.searchsorted(us, ss)
        assert np.all(us[at] == ss), (s.kx, s.ky)
        assert np.array_equal(uo[at], so), (s.kx, s.ky)
    # every union state is its own lowest image in some sector
    for s in all_k:
        rep, _ = s.lowest(us)
        covered |= rep == us
    assert covered.all()


def test_hv_matches_each_sector(torch, ladder6):
    from exactdiag_amd.sector import StoredH

    u, secs, mk = ladder6
    x = mk.block()
    xs = []
    for k, s in enumerate(secs):
        xs.append(_rand(torch, s.dim, 10 + k))
        mk.attach(k, s, xs[-1], x)
    y, dots = mk.hv(x, with_dot=True)
    for k, s in enumerate(secs):
        print("slot", k, (s.kx, s.ky), "dim", s.dim, flush=True)
        hm = StoredH(s, W)
        yk, dk = hm.hv(xs[k].clone(), with_dot=True)
        got = mk.extract(k, s, y)
        err = float((got - yk).abs().max())
        assert err <= 1e-13 * max(1.0, float(yk.abs().max())), (k, err)
        assert abs(dots[k] - dk) <= 1e-11 * max(1.0, abs(dk)), (k, dots[k], dk)
        back = mk.extract(k, s, x)
        assert float((back - xs[k]).abs().max()) <= 1e-14 * float(xs[k].abs().max())
        hm.close()


def test_rows_outside_a_slot_stay_zero(torch, ladder6):
    u, secs, mk = ladder6
    x = mk.block()
    for k, s in enumerate(secs):
        mk.attach(k, s, _rand(torch, s.dim, 30 + k), x)
    y = mk.hv(x).cpu().numpy()
    us, _ = u.basis()
    for k, s in enumerate(secs):
        ss, _ = s.basis()
        member = np.isin(us, ss)
        assert not np.any(y[~member, k]), k
        assert not np.any(y[:, len(secs):])


def test_hv_complex64(torch, ladder6):
    from exactdiag_amd.sector import StoredH

    u, secs, mk = ladder6
    x = mk.block()
    xs = []
    for k, s in enumerate(secs):
        xs.append(_rand(torch, s.dim, 50 + k))
        mk.attach(k, s, xs[-1], x)
    y = mk.hv(x.to(torch.complex64)).to(torch.complex128)
    for k, s in enumerate(secs):
        hm = StoredH(s, W)
        yk = hm.hv(xs[k].clone())
        got = mk.extract(k, s, y)
        err = float((got - yk).abs().max())
        assert err <= 2e-5 * float(yk.abs().max()), (k, err)
        hm.close()


@pytest.mark.parametrize("precision,tol,n_cmp", [(0, 1e-10, 60), (1, 1e-5, 30)])
def test_lanczos_matches_each_sector(torch, ladder6, precision, tol, n_cmp):
    """complex128: all 60 steps to 1e-10.  complex64: two runs that differ in rounding drift apart once the
    extreme Ritz values have converged (the sector has ~1400 states), so the first 30 steps are compared."""
    from exactdiag_amd.sector import StoredH

    u, secs, mk = ladder6
    v0 = mk.block()
    starts = []
    for k, s in enumerate(secs):
        starts.append(_rand(torch, s.dim, 70 + k))
        mk.attach(k, s, starts[-1], v0)
    got = mk.lanczos(v0, 60, precision=precision)
    for k, s in enumerate(secs):
        hm = StoredH(s, W)
        a, b = hm.lanczos(starts[k], 60, precision=precision)
        hm.close()
        print("slot", k, "steps", len(a), len(got[k][0]), flush=True)
        assert len(got[k][0]) == len(a), k
        scale = max(1.0, float(np.abs(a).max()))
        err_a = np.abs(got[k][0][:n_cmp] - a[:n_cmp]).max()
        assert err_a <= tol * scale, (k, err_a)
        assert np.abs(got[k][1][:n_cmp] - b[:n_cmp]).max() <= tol * scale, k


def test_lanczos_matches_oracle_small(torch, oracle_mod):
    """Slot coefficients against the oracle's Lanczos (general/cython_lanczos_diagonalization.pyx:18-77)."""
    u, secs, mk = _setup(4, 2, 3)
    v0 = mk.block()
    starts = []
    for k, s in enumerate(secs):
        rng = np.random.default_rng(90 + k)
        v = rng.standard_normal(s.dim) + 1j * rng.standard_normal(s.dim)
        starts.append(v)
        mk.attach(k, s, torch.as_tensor(v, device="cuda"), v0)
    got = mk.lanczos(v0, 12)
    for k, s in enumerate(secs):
        o = oracle_mod.OracleSector(4, 2, 3, s.kx, s.ky)
        a, b = o.lanczos(W, starts[k], 12)
        n = min(len(a), len(got[k][0]))
        assert len(got[k][0]) == len(a), (k, len(got[k][0]), len(a))
        assert np.allclose(got[k][0][:n], a[:n], atol=1e-10, rtol=0), k
        assert np.allclose(got[k][1][:n], b[:n], atol=1e-10, rtol=0), k


def test_slots_stop_independently(torch):
    """Small sectors exhaust their Krylov space at different steps; each slot stops as its own run does."""
    from exactdiag_amd.sector import StoredH

    u, secs, mk = _setup(3, 1, 1)
    v0 = mk.block()
    starts = []
    for k, s in enumerate(secs):
        starts.append(_rand(torch, s.dim, 110 + k))
        mk.attach(k, s, starts[-1], v0)
    got = mk.lanczos(v0, 40)
    for k, s in enumerate(secs):
        if s.dim == 0:
            assert len(got[k][0]) == 0
            continue
        hm = StoredH(s, W)
        a, b = hm.lanczos(starts[k], 40)
        hm.close()
        assert len(got[k][0]) == len(a) <= s.dim, (k, len(got[k][0]), len(a), s.dim)
        assert np.allclose(got[k][0], a, atol=1e-10, rtol=0), k


def test_attach_checks_the_slot_momentum(torch, ladder6):
    from exactdiag_amd.sector import Sector

    u, secs, mk = ladder6
    other = secs[1]
    x = mk.block()
    with pytest.raises(ValueError):
        mk.attach(0, other, _rand(torch, other.dim, 1), x)
    with pytest.raises(ValueError):
        mk.attach(len(secs), secs[0], _rand(torch, secs[0].dim, 1), x)
    with pytest.raises(ValueError):
        from exactdiag_amd.multik import MultiK

        MultiK(Sector(6, 2, 5, 0, 0), W, [(0, 0)])


def test_2x10_all_q(torch):
    """All 12 distinct q of the 2x10 ladder with 2 holes (union D ~ 4.6e5): products and 30 Lanczos steps."""
    from exactdiag_amd.sector import StoredH

    u, secs, mk = _setup(10, 2, 9)
    print("union dim", u.dim, "slots", len(secs), flush=True)
    v0 = mk.block()
    starts = []
    for k, s in enumerate(secs):
        starts.append(_rand(torch, s.dim, 200 + k))
        mk.attach(k, s, starts[-1], v0)
    y = mk.hv(v0)
    got = mk.lanczos(v0, 30)
    for k, s in enumerate(secs):
        hm = StoredH(s, W)
        yk = hm.hv(starts[k].clone())
        err = float((mk.extract(k, s, y) - yk).abs().max())
        assert err <= 1e-13 * float(yk.abs().max()), (k, err)
        a, b = hm.lanczos(starts[k], 30)
        hm.close()
        assert np.abs(got[k][0] - a).max() <= 1e-10 * max(1.0, np.abs(a).max()), k
        assert np.abs(got[k][1] - b).max() <= 1e-10 * max(1.0, np.abs(a).max()), k


@pytest.mark.parametrize("L,h,nd", [(2, 0, 2), (3, 0, 3), (4, 0, 4), (2, 4, 0), (3, 5, 0), (4, 1, 3), (5, 3, 1),
                                    (6, 4, 4)])
def test_small_and_edge_ladders(torch, L, h, nd):
    """No holes (exchange only), all holes (one state, no elements), one electron, odd electron counts: every
    slot's product and 20 Lanczos steps against its own sector's stored H."""
    from exactdiag_amd.sector import StoredH

    u, secs, mk = _setup(L, h, nd)
    print("ladder", L, h, nd, "union", u.dim, "sectors", [s.dim for s in secs], flush=True)
    assert u.dim >= max(s.dim for s in secs)
    x = mk.block()
    xs = []
    for k, s in enumerate(secs):
        xs.append(_rand(torch, s.dim, 300 + k))
        mk.attach(k, s, xs[-1], x)
    y = mk.hv(x)
    got = mk.lanczos(x, 20)
    for k, s in enumerate(secs):
        if s.dim == 0:
            assert len(got[k][0]) == 0
            continue
        hm = StoredH(s, W)
        yk = hm.hv(xs[k].clone())
        err = float((mk.extract(k, s, y) - yk).abs().max())
        assert err <= 1e-13 * max(1.0, float(yk.abs().max())), (k, err)
        a, b = hm.lanczos(xs[k], 20)
        hm.close()
        assert len(got[k][0]) == len(a), (k, len(got[k][0]), len(a))
        assert np.allclose(got[k][0], a, atol=1e-10, rtol=0), k
        assert np.allclose(got[k][1], b, atol=1e-10, rtol=0), k


def test_per_slot_weights(torch):
    """Slots may carry different couplings: each slot's product against its sector's stored H for its weights."""
    from exactdiag_amd.multik import MultiK
    from exactdiag_amd.sector import Sector, StoredH

    L, h, nd = 6, 2, 5
    u = Sector.union(L, h, nd)
    ws = [[1.0, 1.0, 0.3, 0.3], [0.45, 0.36, 0.15, 0.12], [0.0, 0.0, 1.0, 1.0], [1.0 + 0.3j, 0.7 - 0.2j, 0.4j, 0.25],
          [1.0, 0.0, 0.0, 0.3]]
    slots = [(kx, ky, wi) for kx, ky in [(0, 0), (3, 1), (2, 0)] for wi in range(len(ws))]
    mk = MultiK(u, [ws[wi] for _, _, wi in slots], [(kx, ky) for kx, ky, _ in slots])
    secs = {(kx, ky): Sector(L, h, nd, kx, ky) for kx, ky, _ in slots}
    x = mk.block()
    xs = []
    for k, (kx, ky, wi) in enumerate(slots):
        xs.append(_rand(torch, secs[(kx, ky)].dim, 400 + k))
        mk.attach(k, secs[(kx, ky)], xs[-1], x)
    y = mk.hv(x)
    for k, (kx, ky, wi) in enumerate(slots):
        s = secs[(kx, ky)]
        m = StoredH(s, ws[wi])
        yk = m.hv(xs[k].clone())
        m.close()
        err = float((mk.extract(k, s, y) - yk).abs().max())
        assert err <= 1e-13 * max(1.0, float(yk.abs().max())), (k, slots[k], err)


def test_lowest_eigenpairs_all_slots(torch, oracle_mod):
    """Two-pass lowest eigenpair of every slot: energies against the oracle's dense spectrum (2x4, all 8
    momenta) and against one-sector runs (2x8, 2 holes, 10 momenta x 2 couplings), eigenvectors by residual."""
    from exactdiag_amd.forms import two_pass_ground_state
    from exactdiag_amd.general import krylov
    from exactdiag_amd.multik import MultiK
    from exactdiag_amd.sector import Sector

    # small: dense oracle spectra
    L, h, nd = 4, 2, 3
    ks = [(kx, ky) for kx in range(L) for ky in (0, 1)]
    u = Sector.union(L, h, nd)
    mk = MultiK(u, W, ks)
    secs = [Sector(L, h, nd, *k) for k in ks]
    v0 = mk.block()
    for k, s in enumerate(secs):
        if s.dim:
            mk.attach(k, s, krylov.device_start_vector(s.dim, k), v0)
    e, x, _, res = mk.lowest(v0, tol=1e-10)
    for k, s in enumerate(secs):
        if s.dim == 0:
            assert np.isnan(e[k])
            continue
        exact = np.linalg.eigvalsh(np.conj(oracle_mod.OracleSector(L, h, nd, *ks[k]).toarray(W)))[0]
        assert abs(e[k] - exact) <= 1e-10 * max(1.0, abs(exact)), (ks[k], e[k], exact)
    # larger: one-sector two-pass runs, two couplings
    L, h, nd = 8, 2, 7
    ws = [[1.0, 1.0, 0.3, 0.3], [1.0, 1.0, 0.5, 0.5]]
    slots = [(kx, 0, wi) for kx in range(5) for wi in range(2)] + [(kx, 1, wi) for kx in range(3) for wi in range(2)]
    u = Sector.union(L, h, nd)
    mk = MultiK(u, [ws[wi] for *_, wi in slots], [(kx, ky) for kx, ky, _ in slots])
    secs = {(kx, ky): Sector(L, h, nd, kx, ky) for kx, ky, _ in slots}
    v0 = mk.block()
    for k, (kx, ky, wi) in enumerate(slots):
        mk.attach(k, secs[(kx, ky)], krylov.device_start_vector(secs[(kx, ky)].dim, 50 + k), v0)
    e, x, restarts, res = mk.lowest(v0, tol=1e-10)
    print("restarts", restarts, "max residual", float(np.max(res)), flush=True)
    for k, (kx, ky, wi) in enumerate(slots):
        s = secs[(kx, ky)]
        hm = s.store(ws[wi])
        e1, _, _, _ = two_pass_ground_state(hm, tol=1e-10)
        hm.close()
        assert abs(e[k] - e1) <= 1e-9 * max(1.0, abs(e1)), (slots[k], e[k], e1)
        # the recurrence stops at b < 1e-8 (the reference's threshold), so residuals below that are not pursued
        assert res[k] <= max(1e-10 * abs(e[k]), 1e-8), (slots[k], res[k])
        # the slot's vector, taken back to the sector, is an eigenvector of its own stored H
        psi = mk.extract(k, s, x)
        hm = s.store(ws[wi])
        r = hm.hv(psi.clone()) - e[k] * psi
        hm.close()
        assert float(r.abs().max()) <= 1e-8, slots[k]


def test_lowest_eigenpairs_reference_fixtures(torch):
    """All 51 eigenvalue fixtures of the reference (tests/tJ_spin_half_ladder/test_eigvals.py:9-22, tolerance
    1e-8), grouped by ladder: every ladder's sectors and couplings as slots of multi-momentum products."""
    import glob
    import json
    import pathlib

    from exactdiag_amd.general import krylov
    from exactdiag_amd.multik import SLOTS, MultiK
    from exactdiag_amd.sector import Sector

    gold = pathlib.Path(__file__).parent / "golden" / "precalculated_basis_size_eigenvalues"
    fx = [d for d in (json.load(open(f, encoding="utf-8")) for f in sorted(glob.glob(str(gold / "*.txt"))))
          if "eigvals" in d]
    assert len(fx) == 51
    ladders = {}
    for d in fx:
        ladders.setdefault((d["num_rungs"], d["num_holes"], d["num_down_spins"]), []).append(d)
    checked = 0
    for (L, h, nd), items in ladders.items():
        u = Sector.union(L, h, nd)
        for lo in range(0, len(items), SLOTS):
            grp = items[lo: lo + SLOTS]
            ws = [[d["weights"][t] for t in ("tl", "tr", "jl", "jr")] for d in grp]
            ks = [tuple(d["symmetry_qs"]) for d in grp]
            mk = MultiK(u, ws, ks)
            secs = [Sector(L, h, nd, *k) for k in ks]
            v0 = mk.block()
            for slot, s in enumerate(secs):
                mk.attach(slot, s, krylov.device_start_vector(s.dim, slot), v0)
            e, x, _, res = mk.lowest(v0, tol=1e-10)
            for slot, d in enumerate(grp):
                assert abs(e[slot] - d["eigvals"][0]) < 1e-8, (L, h, nd, ks[slot], e[slot], d["eigvals"][0])
                checked += 1
            mk.close()
        print("ladder", (L, h, nd), "union", u.dim, "fixtures", len(items), flush=True)
    assert checked == 51
