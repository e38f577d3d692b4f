"""The time-to-accuracy Lanczos run (ed_hmat_lanczos_run_tol) on the GPU.

The run uses the same kernels as the fixed-length run, so its coefficients are the first n of a fixed run from
the same start vector bit for bit; it stops at the first checked n where the rule's estimate is below tol; and
the fast all-q form that uses it meets the 1e-3 target against 1000-step FP64 spectra.
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


def _start(torch, dim, seed):
    rng = np.random.default_rng(seed)
    return torch.as_tensor(rng.normal(size=dim) + 1j * rng.normal(size=dim), device="cuda")


@pytest.mark.parametrize("precision", [0, 1])
def test_run_is_a_prefix_of_the_fixed_run(torch, precision):
    from exactdiag_amd.sector import Sector, StopRule

    s = Sector(8, 2, 7, 1, 0)
    hm = s.store(W)
    v0 = _start(torch, s.dim, 11)
    print("2x8 h2 k=(1,0), dim", s.dim, "precision", precision, flush=True)
    af, bf = hm.lanczos(v0, 300, precision=precision)
    e0 = float(np.min(af)) - 2.0
    rule = StopRule(np.linspace(0.0, 4.0, 2001), e0, 0.05, 2e-3, min_steps=40, check_every=5, delta=20)
    a, b, why = hm.lanczos_to_tol(v0, 300, rule, precision=precision)
    n = len(a)
    print("stopped at", n, "reason", why, flush=True)
    assert why == 1 and 40 <= n < 300
    assert np.array_equal(a, af[:n]) and np.array_equal(b, bf[:n])
    assert rule.change(af, bf, n) < 2e-3
    for m in range(40, n, 5):  # every earlier check failed the rule
        assert rule.change(af, bf, m) >= 2e-3
    hm.close()
    s.close()


def test_cap_and_exhausted_space(torch):
    from exactdiag_amd.sector import Sector, StopRule

    s = Sector(8, 2, 7, 1, 0)
    hm = s.store(W)
    v0 = _start(torch, s.dim, 12)
    ws = np.linspace(0.0, 4.0, 401)
    a, b, why = hm.lanczos_to_tol(v0, 33, StopRule(ws, -10.0, 0.05, 1e-14, min_steps=20, check_every=5), precision=0)
    assert why == 0 and len(a) == 33
    af, bf = hm.lanczos(v0, 33)
    assert np.array_equal(a, af) and np.array_equal(b, bf)
    hm.close()
    s.close()
    # a sector of 16 states: the Krylov space ends before the rule can be met
    t = Sector(3, 2, 2, 1, 0)
    assert 0 < t.dim < 40
    hm = t.store(W)
    v0 = _start(torch, t.dim, 13)
    a, b, why = hm.lanczos_to_tol(v0, 100, StopRule(ws, -10.0, 0.05, 1e-14, min_steps=20), precision=0)
    af, bf = hm.lanczos(v0, 100)
    print("small sector dim", t.dim, "steps", len(a), "reason", why, flush=True)
    assert why == 2 and len(a) == len(af) <= t.dim
    assert np.array_equal(a, af) and np.array_equal(b, bf)
    hm.close()
    t.close()


def test_fast_allq_with_the_rule_meets_the_target(torch):
    from exactdiag_amd import allq

    ws = np.linspace(0.0, 4.0, 2001)
    e0, ref = allq.allq_reference(8, 2, 7, W, ws, 0.05, 1000)
    _, fast = allq.allq_fast(8, 2, 7, W, ws, 0.05, 300, spectrum_tol=1e-3)
    steps = dict(allq.LAST_FAST_STEPS)
    err = {q: allq.spectral_l1(fast[q], ref[q], ws) for q in ref}
    print("2x8 all-q: steps", steps, "max L1", max(err.values()), flush=True)
    assert max(err.values()) < 1e-3
    assert all(n < 300 for n in steps.values())
