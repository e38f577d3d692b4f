"""The spectrum-change estimate behind the time-to-accuracy Lanczos run (ed_spectrum_change, host only).

It is compared with an independent numpy/scipy evaluation: the terminator's centre and half-width from scipy's
tridiagonal eigenvalues (closure.bandwidth), the continued fraction level by level in complex numpy arithmetic,
and the relative L1 with numpy.gradient weights (allq.spectral_l1) on the same strided grid.
"""

import numpy as np
import pytest

from exactdiag_amd.allq import spectral_l1
from exactdiag_amd.closure import sqrt_terminator


def _cf(a, b, ws, eta, tail):
    z = ws + 1j * eta
    n = len(a)
    ai, bi = tail
    u = z - ai
    s = np.sqrt(u * u - 4 * bi * bi)
    t = (u - s) / (2 * bi * bi)
    t = np.where(t.imag > 0, (u + s) / (2 * bi * bi), t)
    d = z - a[n - 1] - bi * bi * t
    for i in range(1, n):
        d = z - a[n - 1 - i] - b[n - i] ** 2 / d
    return (1.0 / d).imag


_SEQ = {}


def _coefficients(n, seed):
    """A Lanczos sequence of the 2x6 ladder with 2 holes in the k = (1, 0) sector (t = 1, J = 0.4) from a seeded
    start vector, run by the oracle (a physical sequence: its spectra converge slowly with n)."""
    if seed not in _SEQ:
        from oracle import oracle

        o = oracle.OracleSector(6, 2, 5, 1, 0)
        rng = np.random.default_rng(seed)
        v0 = rng.normal(size=o.dim) + 1j * rng.normal(size=o.dim)
        _SEQ[seed] = o.lanczos([1.0, 1.0, 0.4, 0.4], v0, 160)
    a, b = _SEQ[seed]
    return a[:n], b[:n]


@pytest.mark.parametrize("n,delta,max_points", [(80, 20, 1024), (61, 20, 2001), (150, 10, 256)])
def test_spectrum_change_matches_numpy(n, delta, max_points):
    from exactdiag_amd.sector import StopRule

    a, b = _coefficients(n, 3)
    assert len(a) == n
    ws = np.linspace(0.0, 4.0, 2001)
    e0, eta = float(a.min()) - 1.0, 0.05
    rule = StopRule(ws, e0, eta, 1e-3, delta=delta, max_points=max_points)
    sub = ws[:: rule.c.stride] + e0
    spectra = []
    for m in (n, n - delta, n - delta // 2):
        at, bt, tail = sqrt_terminator(a, b, m)
        spectra.append(_cf(at, bt, sub, eta, tail))
    expect = max(spectral_l1(spectra[0], spectra[1], sub), spectral_l1(spectra[0], spectra[2], sub))
    got = rule.change(a, b)
    assert expect > 1e-6  # a change large enough that a wrong term would show
    assert abs(got - expect) <= 1e-10 * expect


def test_spectrum_change_shrinks_for_a_constant_chain():
    """Constant coefficients: the n-level tridiagonal's extreme eigenvalues are 0.3 +- 2.2 cos(pi / (n + 1)),
    so the terminator approaches the exact continuation and the change over 20 levels falls roughly as 1/n^3."""
    from exactdiag_amd.sector import StopRule

    rule = StopRule(np.linspace(-3.0, 3.0, 801), 0.0, 0.05, 1e-3)
    changes = []
    for n in (60, 120, 240):
        a = np.full(n, 0.3)
        b = np.concatenate([[0.0], np.full(n - 1, 1.1)])
        changes.append(rule.change(a, b))
    assert changes[0] > 4 * changes[1] > 16 * changes[2] > 0


def test_stop_rule_rejects_bad_arguments():
    from exactdiag_amd.sector import StopRule

    a, b = _coefficients(40, 1)
    ws = np.linspace(0.0, 4.0, 101)
    with pytest.raises(ValueError):
        StopRule(ws, 0.0, 0.05, 1e-3, delta=20).change(a, b, 20)  # n must exceed delta
    bad = StopRule(ws, 0.0, 0.05, 1e-3)
    bad.c.stride = 0
    with pytest.raises(ValueError):
        bad.change(a, b)
    bad = StopRule(ws, 0.0, -0.05, 1e-3)
    with pytest.raises(ValueError):
        bad.change(a, b)
