"""Pin the closure-MLP oracle: FP64 gradients against central finite differences,
BF16 rounding against its definition, and the BF16 variant close to FP64."""

import numpy as np

from oracle import mlp_oracle as mo

DIMS = [12, 16, 16, 10]


def test_bf16_rounding_definition():
    x = np.array([1.0, 1.00390625, 1.0078125, 3.14159265, -2.5e-3, 65504.0, 1e-40], np.float32)
    r = mo.bf16(x)
    # bf16 keeps 8 significant bits: representable values stay, others round to nearest (ties to even)
    assert r[0] == 1.0 and r[2] == 1.0078125
    assert r[1] == 1.0  # exactly halfway between 1.0 and 1.0078125 -> even mantissa
    assert abs(r[3] - 3.140625) < 1e-7
    # relative half-ulp bound for normal numbers, absolute half-spacing (2^-134) for subnormals
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0**-8 + 2.0**-134)


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mo.init_params(DIMS, seed=1).astype(np.float64)
    x = rng.normal(size=(7, DIMS[0]))
    t = rng.normal(size=(7, DIMS[-1]))
    _, g = mo.loss_and_grad(p, DIMS, x, t, exact=True)
    idx = rng.choice(p.size, size=60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        lp, _ = mo.loss_and_grad(p + e, DIMS, x, t, exact=True)
        lm, _ = mo.loss_and_grad(p - e, DIMS, x, t, exact=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_variant_close_to_fp64():
    rng = np.random.default_rng(2)
    p = mo.init_params(DIMS, seed=3)
    x = rng.normal(size=(32, DIMS[0])).astype(np.float32)
    t = rng.normal(size=(32, DIMS[-1])).astype(np.float32)
    l16, g16 = mo.loss_and_grad(p, DIMS, x, t)
    l64, g64 = mo.loss_and_grad(p.astype(np.float64), DIMS, x, t, exact=True)
    assert abs(l16 - l64) < 2e-2 * l64
    assert np.linalg.norm(g16 - g64) < 3e-2 * np.linalg.norm(g64)
