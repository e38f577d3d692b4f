"""Device blocks kept for reuse after a free: sectors and stored matrices rebuilt on reused blocks give the same
bases and products bit for bit, before and after the blocks are released (ed_cache_release)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = [1.0, 1.0, 0.4, 0.4]


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _run(torch, L, kx):
    from exactdiag_amd.sector import Sector

    s = Sector(L, 2, L - 1, kx, 0)
    hm = s.store(W)
    rng = np.random.default_rng(L * 10 + kx)
    x = torch.as_tensor(rng.normal(size=s.dim) + 1j * rng.normal(size=s.dim), device="cuda")
    y, dot = hm.hv(x, with_dot=True)
    ym = s.hv(W, x)
    key, orbit = s.basis()
    out = (key, orbit, y.cpu().numpy(), dot, ym.cpu().numpy())
    hm.close()
    s.close()
    return out


def test_rebuilt_on_reused_blocks(torch):
    from exactdiag_amd import _lib

    cases = [(10, 1), (8, 0), (10, 3), (6, 2)]
    first = {c: _run(torch, *c) for c in cases}
    for rnd in range(2):
        for c in reversed(cases):  # other sizes freed in between: blocks are taken by requests of other shapes
            again = _run(torch, *c)
            print("round", rnd, "case", c, "dim", len(again[0]), flush=True)
            for u, v in zip(first[c], again):
                assert np.array_equal(np.asarray(u), np.asarray(v))
        _lib.release_cached_memory()
