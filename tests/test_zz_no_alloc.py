"""ed_vec_dots_noalloc, the entry point that allocates no device memory, against ed_vec_dots.

Both run the same kernels; the no-alloc one keeps its partials in the library's fixed per-GPU array.  The
test calls both through the C ABI twice in a row, the second call with fewer vectors and a shorter vector
than the first, and needs bit-equal results each time (and agreement with numpy).
"""

import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _dots(torch, fn, V, nv, w):
    out = np.full(2 * max(nv, 1), np.nan)
    stream = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(ct.c_void_p(V.data_ptr()), nv, V.stride(0), ct.c_void_p(w.data_ptr()), w.numel(),
            out.ctypes.data_as(ct.POINTER(ct.c_double)), stream)
    return rc, out.view(np.complex128)[:nv].copy()


@pytest.mark.parametrize("first,second", [((40, 300_001, 300_017), (7, 5_000, 5_000)),
                                          ((128, 70_000, 70_000), (1, 1, 1))])
def test_noalloc_dots_equal_allocating_dots_on_two_calls(torch, first, second):
    from exactdiag_amd import _lib

    lib = _lib.lib()
    rng = np.random.default_rng(11)
    for nv, D, ld in (first, second):
        print("noalloc dots: nv", nv, "D", D, "ld", ld, flush=True)
        Vh = rng.normal(size=(nv, ld)) + 1j * rng.normal(size=(nv, ld))
        wh = rng.normal(size=D) + 1j * rng.normal(size=D)
        V = torch.as_tensor(Vh, device="cuda")
        w = torch.as_tensor(wh, device="cuda")
        rc0, ref = _dots(torch, lib.ed_vec_dots, V, nv, w)
        rc1, got = _dots(torch, lib.ed_vec_dots_noalloc, V, nv, w)
        assert rc0 == 0 and rc1 == 0, _lib.lib().ed_last_error()
        assert np.array_equal(got.view(np.float64), ref.view(np.float64)), (nv, D)
        expect = Vh[:, :D].conj() @ wh
        assert np.max(np.abs(got - expect)) < 1e-10 * max(1.0, np.max(np.abs(expect)))


def test_noalloc_dots_limits(torch):
    from exactdiag_amd import _lib

    lib = _lib.lib()
    V = torch.zeros((129, 16), dtype=torch.complex128, device="cuda")
    w = torch.zeros(16, dtype=torch.complex128, device="cuda")
    rc, _ = _dots(torch, lib.ed_vec_dots_noalloc, V, 129, w)
    assert rc == _lib.ED_ERR_CAPACITY
    assert b"at most 128 vectors" in lib.ed_last_error()
    # an empty vector gives zero dots, as ed_vec_dots does
    e = torch.zeros(0, dtype=torch.complex128, device="cuda")
    rc, got = _dots(torch, lib.ed_vec_dots_noalloc, V[:3, :0], 3, e)
    rc0, ref = _dots(torch, lib.ed_vec_dots, V[:3, :0], 3, e)
    assert rc == 0 and rc0 == 0 and np.array_equal(got, ref) and np.all(got == 0)
