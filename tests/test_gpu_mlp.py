"""BF16 MFMA kernels of the closure MLP against the CPU oracle."""

import ctypes as ct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    assert t.cuda.is_available()
    return t


def _bits(a):
    from oracle import mlp_oracle as mo

    return mo.bf16(a).view(np.uint32).astype(np.uint64) >> 16


@pytest.mark.parametrize("shape", [(32, 32, 16), (64, 96, 48), (37, 45, 29), (128, 384, 256), (5, 7, 300)])
def test_gemm_lane_layout_and_values(torch, shape):
    from exactdiag_amd._lib import check, lib
    from oracle import mlp_oracle as mo

    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    ref = mo.bf16(A).astype(np.float64) @ mo.bf16(B).astype(np.float64)
    dA = torch.as_tensor(_bits(A).astype(np.uint16).view(np.int16), device="cuda")
    dB = torch.as_tensor(_bits(B).astype(np.uint16).view(np.int16), device="cuda")
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    s = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().ed_gemm_bf16(M, N, K, ct.c_void_p(dA.data_ptr()), K, 1, ct.c_void_p(dB.data_ptr()), N, 1,
                             ct.c_void_p(C.data_ptr()), N, s))
    got = C.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 1e-4 * max(1.0, np.max(np.abs(ref)))
    # transposed operand through strides: C^T = B^T A^T
    Ct = torch.zeros((N, M), dtype=torch.float32, device="cuda")
    check(lib().ed_gemm_bf16(N, M, K, ct.c_void_p(dB.data_ptr()), 1, N, ct.c_void_p(dA.data_ptr()), 1, K,
                             ct.c_void_p(Ct.data_ptr()), M, s))
    assert np.max(np.abs(Ct.cpu().numpy() - ref.T)) < 1e-4 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("shape", [(200, 136, 88), (264, 512, 1024), (128, 128, 32), (8, 520, 40), (2104, 2056, 72),
                                   (4104, 4104, 40)])
@pytest.mark.parametrize("a_kcontig,b_kcontig", [(True, False), (True, True), (False, False), (False, True)])
def test_tiled_gemm_all_operand_layouts(torch, shape, a_kcontig, b_kcontig):
    """The LDS-tiled kernel (taken when a stride is 1 and the other a multiple of 8) for every layout of A and
    B, with tiles that cross the matrix edges, against float64 products of the same BF16 values; the last two
    shapes have C large enough for the 128 x 256 and the 256 x 256 tiles."""
    from exactdiag_amd._lib import check, lib
    from oracle import mlp_oracle as mo

    M, N, K = shape
    rng = np.random.default_rng(M * N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    ref = mo.bf16(A).astype(np.float64) @ mo.bf16(B).astype(np.float64)
    # A stored as [m][k] (k contiguous) or as its transpose [k][m]; B as [n][k] (k contiguous) or [k][n]
    Ast = _bits(A) if a_kcontig else _bits(np.ascontiguousarray(A.T))
    Bst = _bits(np.ascontiguousarray(B.T)) if b_kcontig else _bits(B)
    sam, sak = (K, 1) if a_kcontig else (1, M)
    sbk, sbn = (1, K) if b_kcontig else (N, 1)
    dA = torch.as_tensor(Ast.astype(np.uint16).view(np.int16), device="cuda")
    dB = torch.as_tensor(Bst.astype(np.uint16).view(np.int16), device="cuda")
    C = torch.full((M, N), float("nan"), dtype=torch.float32, device="cuda")
    s = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().ed_gemm_bf16(M, N, K, ct.c_void_p(dA.data_ptr()), sam, sak, ct.c_void_p(dB.data_ptr()), sbk, sbn,
                             ct.c_void_p(C.data_ptr()), N, s))
    got = C.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) < 1e-4 * max(1.0, np.max(np.abs(ref)))


def test_forward_and_gradients_match_oracle(torch):
    from exactdiag_amd.closure import GpuMLP
    from oracle import mlp_oracle as mo

    dims = [128, 256, 256, 256, 384]
    net = GpuMLP(dims, seed=0)
    p = mo.init_params(dims, seed=5)
    net.set_params(p)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(96, 128)).astype(np.float32)
    t = rng.normal(size=(96, 384)).astype(np.float32)
    y = net.forward(x)
    zs, _ = mo.forward(p, dims, x)
    assert np.max(np.abs(y - zs[-1])) < 2e-2 * max(1.0, np.max(np.abs(zs[-1])))
    loss, g = net.gradients(x, t)
    lref, gref = mo.loss_and_grad(p, dims, x, t)
    assert abs(loss - lref) < 1e-3 * lref
    off = 0
    for l in range(1, len(dims)):
        for size in (dims[l - 1] * dims[l], dims[l]):
            a, b = g[off:off + size], gref[off:off + size]
            assert np.linalg.norm(a - b) < 2e-2 * np.linalg.norm(b) + 1e-8, (l, size)
            off += size


def test_training_reduces_loss(torch):
    from exactdiag_amd.closure import GpuMLP

    rng = np.random.default_rng(8)
    dims = [16, 64, 64, 8]
    x = rng.normal(size=(256, 16)).astype(np.float32)
    W = rng.normal(size=(16, 8)).astype(np.float32) / 4
    t = np.tanh(x @ W).astype(np.float32)
    net = GpuMLP(dims, seed=1)
    first = net.train_step(x, t, 3e-3)
    for _ in range(300):
        last = net.train_step(x, t, 3e-3)
    assert last < 0.1 * first
