"""The read-rate probe behind the stored product's roofline (ed_hmat_line_probe): argument checks on the host,
and on the GPU that both access patterns give a rate inside what the device can serve."""

import ctypes as ct

import pytest

from exactdiag_amd._lib import ED_ERR_ARG, lib


@pytest.mark.parametrize("span", [0, 1 << 19, (1 << 27) + 4096, 1 << 33])
def test_probe_rejects_bad_spans(span):
    r = ct.c_double(-1.0)
    assert lib().ed_hmat_line_probe(1, span, ct.byref(r), None) == ED_ERR_ARG
    assert r.value == -1.0


def test_probe_rejects_null_output():
    assert lib().ed_hmat_line_probe(1, 1 << 27, None, None) == ED_ERR_ARG


@pytest.mark.gpu
def test_probe_rates_are_plausible():
    import torch

    from exactdiag_amd._lib import check

    sp = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    rates = {}
    for scattered in (1, 0):
        r = ct.c_double(0.0)
        check(lib().ed_hmat_line_probe(scattered, 1 << 27, ct.byref(r), sp))
        rates[scattered] = r.value
    for v in rates.values():
        # at least 1 TB/s of 128-byte lines, at most the 8 TB/s HBM peak plus what the caches can add
        assert 1e12 / 128 < v < 64e12 / 128
