"""vd_dot_f64_noalloc (the allocation-free dot product) against vd_dot_f64,
through the C ABI: the same kernels, so the results must be identical bit
for bit; two calls in a row, the second with a smaller input (the shared
device scratch keeps the first call's partials, which the second must not
see), then the size check."""
import ctypes as C

import numpy as np
import pytest
import torch

from vardyn_amd import _lib

pytestmark = pytest.mark.gpu


def _dot(lib, n, x, y, alpha, out, scratch=None):
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if scratch is None:
        return lib.vd_dot_f64_noalloc(n, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                      alpha, C.c_void_p(out.data_ptr()), s)
    return lib.vd_dot_f64(n, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), alpha,
                          C.c_void_p(out.data_ptr()), C.c_void_p(scratch.data_ptr()), s)


def test_noalloc_dot_matches_dot_on_two_calls():
    lib = _lib.load()
    dev = torch.device("cuda")
    g = torch.Generator(device=dev).manual_seed(3)
    scratch = torch.empty(_lib.VD_REDUCE_SCRATCH, dtype=torch.float64, device=dev)
    for n in (3_000_001, 777):  # second call smaller: fewer partials than the first
        x = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
        y = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
        a = torch.full((1,), 0.25, dtype=torch.float64, device=dev)
        b = a.clone()
        assert _dot(lib, n, x, y, 0.5, a) == 0, _lib.last_error()
        assert _dot(lib, n, x, y, 0.5, b, scratch) == 0, _lib.last_error()
        torch.cuda.synchronize()
        assert torch.equal(a, b), (n, a.item(), b.item())
        ref = 0.25 + 0.5 * float(np.dot(x.cpu().numpy(), y.cpu().numpy()))
        assert a.item() == pytest.approx(ref, rel=1e-12, abs=1e-12)


def test_noalloc_dot_zero_length_and_size_check():
    lib = _lib.load()
    x = torch.ones(4, dtype=torch.float64, device="cuda")
    out = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
    assert _dot(lib, 0, x, x, 1.0, out) == 0
    torch.cuda.synchronize()
    assert out.item() == 7.0
    assert _dot(lib, (1 << 36) + 1, x, x, 1.0, out) == _lib.VD_OK - 1
    assert b"VD_DOT_NOALLOC_MAX_N" in lib.vd_last_error()
    assert _dot(lib, -1, x, x, 1.0, out) == -1
