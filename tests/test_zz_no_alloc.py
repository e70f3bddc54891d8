"""rsp_xterm_gemm_static (the split-BF16 cross-term GEMM with its scratch in a static device
array, for callers that cannot allocate) against rsp_xterm_gemm mode 1 (the same kernels
with scratch allocated per call): bit-identical outputs on two calls in a row, the second
smaller, and a clear size error above the scratch's 32 MiB."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lib():
    from refspi_amd import _lib as L
    return L, L.load()


def _pair(L, lib, M, N, K, seed):
    import torch
    rng = np.random.default_rng(seed)
    a = torch.as_tensor(rng.standard_normal((M, K)).astype('f4'), device='cuda')
    b = torch.as_tensor(rng.standard_normal((N, K)).astype('f4'), device='cuda')
    c_alloc = torch.full((N, M), np.nan, dtype=torch.float32, device='cuda')
    c_static = torch.full((N, M), np.nan, dtype=torch.float32, device='cuda')
    L.call('rsp_xterm_gemm', 1, M, N, K, L.ptr(a), L.ptr(b), L.ptr(c_alloc), L.stream_handle())
    print('rsp_xterm_gemm_static M=%d N=%d K=%d' % (M, N, K), flush=True)
    L.call('rsp_xterm_gemm_static', M, N, K, L.ptr(a), L.ptr(b), L.ptr(c_static),
           L.stream_handle())
    torch.cuda.synchronize()
    return c_alloc.cpu().numpy(), c_static.cpu().numpy(), a.cpu().numpy(), b.cpu().numpy()


def test_static_scratch_gemm_equals_allocating_gemm():
    L, lib = _lib()
    # second call smaller: the scratch keeps the first call's contents
    for (M, N, K, seed) in ((720, 300, 5000, 1), (100, 37, 700, 2)):
        ca, cs, a, b = _pair(L, lib, M, N, K, seed)
        assert np.isfinite(cs).all()
        np.testing.assert_array_equal(cs, ca)
        ref = b.astype('f8') @ a.astype('f8').T
        bound = (3 * 2. ** -18 + K * 2. ** -24) * (np.abs(b).astype('f8') @ np.abs(a).astype('f8').T)
        assert np.all(np.abs(cs - ref) <= bound)


def test_static_scratch_gemm_rejects_sizes_above_its_scratch():
    L, lib = _lib()
    # planes (2*768 + 2*4096) * 4096 * 2 bytes > 32 MiB: refused before any device call
    rc = lib.rsp_xterm_gemm_static(720, 4000, 4096, 1, 1, 1, None)
    assert rc == -2
    assert b'scratch' in lib.rsp_last_error()
