"""GPU parity of the cross-term GEMM (rsp_xterm_gemm) against an FP64 product.

Split-BF16 bound: with hi = bf16(x) (|x - hi| <= 2^-9 |x|) and lo = bf16(x - hi), the dropped
part of each product is lo_a*lo_b plus the rounding of lo, at most 3 * 2^-18 |a b|; FP32
accumulation over K terms adds at most K * 2^-24 sum |a b|.  The test asserts
|C - C64| <= (3 * 2^-18 + K * 2^-24) * (|A| @ |B|^T) elementwise.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(mode, A, B):
    import torch
    from refspi_amd import _lib
    dt = torch.float64 if mode == 2 else torch.float32
    a = torch.as_tensor(A, device='cuda').to(dt).contiguous()
    b = torch.as_tensor(B, device='cuda').to(dt).contiguous()
    c = torch.empty(B.shape[0], A.shape[0], dtype=dt, device='cuda')
    _lib.call('rsp_xterm_gemm', mode, A.shape[0], B.shape[0], A.shape[1], _lib.ptr(a),
              _lib.ptr(b), _lib.ptr(c), _lib.stream_handle())
    torch.cuda.synchronize()
    return c.cpu().numpy().astype('f8')


@pytest.mark.parametrize('M,N,K', [(300, 200, 1000), (720, 390, 4096), (40, 1, 70),
                                   (256, 128, 32)])
def test_split_bf16_within_bound(M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype('f4').astype('f8')
    B = rng.standard_normal((N, K)).astype('f4').astype('f8')
    C64 = B @ A.T
    bound = (3 * 2. ** -18 + K * 2. ** -24) * (np.abs(B) @ np.abs(A).T)
    Cs = _run(1, A, B)
    assert np.all(np.abs(Cs - C64) <= bound)
    C32 = _run(0, A, B)
    assert np.all(np.abs(C32 - C64) <= K * 2. ** -24 * (np.abs(B) @ np.abs(A).T) + 1e-30)
    C2 = _run(2, A, B)
    np.testing.assert_allclose(C2, C64, rtol=1e-12, atol=1e-12 * np.abs(C64).max())


def test_split_bf16_exact_on_bf16_inputs():
    """Inputs that are exactly bf16 have lo = 0: every product is exact in FP32, so the
    result equals an FP32-accumulated sum within K * 2^-24 and is exact for small ints."""
    rng = np.random.default_rng(9)
    A = rng.integers(-8, 9, (64, 96)).astype('f8')
    B = rng.integers(-8, 9, (33, 96)).astype('f8')
    np.testing.assert_array_equal(_run(1, A, B), B @ A.T)


def test_split_bf16_layout_rows_and_columns():
    """One-hot rows/columns place each output at C[d, r] (checks the MFMA output layout)."""
    M, N, K = 260, 140, 64
    A = np.zeros((M, K))
    B = np.zeros((N, K))
    A[np.arange(M), np.arange(M) % K] = np.arange(1, M + 1)
    B[np.arange(N), np.arange(N) % K] = 1.
    C = _run(1, A, B)
    ref = B @ A.T
    np.testing.assert_array_equal(C, ref)
