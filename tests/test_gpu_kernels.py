"""Device kernels against LAPACK-grade references through the C ABI:
the MFMA GEMM (FP64 and FP32), the batched blocked Cholesky and the
Householder tridiagonalisation + bisection eigenvalues."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _e():
    from mbd_amd import engine
    return engine


# n <= 32 with k <= 128 takes the tall-skinny kernel (64-row workgroups, K in
# one LDS piece), the rest the 128 x 128 tiles
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (128, 128, 16), (130, 70, 33),
                                   (300, 257, 129), (1000, 32, 128), (777, 17, 33), (65, 32, 1),
                                   (300, 20, 200)])
def test_gemm_nt_fp64(m, n, k):
    rng = np.random.default_rng(m * 7 + n)
    A = rng.standard_normal((m, k))
    B = rng.standard_normal((n, k))
    C = rng.standard_normal((m, n))
    got = _e().device_gemm_nt(A, B, C, alpha=-0.5, beta_=2.0)
    want = -0.5 * A @ B.T + 2.0 * C
    assert np.allclose(got, want, rtol=1e-13, atol=1e-12 * np.sqrt(k))


def test_gemm_nt_fp32():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((200, 64)).astype(np.float32).astype(np.float64)
    B = rng.standard_normal((150, 64)).astype(np.float32).astype(np.float64)
    C = np.zeros((200, 150))
    got = _e().device_gemm_nt(A, B, C, fp32=True)
    assert np.allclose(got, A @ B.T, rtol=0, atol=2e-5 * 8)


@pytest.mark.parametrize("n,batch", [(1, 1), (3, 2), (36, 1), (127, 1), (128, 3), (129, 2),
                                     (300, 2), (517, 1)])
def test_cholesky_batched(n, batch):
    rng = np.random.default_rng(n)
    As = []
    for _ in range(batch):
        M = rng.standard_normal((n, n))
        As.append(M @ M.T / n + np.eye(n))
    A = np.array(As)
    L, info = _e().device_cholesky(A)
    assert list(info) == [0] * batch
    for b in range(batch):
        assert np.allclose(L[b], np.linalg.cholesky(A[b]), rtol=1e-12, atol=1e-12)


def test_cholesky_reports_first_non_positive_pivot():
    A = np.diag([4.0] * 130 + [-1.0] + [2.0] * 9)
    _, info = _e().device_cholesky(A)
    assert info[0] == 131


@pytest.mark.parametrize("n", [1, 2, 5, 100, 128, 129, 200, 257, 517])
def test_sym_eigvals(n):
    rng = np.random.default_rng(100 + n)
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2 + np.diag(rng.uniform(1, 3, n))
    w = _e().device_sym_eigvals(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), rtol=0, atol=1e-12 * np.abs(A).max() * n ** 0.5)
