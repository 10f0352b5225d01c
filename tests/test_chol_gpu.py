"""?potrf_, ?potrs_ and ?posv_ on the GPU through the C-ABI, against the
oracle's xPOTRF2 restatement and the criteria of hwyBLAS's tests:
||L L^T - A||_1 / (n ||A||_1 eps) < 30 (tests/lapack/posv_test.cc:302-347,
lapack_test_util.h:243-249) and the solve ratio
||B - A X||_1 / (||A||_1 ||X||_1 n eps) < 30; the other triangle and the
padding keep their sentinels; info names the first failing minor."""
import numpy as np
import pytest

import oracle
from helpers import eps

pytestmark = pytest.mark.gpu
SENTINEL = 7.25


def spd(n, seed, dtype):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, n))
    return (m @ m.T / n + np.eye(n)).astype(dtype)


def ratio(prod, a, n, dtype):
    a = a.astype(np.float64)
    return np.abs(prod - a).sum(0).max() / (max(1, n) * np.abs(a).sum(0).max() * eps(dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf_residual_and_oracle(H, dtype, uplo):
    for n in (1, 5, 64, 65, 130, 300, 1000):
        a = spd(n, n, dtype)
        lda = n + 2
        buf = np.full((lda, n), SENTINEL, dtype=dtype, order="F")
        tri = np.tril(np.ones((n, n), dtype=bool)) if uplo == "L" else \
            np.triu(np.ones((n, n), dtype=bool))
        buf[:n, :n][tri] = a[tri]
        before = buf.copy()
        info = H.potrf(uplo, n, buf, lda)
        assert info == 0
        assert np.all(buf[n:, :] == SENTINEL) and np.array_equal(buf[:n, :n][~tri],
                                                                  before[:n, :n][~tri])
        f = np.where(tri, buf[:n, :n], 0).astype(np.float64)
        prod = f @ f.T if uplo == "L" else f.T @ f
        assert ratio(prod, a, n, dtype) < 30, n
        if dtype == np.float64:
            want, oinfo = oracle.potrf2(uplo, a)
            assert oinfo == 0
            w = np.where(tri, want, 0)
            assert np.max(np.abs(f - w)) <= 1e-12 * max(1.0, np.max(np.abs(w))), n


def test_potrf_exact_on_unit_triangular_products(H):
    """A = L L^T with an integer unit lower L: every pivot is 1, the leaves
    divide by 1, the TRSMs multiply by reciprocals of 1 and the updates are
    integer sums, so the factor is L itself."""
    rng = np.random.default_rng(4)
    for n in (7, 64, 200):
        low = np.tril(rng.integers(-2, 3, (n, n)).astype(float), -1) + np.eye(n)
        a = np.asfortranarray(low @ low.T)
        info = H.potrf("L", n, a, n)
        assert info == 0 and np.array_equal(np.tril(a), low), n


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf_info_first_failing_minor(H, uplo):
    """Minor 70 (past the first 64-column leaf) is not positive: info = 70;
    a NaN on the diagonal at 3: info = 3."""
    n = 150
    a = np.asfortranarray(np.diag(np.arange(1.0, n + 1)))
    a[69, 69] = -5.0
    assert H.potrf(uplo, n, a, n) == 70
    b = np.asfortranarray(np.eye(10))
    b[2, 2] = np.nan
    assert H.potrf(uplo, 10, b, 10) == 3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_posv_and_potrs(H, dtype, uplo):
    n, nrhs = 517, 9
    a = spd(n, 7, dtype)
    b = np.random.default_rng(8).uniform(-1, 1, (n, nrhs)).astype(dtype)
    af = np.asfortranarray(a.copy())
    x = np.asfortranarray(b.copy())
    assert H.posv(uplo, n, nrhs, af, n, x, n) == 0
    r = np.abs(b.astype(np.float64) - a.astype(np.float64) @ x.astype(np.float64)).sum(0).max()
    den = np.abs(a.astype(np.float64)).sum(0).max() * np.abs(x.astype(np.float64)).sum(0).max()
    assert r / (den * n * eps(dtype)) < 30
    # potrs with the factor posv left gives the same bits.
    x2 = np.asfortranarray(b.copy())
    assert H.potrs(uplo, n, nrhs, af, n, x2, n) == 0
    assert np.array_equal(x, x2)
    # posv on a matrix that is not positive definite: info, B untouched.
    c = np.asfortranarray(np.diag(np.r_[np.ones(99), -1.0, np.ones(n - 100)]).astype(dtype))
    y = np.asfortranarray(b.copy())
    assert H.posv(uplo, n, nrhs, c, n, y, n) == 100
    assert np.array_equal(y, b)


def test_device_operands(H):
    n = 300
    a = spd(n, 9, np.float64)
    da = H.DeviceArray.from_host(np.asfortranarray(a))
    assert H.potrf("L", n, da, n, dtype=np.float64) == 0
    host = np.asfortranarray(a.copy())
    assert H.potrf("L", n, host, n) == 0
    assert np.array_equal(da.to_host(), host)
