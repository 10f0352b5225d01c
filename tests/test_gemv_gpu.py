"""?gemv_ and cblas_?gemv on the GPU through the C-ABI (hwyBLAS src/gemv.cc
semantics): every result within the componentwise bound 4 (k + 2) eps of
the double-double reference (the GEMM reference with one column), for both
transpositions, increments 1, 2 and -3, beta in {0, 1, -0.5} and alpha in
{1.5, 0}; y between its increments, and y itself when beta = 0 is NaN, must
not leak into the result; host and device operands give the same bits."""
import numpy as np
import pytest

import oracle
from helpers import eps

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32]


def vec(len_, inc, seed, dtype, fill=7.25):
    """A BLAS vector of length len_ and increment inc in a buffer of
    1 + (len_ - 1) |inc| elements (the gaps hold `fill`), and its logical
    values."""
    buf = np.full(max(1, 1 + (len_ - 1) * abs(inc)), fill, dtype=dtype)
    vals = oracle.fill_uniform(len_, 1, seed, dtype=dtype).ravel()
    idx = np.arange(len_) * inc if inc > 0 else (len_ - 1 - np.arange(len_)) * (-inc)
    buf[idx] = vals
    return buf, vals, idx


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("trans", ["N", "T"])
def test_gemv_bound(H, dtype, trans):
    for (m, n) in ((300, 257), (1, 700), (700, 1), (2049, 513)):
        lda = m + 3
        a = oracle.fill_uniform(m, n, 5, dtype=dtype, ld=lda)
        len_x, len_y = (n, m) if trans == "N" else (m, n)
        for incx, incy in ((1, 1), (2, -3), (-1, 2)):
            for alpha, beta in ((1.5, -0.5), (1.0, 0.0), (1.0, 1.0), (0.0, -0.5)):
                x, xv, _ = vec(len_x, incx, 6, dtype)
                y, yv, yi = vec(len_y, incy, 7, dtype)
                if beta == 0.0:
                    y[yi] = np.nan  # beta = 0: y is not read
                y0 = y.copy()
                H.gemv(trans, m, n, alpha, a, lda, x, incx, beta, y, incy)
                gaps = np.ones(y.size, dtype=bool)
                gaps[yi] = False
                assert np.array_equal(y[gaps], y0[gaps])
                got = y[yi].reshape(len_y, 1, order="F")
                ym = yv.reshape(len_y, 1, order="F").astype(dtype)
                if beta == 0.0:
                    ym = np.zeros_like(ym)
                exp, bnd = oracle.gemm_ref(trans, "N", len_y, 1, len_x, alpha, a, lda,
                                           np.asfortranarray(xv.reshape(len_x, 1)), len_x, beta,
                                           np.asfortranarray(ym), len_y)
                err = np.abs(got.astype(np.float64) - exp)
                tol = 4.0 * (len_x + 2) * eps(dtype)
                assert np.all(err <= tol * bnd), (m, n, incx, incy, alpha, beta)


@pytest.mark.parametrize("dtype", DTYPES)
def test_gemv_device_operands_and_cblas(H, dtype):
    m, n = 1000, 777
    a = oracle.fill_uniform(m, n, 8, dtype=dtype)
    x = oracle.fill_uniform(n, 1, 9, dtype=dtype).ravel()
    y = oracle.fill_uniform(m, 1, 10, dtype=dtype).ravel()
    want = y.copy()
    H.gemv("N", m, n, 2.0, a, m, x, 1, 0.5, want, 1)
    da, dx, dy = (H.DeviceArray.from_host(np.asfortranarray(v.reshape(v.shape[0], -1)))
                  for v in (a, x, y))
    H.gemv("N", m, n, 2.0, da, m, dx, 1, 0.5, dy, 1, dtype=dtype)
    assert np.array_equal(dy.to_host().ravel(), want)
    # cblas: column-major is the Fortran call; row-major A (m x n, C order) is
    # the column-major A^T, so row-major N equals Fortran T on the same buffer.
    got = y.copy()
    H.cblas_level3("gemv", H.ColMajor, H.NoTrans, m, n, 2.0, a, m, x, 1, 0.5, got, 1, dtype=dtype)
    assert np.array_equal(got, want)
    ar = np.ascontiguousarray(a)  # row-major m x n
    got = y.copy()
    H.cblas_level3("gemv", H.RowMajor, H.NoTrans, m, n, 2.0, ar, n, x, 1, 0.5, got, 1,
                   dtype=dtype)
    want_t = y.copy()
    H.gemv("T", n, m, 2.0, ar.ravel(), n, x, 1, 0.5, want_t, 1)
    assert np.array_equal(got, want_t)
    ref = 2.0 * a.astype(np.float64) @ x.astype(np.float64) + 0.5 * y.astype(np.float64)
    assert np.max(np.abs(got - ref)) <= 4 * (n + 2) * eps(dtype) * np.max(np.abs(a)) * n
