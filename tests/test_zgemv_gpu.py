"""Complex GEMV (cgemv_, zgemv_, cblas_?gemv) on the GPU, checked against the
complex128 product written out with the componentwise bound 4 (k + 2) eps
(|alpha| |op(A)| |x| + |beta| |y|), k the dot length: every op (N, T, C),
increments 1, 2, -3, beta in {0 (y poisoned), 1, complex}, device operands
bitwise equal to host ones, and row-major CBLAS in every op against the
column-major routine on the same logical matrix."""
import numpy as np
import pytest

import hwyblas_amd as H
from test_zgemm_gpu import zeps, zrand

pytestmark = pytest.mark.gpu
ZDTYPES = [np.complex128, np.complex64]


def vec(v, length, inc):
    """Logical elements of a BLAS vector with increment inc."""
    idx = np.arange(length) * inc if inc > 0 else (length - 1 - np.arange(length)) * (-inc)
    return v[idx], idx


def reference(trans, a, x, alpha, beta, y0, m, n):
    aa = a[:m, :n].astype(np.complex128)
    op = aa if trans == "N" else (aa.T if trans == "T" else np.conj(aa.T))
    xx = x.astype(np.complex128)
    exp = alpha * (op @ xx)
    scale = abs(alpha) * (np.abs(op) @ np.abs(xx))
    if beta != 0:
        exp = exp + beta * y0.astype(np.complex128)
        scale = scale + abs(beta) * np.abs(y0)
    return exp, scale + np.finfo(np.float64).tiny


def check(got, exp, scale, k, dtype):
    err = np.abs(got.astype(np.complex128) - exp)
    err[np.isnan(err)] = np.inf
    r = float(np.max(err / (4.0 * (k + 2) * zeps(dtype) * scale)))
    assert r <= 1.0, r


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -3), (-1, 2)])
@pytest.mark.parametrize("beta", [0.0, 1.0, 0.5 - 2.0j])
def test_zgemv(dtype, trans, incx, incy, beta):
    m, n = 300, 517
    a = zrand(m + 3, n, 1, dtype)
    lx, ly = (n, m) if trans == "N" else (m, n)
    x = zrand(lx * abs(incx), 1, 2, dtype).ravel()
    y = zrand(ly * abs(incy), 1, 3, dtype).ravel()
    xs, _ = vec(x, lx, incx)
    y0, yidx = vec(y, ly, incy)
    if beta == 0:
        y[yidx] = np.nan
    alpha = 1.5 + 0.75j
    yk = y.copy()
    H.zgemv(trans, m, n, alpha, a, m + 3, x, incx, beta, y, incy)
    exp, scale = reference(trans, a, xs, alpha, beta, y0, m, n)
    check(y[yidx], exp, scale, lx, dtype)
    untouched = np.ones(len(y), bool)
    untouched[yidx] = False
    assert np.array_equal(y[untouched], yk[untouched])


@pytest.mark.parametrize("dtype", ZDTYPES)
def test_zgemv_device_operands(dtype):
    m, n = 129, 77
    a = zrand(m, n, 4, dtype)
    x = zrand(m, 1, 5, dtype)
    y = zrand(n, 1, 6, dtype)
    host = y.copy(order="F")
    H.zgemv("C", m, n, 2.0 - 1.0j, a, m, x, 1, 0.25 + 0.5j, host, 1)
    da, dx, dy = H.DeviceArray.from_host(a), H.DeviceArray.from_host(x), H.DeviceArray.from_host(y)
    H.zgemv("C", m, n, 2.0 - 1.0j, da, m, dx, 1, 0.25 + 0.5j, dy, 1, dtype=dtype)
    assert np.array_equal(dy.to_host(), host)


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("trans", [H.NoTrans, H.Trans, H.ConjTrans])
def test_cblas_zgemv_row_major(dtype, trans):
    m, n = 90, 140
    a = zrand(m, n, 7, dtype)
    a_row = np.ascontiguousarray(a)
    tch = {H.NoTrans: "N", H.Trans: "T", H.ConjTrans: "C"}[trans]
    lx, ly = (n, m) if tch == "N" else (m, n)
    x = zrand(lx, 1, 8, dtype).ravel()
    y = zrand(ly, 1, 9, dtype).ravel()
    y0 = y.copy()
    alpha, beta = -0.5 + 1.0j, 1.5 + 0.0j
    H.cblas_zgemv(H.RowMajor, trans, m, n, alpha, a_row, n, x, 1, beta, y, 1, dtype=dtype)
    exp, scale = reference(tch, a, x, alpha, beta, y0, m, n)
    check(y, exp, scale, lx, dtype)
    yc = y0.copy()
    H.cblas_zgemv(H.ColMajor, trans, m, n, alpha, a, m, x, 1, beta, yc, 1, dtype=dtype)
    check(yc, exp, scale, lx, dtype)
