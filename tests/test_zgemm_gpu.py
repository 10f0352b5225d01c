"""Complex GEMM (cgemm_, zgemm_, cblas_?gemm, hwyblas_?gemm_dev) on the GPU,
checked as hwyBLAS checks it (tests/gemm_test.cc:47-120): against a
higher-precision reference with the bound 4 (k + 2) eps (|alpha| sum |a||b| +
|beta| |c|), beta = 0 with C poisoned by NaN, padding rows of C untouched.
The reference here is built from the oracle's double-double real GEMM: the
four real products of op(A) and op(B) written out, each to about one
rounding, combined in float64."""
import itertools

import numpy as np
import pytest

import hwyblas_amd as H
import oracle

pytestmark = pytest.mark.gpu
ZDTYPES = [np.complex128, np.complex64]
OPS = ["N", "T", "C"]


def zeps(dtype):
    return float(np.finfo(np.float64 if np.dtype(dtype) == np.complex128 else np.float32).eps)


def zrand(rows, cols, seed, dtype):
    rt = np.float64 if np.dtype(dtype) == np.complex128 else np.float32
    re = oracle.fill_uniform(rows, cols, seed, dtype=rt)
    im = oracle.fill_uniform(rows, cols, seed + 1000, dtype=rt)
    return np.asfortranarray(re + 1j * im).astype(dtype)


def op(x, t, rows, cols):
    """op(X) written out (rows x cols of the operation), complex128."""
    x = x.astype(np.complex128)
    if t == "N":
        return x[:rows, :cols]
    y = x[:cols, :rows].T
    return np.conj(y) if t == "C" else y


def real_product(x, y):
    """Double-double x @ y for real float64 x, y (compact, Fortran order)."""
    m, k = x.shape
    n = y.shape[1]
    xa, ya = np.asfortranarray(x), np.asfortranarray(y)
    exp, _ = oracle.gemm_ref("N", "N", m, n, k, 1.0, xa, m, ya, k, 0.0, None, m)
    return exp


def reference(ta, tb, m, n, k, alpha, a, b, beta, c0):
    x, y = op(a, ta, m, k), op(b, tb, k, n)
    if k == 0:
        s = np.zeros((m, n), dtype=np.complex128)
    else:
        rr = real_product(x.real, y.real)
        ii = real_product(x.imag, y.imag)
        ri = real_product(x.real, y.imag)
        ir = real_product(x.imag, y.real)
        s = (rr - ii) + 1j * (ir + ri)
    scale = abs(alpha) * (np.abs(x) @ np.abs(y))
    exp = alpha * s
    if beta != 0:
        cc = c0[:m, :n].astype(np.complex128)
        exp = exp + beta * cc
        scale = scale + abs(beta) * np.abs(cc)
    return exp, scale + np.finfo(np.float64).tiny


def check(got, exp, scale, k, dtype):
    tol = 4.0 * (k + 2) * zeps(dtype)
    err = np.abs(got.astype(np.complex128) - exp)
    err[np.isnan(err)] = np.inf
    r = float(np.max(err / (tol * scale)))
    assert r <= 1.0, r


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("ta,tb", list(itertools.product(OPS, OPS)))
def test_zgemm_every_op_pair(dtype, ta, tb):
    m, n, k, pad = 67, 45, 130, 3
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    a = zrand(ra + pad, ca, 1, dtype)
    b = zrand(rb + pad, cb, 2, dtype)
    c = zrand(m + pad, n, 3, dtype)
    c0 = c.copy()
    alpha, beta = 0.75 - 0.5j, -0.25 + 1.5j
    H.zgemm(ta, tb, m, n, k, alpha, a, ra + pad, b, rb + pad, beta, c, m + pad)
    exp, scale = reference(ta, tb, m, n, k, alpha, a, b, beta, c0)
    check(c[:m], exp, scale, k, dtype)
    assert np.array_equal(c[m:], c0[m:])


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("beta", [0.0, 1.0, 2.0 - 1.0j])
@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 17, 5), (128, 64, 256), (300, 257, 70)])
def test_zgemm_beta_cases_and_shapes(dtype, beta, shape):
    m, n, k = shape
    a = zrand(m, k, 4, dtype)
    b = zrand(n, k, 5, dtype)
    c = zrand(m, n, 6, dtype)
    if beta == 0:
        c[:] = np.nan  # beta = 0 must not read C
    c0 = c.copy()
    alpha = -1.25 + 0.5j
    H.zgemm("N", "C", m, n, k, alpha, a, m, b, n, beta, c, m)
    exp, scale = reference("N", "C", m, n, k, alpha, a, b, beta, c0)
    check(c, exp, scale, k, dtype)


@pytest.mark.parametrize("dtype", ZDTYPES)
def test_zgemm_no_product_scales_c(dtype):
    """k = 0 or alpha = 0: C := beta C (beta = 0 writes zeros without
    reading C)."""
    m, n = 40, 30
    a = zrand(m, 8, 7, dtype)
    c = zrand(m, n, 8, dtype)
    c0 = c.copy()
    beta = 0.5 + 2.0j
    H.zgemm("N", "N", m, n, 0, 1.0, a, m, a, 8, beta, c, m)
    exp = (beta * c0.astype(np.complex128))
    check(c, exp, np.abs(beta) * np.abs(c0) + 1e-300, 0, dtype)
    c[:] = np.nan
    H.zgemm("N", "N", m, n, 8, 0.0, a, m, a, 8, 0.0, c, m)
    assert np.all(c == 0)


@pytest.mark.parametrize("dtype", ZDTYPES)
def test_zgemm_device_operands_and_dev_entry(dtype):
    """Device operands through zgemm_ and through hwyblas_?gemm_dev give the
    same bits as host operands."""
    m, n, k = 129, 66, 97
    a = zrand(k, m, 9, dtype)
    b = zrand(k, n, 10, dtype)
    c = zrand(m, n, 11, dtype)
    alpha, beta = 1.5 + 0.25j, 0.5 - 0.5j
    host = c.copy(order="F")
    H.zgemm("C", "N", m, n, k, alpha, a, k, b, k, beta, host, m)
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    H.zgemm("C", "N", m, n, k, alpha, da, k, db, k, beta, dc, m, dtype=dtype)
    assert np.array_equal(dc.to_host(), host)
    dc2 = H.DeviceArray.from_host(c)
    assert H.zgemm_dev("C", "N", m, n, k, alpha, da, k, db, k, beta, dc2, m, dtype=dtype) == 0
    H.synchronize()
    assert np.array_equal(dc2.to_host(), host)
    exp, scale = reference("C", "N", m, n, k, alpha, a, b, beta, c)
    check(host, exp, scale, k, dtype)


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "C"), ("C", "N")])
def test_cblas_zgemm_row_major_is_the_transposed_problem(dtype, ta, tb):
    """Row-major cblas_?gemm equals the column-major call on the transposes,
    bit for bit (the mapping only swaps operands)."""
    m, n, k = 37, 29, 41
    enum = {"N": H.NoTrans, "T": H.Trans, "C": H.ConjTrans}
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    a_row = np.ascontiguousarray(zrand(ra, ca, 12, dtype))   # row-major storage
    b_row = np.ascontiguousarray(zrand(rb, cb, 13, dtype))
    c_row = np.ascontiguousarray(zrand(m, n, 14, dtype))
    c0 = c_row.copy()
    alpha, beta = 0.5 + 1.0j, -1.0 + 0.0j
    H.cblas_zgemm(H.RowMajor, enum[ta], enum[tb], m, n, k, alpha, a_row, ca, b_row, cb, beta,
                  c_row, n, dtype=dtype)
    # Column-major: C^T = op(B)^T op(A)^T, with A^T and B^T as stored.
    ct = np.array(c0.T, order="F")  # a copy: c0.T is already F-contiguous
    H.zgemm(tb, ta, n, m, k, alpha, np.asfortranarray(b_row.T), cb,
            np.asfortranarray(a_row.T), ca, beta, ct, n)
    assert np.array_equal(c_row, np.ascontiguousarray(ct.T))
    exp, scale = reference(ta, tb, m, n, k, alpha, np.asfortranarray(a_row), np.asfortranarray(b_row), beta,
                           np.asfortranarray(c0))
    check(np.asfortranarray(c_row), exp, scale, k, dtype)
