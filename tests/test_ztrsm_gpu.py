"""Complex TRSM (ctrsm_, ztrsm_, cblas_?trsm, hwyblas_?trsm_dev) on the GPU.
Every side x uplo x op (N, T, C) x diag case is checked by its residual,
|op(A) X - alpha B| <= tol (|op(A)| |X| + |alpha| |B|) componentwise in
complex128 (left; X op(A) for the right side), tol = 8 (k + 2) eps with k the
triangle's order, the bound shape of hwyBLAS's TRSM tests. The triangle not
referenced and a unit diagonal hold NaN (never read); B's padding rows must
come back unchanged. Orders above 64 exercise the recursion."""
import itertools

import numpy as np
import pytest

import hwyblas_amd as H
from test_zgemm_gpu import zeps, zrand

pytestmark = pytest.mark.gpu
ZDTYPES = [np.complex128, np.complex64]


def triangle(k, lower, unit, seed, dtype, pad=2):
    """A (k + pad) x k array whose referenced triangle is random with a
    dominant diagonal; the other triangle (and a unit diagonal) hold NaN."""
    a = zrand(k + pad, k, seed, dtype)
    i, j = np.indices((k + pad, k))
    a[i == j] += (k + 4)
    if lower:
        a[i < j] = np.nan
    else:
        a[(i > j) & (i < k)] = np.nan
    if unit:
        a[i == j] = np.nan
    return np.asfortranarray(a)


def dense_op(a, k, lower, unit, ta):
    t = a[:k, :k].astype(np.complex128).copy()
    i, j = np.indices((k, k))
    t[(i < j) if lower else (i > j)] = 0
    if unit:
        t[i == j] = 1
    if ta == "T":
        t = t.T
    elif ta == "C":
        t = np.conj(t.T)
    return t


def check_residual(side, t, x, alpha, b0, k, dtype):
    x = x.astype(np.complex128)
    b0 = b0.astype(np.complex128)
    lhs = t @ x if side == "L" else x @ t
    scale = (np.abs(t) @ np.abs(x) if side == "L" else np.abs(x) @ np.abs(t))
    scale = scale + abs(alpha) * np.abs(b0) + np.finfo(np.float64).tiny
    r = np.abs(lhs - alpha * b0) / (8.0 * (k + 2) * zeps(dtype) * scale)
    assert np.all(np.isfinite(r)) and float(np.max(r)) <= 1.0, float(np.max(r))


@pytest.mark.parametrize("dtype", ZDTYPES)
@pytest.mark.parametrize("side,uplo,ta,diag",
                         list(itertools.product("LR", "UL", "NTC", "NU")))
def test_ztrsm_every_case(dtype, side, uplo, ta, diag):
    m, n, pad = 150, 70, 3
    k = m if side == "L" else n
    a = triangle(k, uplo == "L", diag == "U", 1, dtype)
    b = zrand(m + pad, n, 2, dtype)
    b0 = b.copy(order="F")
    alpha = 0.75 - 1.25j
    H.ztrsm(side, uplo, ta, diag, m, n, alpha, a, k + 2, b, m + pad)
    t = dense_op(a, k, uplo == "L", diag == "U", ta)
    check_residual(side, t, b[:m], alpha, b0[:m], k, dtype)
    assert np.array_equal(b[m:], b0[m:])


@pytest.mark.parametrize("dtype", ZDTYPES)
def test_ztrsm_alpha_zero_and_device_forms(dtype):
    """alpha = 0 writes zeros without reading A (A all NaN); device operands
    through ztrsm_ and hwyblas_?trsm_dev give the host result's bits."""
    m, n = 97, 130
    a_nan = np.full((m, m), np.nan, dtype=dtype, order="F")
    b = zrand(m, n, 3, dtype)
    H.ztrsm("L", "L", "N", "N", m, n, 0.0, a_nan, m, b, m)
    assert np.all(b == 0)
    a = triangle(n, False, False, 4, dtype, pad=0)
    b = zrand(m, n, 5, dtype)
    host = b.copy(order="F")
    H.ztrsm("R", "U", "C", "N", m, n, 2.0 + 0.5j, a, n, host, m)
    da, db = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b)
    H.ztrsm("R", "U", "C", "N", m, n, 2.0 + 0.5j, da, n, db, m, dtype=dtype)
    assert np.array_equal(db.to_host(), host)
    db2 = H.DeviceArray.from_host(b)
    assert H.ztrsm_dev("R", "U", "C", "N", m, n, 2.0 + 0.5j, da, n, db2, m, dtype=dtype) == 0
    H.synchronize()
    assert np.array_equal(db2.to_host(), host)


@pytest.mark.parametrize("dtype", ZDTYPES)
def test_cblas_ztrsm_row_major_is_the_flipped_problem(dtype):
    """Row-major cblas_?trsm equals the column-major call with the other side
    and triangle on the transposes, bit for bit."""
    m, n = 90, 70
    a = triangle(m, True, False, 6, dtype, pad=0)
    a_row = np.ascontiguousarray(a)              # the same logical A, row-major
    b_row = np.ascontiguousarray(zrand(m, n, 7, dtype))
    b0 = b_row.copy()
    alpha = -0.5 + 0.25j
    H.cblas_ztrsm(H.RowMajor, H.Left, H.Lower, H.ConjTrans, H.NonUnit, m, n, alpha, a_row, m,
                  b_row, n, dtype=dtype)
    bt = np.array(b0.T, order="F")
    H.ztrsm("R", "U", "C", "N", n, m, alpha, np.asfortranarray(a_row.T), m, bt, n)
    assert np.array_equal(b_row, np.ascontiguousarray(bt.T))
    t = dense_op(a, m, True, False, "C")
    check_residual("L", t, b_row, alpha, b0, m, dtype)
