"""LAPACKE_{s,d}potrf/potrs/posv on the GPU: column-major calls give the
Fortran routines' bits; row-major calls give the bits of the column-major
routine on the transposed data, with only the referenced triangle of A
written back; info is passed through (positive for a non-positive minor,
one lower for an argument the routine rejects)."""
import numpy as np
import pytest

import hwyblas_amd as H
import oracle

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32]
SENTINEL = 7.25


def spd(n, seed, dtype):
    g = oracle.fill_uniform(n, n, seed)
    a = g @ g.T / n + np.eye(n)
    return np.asfortranarray(a.astype(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_lapacke_potrf_both_layouts(dtype, uplo):
    n = 150
    a = spd(n, 1, dtype)
    lower = uplo == "L"
    i, j = np.indices((n, n))
    other = (i < j) if lower else (i > j)
    ref = a.copy(order="F")
    assert H.potrf(uplo, n, ref, n) == 0
    col = a.copy(order="F")
    col[other] = SENTINEL
    assert H.lapacke("potrf", H.ColMajor, uplo, n, col, n) == 0
    tri = ~other
    assert np.array_equal(col[tri], ref[tri]) and np.all(col[other] == SENTINEL)
    # Row-major: the logical matrix's triangle uplo, stored by rows; the
    # routine factors the column-major copy of that triangle.
    row = np.ascontiguousarray(a)
    row[other] = SENTINEL
    assert H.lapacke("potrf", H.RowMajor, uplo, n, row, n) == 0
    assert np.array_equal(row[tri], ref[tri]) and np.all(row[other] == SENTINEL)


@pytest.mark.parametrize("dtype", DTYPES)
def test_lapacke_posv_and_potrs_row_major(dtype):
    n, nrhs = 130, 5
    a = spd(n, 2, dtype)
    b = oracle.fill_uniform(n, nrhs, 3).astype(dtype)
    af = a.copy(order="F")
    xf = np.asfortranarray(b.copy())
    assert H.posv("U", n, nrhs, af, n, xf, n) == 0
    ar = np.ascontiguousarray(a)
    xr = np.ascontiguousarray(b.copy())
    assert H.lapacke("posv", H.RowMajor, "U", n, nrhs, ar, n, xr, nrhs) == 0
    assert np.array_equal(xr, xf)
    # potrs with the row-major factor just computed.
    x2 = np.ascontiguousarray(b.copy())
    assert H.lapacke("potrs", H.RowMajor, "U", n, nrhs, ar, n, x2, nrhs) == 0
    assert np.array_equal(x2, xf)


@pytest.mark.parametrize("dtype", DTYPES)
def test_lapacke_info_passthrough(dtype):
    n = 70
    a = spd(n, 4, dtype)
    a[40, 40] = -1.0
    assert H.lapacke("potrf", H.ColMajor, "L", n, a.copy(order="F"), n) == 41
    with H.ErrorCapture() as cap:
        assert H.lapacke("potrf", H.ColMajor, "X", n, a.copy(order="F"), n) == -2
    name = "DPOTRF" if dtype == np.float64 else "SPOTRF"
    assert cap.calls == [(name, 1)]
