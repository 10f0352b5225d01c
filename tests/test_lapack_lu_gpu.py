"""dgetrf_, dgetrs_, dgesv_ and LAPACKE_d{getrf,getrs,gesv} on the GPU: the
Fortran entries with host operands give the device entries' bits (factors,
pivots, solutions); dgesv_ = dgetrf_ then dgetrs_; LAPACKE row-major gives the
column-major routine's bits on the same logical matrices; singular info is
passed through; the solution's residual is within the LU bound."""
import numpy as np
import pytest

import hwyblas_amd as H
import oracle

pytestmark = pytest.mark.gpu


def system(n, nrhs, seed):
    a = oracle.fill_uniform(n, n, seed) + np.eye(n) * 0.5
    b = oracle.fill_uniform(n, nrhs, seed + 1)
    return np.asfortranarray(a), np.asfortranarray(b)


def test_dgetrf_host_matches_device_entry():
    m, n = 300, 260
    a = np.asfortranarray(oracle.fill_uniform(m, n, 5))
    host = a.copy(order="F")
    piv = np.zeros(min(m, n), dtype=np.int32)
    assert H.getrf(m, n, host, m, piv) == 0
    da = H.DeviceArray.from_host(a)
    dpiv, info = H.getrf_dev(da)
    assert info == 0
    assert np.array_equal(da.to_host(), host)
    assert np.array_equal(dpiv.to_host().ravel(), piv)
    # Device operands through the Fortran entry: the same bits.
    db = H.DeviceArray.from_host(a)
    dp2 = H.DeviceArray(min(m, n), 1, np.int32)
    assert H.getrf(m, n, db, m, dp2) == 0
    assert np.array_equal(db.to_host(), host)
    assert np.array_equal(dp2.to_host().ravel(), piv)


@pytest.mark.parametrize("trans", ["N", "T"])
def test_dgesv_and_dgetrs(trans):
    n, nrhs = 257, 7
    a, b = system(n, nrhs, 11)
    lu = a.copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    x = b.copy(order="F")
    if trans == "N":
        assert H.gesv(n, nrhs, lu, n, piv, x, n) == 0
        lu2, piv2, x2 = a.copy(order="F"), np.zeros(n, dtype=np.int32), b.copy(order="F")
        assert H.getrf(n, n, lu2, n, piv2) == 0
        assert H.getrs("N", n, nrhs, lu2, n, piv2, x2, n) == 0
        assert np.array_equal(x, x2) and np.array_equal(lu, lu2)
    else:
        assert H.getrf(n, n, lu, n, piv) == 0
        assert H.getrs("T", n, nrhs, lu, n, piv, x, n) == 0
    op = a if trans == "N" else a.T
    r = np.abs(op @ x - b)
    bound = (np.abs(op) @ np.abs(x) + np.abs(b)) * 8 * (n + 2) * np.finfo(np.float64).eps * 64
    assert np.all(r <= bound)


def test_lapacke_lu_row_major_and_info():
    n, nrhs = 200, 3
    a, b = system(n, nrhs, 21)
    lu, piv, x = a.copy(order="F"), np.zeros(n, dtype=np.int32), b.copy(order="F")
    assert H.gesv(n, nrhs, lu, n, piv, x, n) == 0
    ar, br = np.ascontiguousarray(a), np.ascontiguousarray(b)
    pr = np.zeros(n, dtype=np.int32)
    assert H.lapacke_lu("gesv", H.RowMajor, n, nrhs, ar, n, pr, br, nrhs) == 0
    assert np.array_equal(br, x) and np.array_equal(ar, lu) and np.array_equal(pr, piv)
    ac = a.copy(order="F")
    pc = np.zeros(n, dtype=np.int32)
    assert H.lapacke_lu("getrf", H.ColMajor, n, n, ac, n, pc) == 0
    assert np.array_equal(ac, lu) and np.array_equal(pc, piv)
    bt = b.copy(order="F")
    assert H.lapacke_lu("getrs", H.ColMajor, "N", n, nrhs, ac, n, pc, bt, n) == 0
    assert np.array_equal(bt, x)
    # Singular: a zero column gives info = its index (1-based) or earlier.
    s = a.copy(order="F")
    s[:, 17] = 0.0
    ps = np.zeros(n, dtype=np.int32)
    info = H.lapacke_lu("getrf", H.ColMajor, n, n, s, n, ps)
    assert 0 < info <= 18
