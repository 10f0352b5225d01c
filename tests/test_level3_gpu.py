"""SYMM, SYRK, SYR2K, GEMMTR/GEMMT and TRMM on the GPU through the Fortran
C-ABI (src/level3.cc semantics). Each result is checked against the
double-double reference of the same product with the structured operand
written out densely, with hwyBLAS's GEMM bound 4 (k + 2) eps
(tests/gemm_test.cc:72-101). The unreferenced triangle of a structured
operand holds NaN (it must not be read), a unit diagonal holds NaN, the
other triangle of C and the padding rows of C or B hold sentinels that must
come back unchanged."""
import numpy as np
import pytest

import oracle
from helpers import eps

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32]
SENTINEL = 7.25


def bound_ratio(got, mask, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc):
    """max |got - ref| / (4 (k + 2) eps |ref scale|) over the entries in mask."""
    exp, bnd = oracle.gemm_ref(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc)
    tol = 4.0 * (k + 2) * eps(got.dtype)
    err = np.abs(got[:m, :n].astype(np.float64) - exp)
    err[np.isnan(err)] = np.inf
    r = err / (tol * bnd)
    return float(np.max(r[mask])) if mask.any() else 0.0


def tri_mask(n, lower):
    i, j = np.indices((n, n))
    return i >= j if lower else i <= j


def rand(rows, cols, seed, dtype, ld=None):
    return oracle.fill_uniform(rows, cols, seed, dtype=dtype, ld=ld or rows)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_symm(H, dtype, side, uplo):
    for (m, n, alpha, beta) in ((70, 45, 1.5, -0.5), (257, 130, 1.0, 0.0), (1, 3, -2.0, 1.0)):
        ra = m if side == "L" else n
        lda, ldb, ldc = ra + 2, m + 1, m + 3
        a = rand(ra, ra, 1, dtype, lda)
        stored = tri_mask(ra, uplo == "L")
        sym = np.where(stored, a[:ra, :ra], a[:ra, :ra].T).astype(dtype)  # mirrored
        a[:ra, :ra][~stored] = np.nan  # never read
        b = rand(m, n, 2, dtype, ldb)
        c = rand(m, n, 3, dtype, ldc)
        c[m:, :] = SENTINEL
        if beta == 0.0:
            c[:m, :] = np.nan  # beta = 0: C is not read
        c0 = c.copy(order="F")
        H.symm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc)
        assert np.all(c[m:, :] == SENTINEL)
        symf = np.asfortranarray(sym)
        full = np.ones((m, n), dtype=bool)
        if side == "L":
            r = bound_ratio(c, full, "N", "N", m, n, ra, alpha, symf, ra, b, ldb, beta, c0, ldc)
        else:
            r = bound_ratio(c, full, "N", "N", m, n, ra, alpha, b, ldb, symf, ra, beta, c0, ldc)
        assert r <= 1.0, (m, n, r)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T", "C"])
def test_syrk(H, dtype, uplo, trans):
    for (n, k, alpha, beta) in ((70, 33, 1.5, -0.5), (300, 200, 1.0, 0.0), (5, 1, 2.0, 1.0)):
        ra, ca = (n, k) if trans == "N" else (k, n)
        lda, ldc = ra + 1, n + 2
        a = rand(ra, ca, 4, dtype, lda)
        c = rand(n, n, 5, dtype, ldc)
        c[n:, :] = SENTINEL
        low = tri_mask(n, uplo == "L")
        c[:n, :n][~low] = SENTINEL
        if beta == 0.0:
            c[:n, :n][low] = np.nan
        c0 = c.copy(order="F")
        H.syrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc)
        assert np.all(c[n:, :] == SENTINEL) and np.all(c[:n, :n][~low] == SENTINEL)
        ta, tb = ("N", "T") if trans == "N" else ("T", "N")
        r = bound_ratio(c, low, ta, tb, n, n, k, alpha, a, lda, a, lda, beta, c0, ldc)
        assert r <= 1.0, (n, k, r)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_syr2k(H, dtype, uplo, trans):
    n, k, alpha, beta = 150, 70, 1.25, 0.5
    ra, ca = (n, k) if trans == "N" else (k, n)
    lda, ldb, ldc = ra + 1, ra + 2, n + 1
    a = rand(ra, ca, 6, dtype, lda)
    b = rand(ra, ca, 7, dtype, ldb)
    c = rand(n, n, 8, dtype, ldc)
    c[n:, :] = SENTINEL
    low = tri_mask(n, uplo == "L")
    c[:n, :n][~low] = SENTINEL
    c0 = c.copy(order="F")
    H.syr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert np.all(c[n:, :] == SENTINEL) and np.all(c[:n, :n][~low] == SENTINEL)
    # op(A) op(B)^T + op(B) op(A)^T as one product of depth 2k: [op(A) op(B)] [op(B) op(A)]^T.
    opa = a[:ra, :ca] if trans == "N" else a[:ra, :ca].T
    opb = b[:ra, :ca] if trans == "N" else b[:ra, :ca].T
    left = np.asfortranarray(np.hstack([opa, opb]))
    right = np.asfortranarray(np.hstack([opb, opa]))
    r = bound_ratio(c, low, "N", "T", n, n, 2 * k, alpha, left, n, right, n, beta, c0, ldc)
    assert r <= 1.0, r


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_gemmtr(H, dtype, uplo, ta, tb):
    for name in ("gemmtr", "gemmt"):
        n, k, alpha, beta = 211, 90, -1.5, 0.75
        ra, ca = (n, k) if ta == "N" else (k, n)
        rb, cb = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = ra + 1, rb + 3, n + 2
        a = rand(ra, ca, 9, dtype, lda)
        b = rand(rb, cb, 10, dtype, ldb)
        c = rand(n, n, 11, dtype, ldc)
        c[n:, :] = SENTINEL
        low = tri_mask(n, uplo == "L")
        c[:n, :n][~low] = SENTINEL
        c0 = c.copy(order="F")
        H.gemmtr(uplo, ta, tb, n, k, alpha, a, lda, b, ldb, beta, c, ldc, name=name)
        assert np.all(c[n:, :] == SENTINEL) and np.all(c[:n, :n][~low] == SENTINEL)
        r = bound_ratio(c, low, ta, tb, n, n, k, alpha, a, lda, b, ldb, beta, c0, ldc)
        assert r <= 1.0, (name, r)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("ta", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trmm(H, dtype, side, uplo, ta, diag):
    for (m, n, alpha) in ((70, 45, 1.5), (130, 257, -1.0)):
        ra = m if side == "L" else n
        lda, ldb = ra + 1, m + 2
        a = rand(ra, ra, 12, dtype, lda)
        low = tri_mask(ra, uplo == "L")
        dense = np.where(low, a[:ra, :ra], 0).astype(dtype)
        if diag == "U":
            np.fill_diagonal(dense, 1)
            np.fill_diagonal(a[:ra, :ra], np.nan)  # a unit diagonal is not read
        a[:ra, :ra][~low] = np.nan
        b = rand(m, n, 13, dtype, ldb)
        b[m:, :] = SENTINEL
        x = np.asfortranarray(b[:m, :n].copy())
        H.trmm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb)
        assert np.all(b[m:, :] == SENTINEL)
        dense = np.asfortranarray(dense)
        zero = np.zeros((m, n), dtype=dtype, order="F")
        full = np.ones((m, n), dtype=bool)
        if side == "L":
            r = bound_ratio(b, full, ta, "N", m, n, m, alpha, dense, ra, x, m, 0.0, zero, m)
        else:
            r = bound_ratio(b, full, "N", ta, m, n, n, alpha, x, m, dense, ra, 0.0, zero, m)
        assert r <= 1.0, (m, n, r)


@pytest.mark.parametrize("dtype", DTYPES)
def test_quick_returns(H, dtype):
    """alpha = 0: SYRK scales only its triangle (beta = 0 writes zeros without
    reading); TRMM writes zeros without reading A or B; SYMM scales C."""
    n = 40
    c = rand(n, n, 14, dtype)
    low = tri_mask(n, True)
    c0 = c.copy()
    H.syrk("L", "N", n, 7, 0.0, np.full((n, 7), np.nan, dtype=dtype, order="F"), n, 2.0, c, n)
    assert np.array_equal(c[low], (c0 * 2)[low]) and np.array_equal(c[~low], c0[~low])
    c[:] = np.nan
    H.gemmtr("U", "N", "N", n, 0, 1.0, c, n, c, 1, 0.0, c, n)
    assert np.all(c[tri_mask(n, False)] == 0) and np.isnan(c[~tri_mask(n, False)]).all()
    b = np.full((n, 3), np.nan, dtype=dtype, order="F")
    H.trmm("L", "L", "N", "N", n, 3, 0.0, np.full((n, n), np.nan, dtype=dtype, order="F"), n, b, n)
    assert np.all(b == 0)
    c = rand(n, 5, 15, dtype)
    c0 = c.copy()
    H.symm("L", "U", n, 5, 0.0, np.full((n, n), np.nan, dtype=dtype, order="F"), n, c0, n, 3.0,
           c, n)
    assert np.array_equal(c, c0 * dtype(3.0))


def test_device_operands(H):
    """Device-resident operands are used in place (SYRK and TRMM)."""
    n, k = 200, 64
    a = rand(n, k, 16, np.float64)
    c = rand(n, n, 17, np.float64)
    da, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(c)
    H.syrk("U", "N", n, k, 1.0, da, n, 0.5, dc, n, dtype=np.float64)
    want = c.copy(order="F")
    H.syrk("U", "N", n, k, 1.0, a, n, 0.5, want, n)
    assert np.array_equal(dc.to_host(), want)
    t = rand(n, n, 18, np.float64)
    b = rand(n, 9, 19, np.float64)
    dt, db = H.DeviceArray.from_host(t), H.DeviceArray.from_host(b)
    H.trmm("L", "U", "T", "N", n, 9, 2.0, dt, n, db, n, dtype=np.float64)
    want = b.copy(order="F")
    H.trmm("L", "U", "T", "N", n, 9, 2.0, t, n, want, n)
    assert np.array_equal(db.to_host(), want)


@pytest.mark.parametrize("dtype,n", [(np.float64, 2048), (np.float32, 4096)])
def test_large_triangle_and_k_range_products(H, dtype, n):
    """Sizes where the default configuration is the DMA-fed kernel: SYRK's
    triangle (a triangle-of-C launch) equals the full product A A^T of the
    same library bit for bit and leaves the other triangle alone; TRMM (a
    k-range launch that skips the factor's zero slices) equals the product
    with the triangle written out densely."""
    L = H.lib()
    assert L.hwyblas_gemm_config_stages(
        L.hwyblas_gemm_default_config(np.dtype(dtype).itemsize, b"N", b"T", n, n, n)) > 0
    k = 512
    a = H.DeviceArray(n, k, dtype)
    a.fill_uniform(n, k, 31)
    full = H.DeviceArray(n, n, dtype)
    H.gemm("N", "T", n, n, k, 1.5, a, n, a, n, 0.0, full, n, dtype=dtype)
    want = full.to_host()
    for uplo in "LU":
        c = H.DeviceArray(n, n, dtype)
        c.fill_uniform(n, n, 32)
        c0 = c.to_host()
        H.syrk(uplo, "N", n, k, 1.5, a, n, 0.0, c, n, dtype=dtype)
        got = c.to_host()
        tri = tri_mask(n, uplo == "L")
        assert np.array_equal(got[tri], want[tri]), uplo
        assert np.array_equal(got[~tri], c0[~tri]), uplo
    t = H.DeviceArray(n, n, dtype)
    t.fill_uniform(n, n, 33)
    th = t.to_host()
    b = H.DeviceArray(n, n, dtype)
    b.fill_uniform(n, n, 34)
    bh = b.to_host()
    for uplo, ta in (("L", "N"), ("U", "T")):
        dense = np.asfortranarray(np.where(tri_mask(n, uplo == "L"), th, 0).astype(dtype))
        dd = H.DeviceArray.from_host(dense)
        ref = H.DeviceArray(n, n, dtype)
        H.gemm(ta, "N", n, n, n, 1.0, dd, n, b, n, 0.0, ref, n, dtype=dtype)
        bb = H.DeviceArray.from_host(bh)
        H.trmm("L", uplo, ta, "N", n, n, 1.0, t, n, bb, n, dtype=dtype)
        assert np.array_equal(bb.to_host(), ref.to_host()), (uplo, ta)


def _rowmajor(rows, cols, seed, dtype):
    """A row-major rows x cols matrix (C order) with uniform[-1, 1] entries."""
    return np.ascontiguousarray(rand(rows, cols, seed, dtype))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("layout", ["row", "col"])
def test_cblas_level3(H, dtype, layout):
    """cblas_?symm/syrk/syr2k/gemmtr/gemmt/trmm in both layouts against a
    float64 numpy evaluation of the same mathematics (relative 1e-12 for
    FP64, 2e-5 for FP32, scaled by the inputs' magnitudes); untouched parts
    stay untouched."""
    lay = H.RowMajor if layout == "row" else H.ColMajor
    order = "C" if layout == "row" else "F"
    tol = 1e-12 if dtype == np.float64 else 2e-5
    mk = (lambda r, c, sd: np.asarray(rand(r, c, sd, dtype), order=order).copy(order=order))

    def ld(x):
        return x.shape[1] if layout == "row" else x.shape[0]

    m, n, k = 37, 29, 23
    f64 = lambda x: x.astype(np.float64)
    # SYMM, side L, uplo L: only the lower triangle of A is read.
    a = mk(m, m, 21)
    low = tri_mask(m, True)
    s_full = np.where(low, f64(a), f64(a).T)
    a[~low] = np.nan
    b, c = mk(m, n, 22), mk(m, n, 23)
    want = 1.5 * s_full @ f64(b) - 0.5 * f64(c)
    H.cblas_level3("symm", lay, H.Left, H.Lower, m, n, 1.5, a, ld(a), b, ld(b), -0.5, c, ld(c),
                   dtype=dtype)
    assert np.max(np.abs(f64(c) - want)) <= tol * m * 4
    # SYRK, uplo U, trans T: C (n x n) upper := A^T A + C, A is k x n.
    a = mk(k, n, 24)
    c = mk(n, n, 25)
    c0 = f64(c)
    up = tri_mask(n, False)
    H.cblas_level3("syrk", lay, H.Upper, H.Trans, n, k, 1.0, a, ld(a), 1.0, c, ld(c), dtype=dtype)
    want = f64(a).T @ f64(a) + c0
    assert np.max(np.abs(f64(c)[up] - want[up])) <= tol * k * 4
    assert np.array_equal(f64(c)[~up], c0[~up])
    # SYR2K, uplo L, trans N: C lower := A B^T + B A^T (A, B n x k).
    a, b = mk(n, k, 26), mk(n, k, 27)
    c = mk(n, n, 28)
    c0 = f64(c)
    lo = tri_mask(n, True)
    H.cblas_level3("syr2k", lay, H.Lower, H.NoTrans, n, k, 1.0, a, ld(a), b, ld(b), 0.0, c, ld(c),
                   dtype=dtype)
    want = f64(a) @ f64(b).T + f64(b) @ f64(a).T
    assert np.max(np.abs(f64(c)[lo] - want[lo])) <= tol * k * 8
    assert np.array_equal(f64(c)[~lo], c0[~lo])
    # GEMMTR / GEMMT, uplo U, A^T (k x n stored) times B (k x n).
    for name in ("gemmtr", "gemmt"):
        a, b = mk(k, n, 29), mk(k, n, 30)
        c = mk(n, n, 31)
        c0 = f64(c)
        H.cblas_level3(name, lay, H.Upper, H.Trans, H.NoTrans, n, k, 2.0, a, ld(a), b, ld(b), 0.5,
                       c, ld(c), dtype=dtype)
        want = 2.0 * f64(a).T @ f64(b) + 0.5 * c0
        assert np.max(np.abs(f64(c)[up] - want[up])) <= tol * k * 8, name
        assert np.array_equal(f64(c)[~up], c0[~up])
    # TRMM, side R, uplo U, A^T, unit diagonal: B (m x n) := B op(A).
    a = mk(n, n, 32)
    t = np.where(tri_mask(n, False), f64(a), 0.0)
    np.fill_diagonal(t, 1.0)
    a[~tri_mask(n, False)] = np.nan
    np.fill_diagonal(a, np.nan)
    b = mk(m, n, 33)
    want = f64(b) @ t.T
    H.cblas_level3("trmm", lay, H.Right, H.Upper, H.Trans, H.Unit, m, n, 1.0, a, ld(a), b, ld(b),
                   dtype=dtype)
    assert np.max(np.abs(f64(b) - want)) <= tol * n * 4
