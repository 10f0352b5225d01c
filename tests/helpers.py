"""Shared test helpers: seeded operands and the componentwise bound checks.

The bounds are hwyBLAS's own test tolerances:
 - GEMM: |C - C_ref| <= 4 (k + 2) eps * (|alpha| |A||B| + |beta| |C| + tiny)
   (tests/gemm_test.cc:72-101);
 - TRSM: |op(A) X - alpha B| <= 8 (ka + 2) eps * (|op(A)||X| + |alpha B|)
   (tests/level3_test.cc:482-520).
C_ref and the error scales come from the oracle's double-double recomputation.
"""
import numpy as np

import oracle

EPS = {np.dtype(np.float64): np.finfo(np.float64).eps, np.dtype(np.float32): np.finfo(np.float32).eps}


def eps(dtype):
    return EPS[np.dtype(dtype)]


def op_shape(t, rows, cols):
    """Stored shape of X for op(X) of shape rows x cols."""
    return (rows, cols) if t == "N" else (cols, rows)


def gemm_operands(ta, tb, m, n, k, pad, dtype, seed):
    ra, ca = op_shape(ta, m, k)
    rb, cb = op_shape(tb, k, n)
    lda, ldb, ldc = max(1, ra + pad), max(1, rb + pad), max(1, m + pad)
    a = oracle.fill_uniform(ra, ca, seed, dtype=dtype, ld=lda)
    b = oracle.fill_uniform(rb, cb, seed + 1, dtype=dtype, ld=ldb)
    c = oracle.fill_uniform(m, n, seed + 2, dtype=dtype, ld=ldc)
    # Padding rows get a recognisable value so that writes to them show.
    if pad:
        c[m:, :] = 7.25
    return a, lda, b, ldb, c, ldc


def gemm_bound_ratio(got, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc):
    """max |got - ref| / (tol * scale) over the m x n result (<= 1 passes)."""
    exp, bnd = oracle.gemm_ref(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c0, ldc)
    tol = 4.0 * (k + 2) * eps(got.dtype)
    g = got[:m, :n].astype(np.float64)
    err = np.abs(g - exp)
    both_nan = np.isnan(g) & np.isnan(exp)
    err[both_nan] = 0
    err[np.isnan(err)] = np.inf
    if m == 0 or n == 0:
        return 0.0
    return float(np.max(err / (tol * bnd)))


def trsm_operands(side, uplo, diag, m, n, dtype, seed, lda_pad=1, ldb=None, scale_offdiag=True):
    """A triangle as in hwyBLAS's tests (tests/test_util.h:266-304): off-diagonal
    entries uniform[-1,1] scaled by 1/ka, diagonal +-(1.5 +- 0.5) (unit: NaN,
    it must not be read), unreferenced triangle NaN."""
    ka = m if side == "L" else n
    lda = ka + lda_pad
    ldb = ldb or (m + 2)
    a = oracle.fill_uniform(ka, ka, seed, dtype=dtype, ld=lda)
    rng = np.random.default_rng(seed)
    if scale_offdiag:
        a[:ka, :ka] /= ka
    iu = np.triu_indices(ka, 1)
    il = np.tril_indices(ka, -1)
    if uplo == "L":
        a[iu] = np.nan
    else:
        a[il] = np.nan
    d = (1.5 + rng.uniform(-0.5, 0.5, ka)) * rng.choice([-1.0, 1.0], ka)
    a[np.arange(ka), np.arange(ka)] = np.nan if diag == "U" else d
    a[ka:, :] = np.nan
    b = oracle.fill_uniform(m, n, seed + 1, dtype=dtype, ld=ldb)
    b[m:, :] = 7.25
    return a, lda, b, ldb


def trsm_residual_ratio(side, uplo, ta, diag, m, n, alpha, a, lda, b0, ldb, x, which=None):
    ka = m if side == "L" else n
    other = n if side == "L" else m
    if which is None:
        which = np.arange(other)
    r = oracle.trsm_residual(side, uplo, ta, diag, m, n, alpha, a, lda, b0, ldb, x, ldb, which)
    tol = 8.0 * (ka + 2) * eps(b0.dtype)
    return float(np.max(r) / tol) if len(r) else 0.0


def permuted_triangular_system(n, nrhs, seed, trans="N"):
    """(A, X, B) with A = P U: U upper triangular with integer entries, a
    power-of-two diagonal at least the row's off-diagonal sum and
    off-diagonals in {0, -1, -2}, none inside the 64 x 64 diagonal blocks
    (so a block solve by substitution or by the block's inverse is exact);
    P a random row permutation; X integer, nonzero, |X| <= 50;
    B = op(A) X, exact. Partial pivoting recovers P with L = I, every solve
    is exact integer arithmetic (so refinement stops with R = 0), and
    inv(op(A)) >= 0 entrywise."""
    rng = np.random.default_rng(seed)
    u = np.triu(-rng.integers(0, 3, (n, n)).astype(float) * (rng.random((n, n)) < 0.2), 1)
    blk = np.arange(n) // 64
    u[blk[:, None] == blk[None, :]] = 0.0
    rs = -u.sum(1)
    u[np.arange(n), np.arange(n)] = 2.0 ** np.ceil(np.log2(rs + 1)) * rng.choice([1, 2], n)
    a = np.asfortranarray(u[rng.permutation(n)])
    x = rng.integers(1, 51, (n, nrhs)).astype(float) * rng.choice([-1.0, 1.0], (n, nrhs))
    opa = a if trans == "N" else a.T
    return a, np.asfortranarray(x), np.asfortranarray(opa @ x)


def forward_bound(opa, x, b):
    """gerfs's forward error bound for an exact solution (R = 0), by its
    defining formula (src/lapack/gesvx.cc:469-476) with an explicit
    inverse: || |inv(op(A))| W ||inf / max|X| per column, W = nz eps
    (|op(A)| |X| + |B|) + safe1 where that is <= safe2, nz = n + 1."""
    n = opa.shape[0]
    e = np.finfo(np.float64).eps / 2
    safe1 = (n + 1) * np.finfo(np.float64).tiny
    d = np.abs(opa) @ np.abs(x) + np.abs(b)
    w = (n + 1) * e * d + np.where(d > safe1 / e, 0.0, safe1)
    return (np.abs(np.linalg.inv(opa)) @ w).max(0) / np.abs(x).max(0)
