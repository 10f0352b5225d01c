"""Robustness of the drop-in entries on the GPU, through the C-ABI.

- Reentrancy: calls from several host threads at once (each thread has its
  own stream and staging buffers; hwyBLAS's per-thread state is at
  src/common.cc:19-20, :46-64) give each thread's results bit for bit as
  the same calls made one at a time.
- 64-bit offsets (after hwyBLAS's tests/semantics_test.cc:616-682): device
  operands with leading dimension 2^30 and 2^30 + 3 (column strides of 8 GiB),
  so every element past the first column lies beyond 2^31 elements.
- Differential fuzzing (after tests/fuzz/blas_fuzz.cc:371-414, :1155-1166):
  random shapes, transpositions, leading dimensions and alpha/beta for
  dgemm_/sgemm_/dtrsm_/strsm_ on host operands, with every unreferenced
  element poisoned with NaN (padding rows of A, B and C, the other triangle
  and the unit diagonal of a triangular A, C when beta = 0, A and B when
  alpha = 0); the results within hwyBLAS's bounds against the oracle, the
  padding of C / B untouched, and A and B (and C's padding) bitwise
  unchanged on return.
"""
import threading

import numpy as np
import pytest

import oracle
from helpers import gemm_bound_ratio, trsm_residual_ratio

pytestmark = pytest.mark.gpu


def test_concurrent_calls_from_threads(H):
    rng = np.random.default_rng(5)
    jobs = []
    for t in range(8):
        m, n, k = (int(x) for x in rng.integers(200, 900, 3))
        a = oracle.fill_uniform(m, k, 100 + t)
        b = oracle.fill_uniform(k, n, 200 + t)
        c = oracle.fill_uniform(m, n, 300 + t)
        tri = oracle.fill_uniform(m, m, 400 + t) / m + np.eye(m) * 2.0
        rhs = oracle.fill_uniform(m, 50, 500 + t)
        jobs.append((m, n, k, a, b, c, tri, rhs))

    def run(job):
        m, n, k, a, b, c, tri, rhs = job
        out = c.copy(order="F")
        H.gemm("N", "T" if m % 2 else "N", m, n, k, 1.25, a, m, b if m % 2 == 0 else
               np.asfortranarray(b.T), k if m % 2 == 0 else n, -0.5, out, m)
        x = rhs.copy(order="F")
        H.trsm("L", "L", "N", "N", m, 50, 1.0, tri, m, x, m)
        return out, x

    serial = [run(j) for j in jobs]
    results = [None] * len(jobs)

    def worker(i):
        for _ in range(3):
            results[i] = run(jobs[i])

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(jobs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for i, (want, got) in enumerate(zip(serial, results)):
        assert np.array_equal(want[0], got[0]) and np.array_equal(want[1], got[1]), i


@pytest.mark.parametrize("extra", [0, 3])
def test_leading_dimension_two_to_the_thirty(H, extra):
    """A (m x k) and C (m x n) on the device with ld = 2^30 + extra: column j
    starts 8 GiB x j in; B is compact. One allocation of (cols - 1) ld + m
    elements per matrix (16 GiB + a little for three columns)."""
    ld = (1 << 30) + extra
    m, n, k = 5, 3, 3
    a = oracle.fill_uniform(m, k, 7)
    b = oracle.fill_uniform(k, n, 8)
    c = oracle.fill_uniform(m, n, 9)
    L = H.lib()
    elems_a = (k - 1) * ld + m
    elems_c = (n - 1) * ld + m
    pa = L.hwyblas_amd_malloc(elems_a * 8)
    pc = L.hwyblas_amd_malloc(elems_c * 8)
    assert pa and pc
    try:
        for j in range(k):
            L.hwyblas_amd_memcpy(pa + j * ld * 8, np.ascontiguousarray(a[:, j]).ctypes.data,
                                 m * 8, 1)
        for j in range(n):
            L.hwyblas_amd_memcpy(pc + j * ld * 8, np.ascontiguousarray(c[:, j]).ctypes.data,
                                 m * 8, 1)
        db = H.DeviceArray.from_host(b)
        assert H.gemm_dev("N", "N", m, n, k, 1.5, pa, ld, db, k, -0.5, pc, ld,
                          dtype=np.float64) == 0
        H.synchronize()
        got = np.zeros((m, n), order="F")
        col = np.zeros(m)
        for j in range(n):
            L.hwyblas_amd_memcpy(col.ctypes.data, pc + j * ld * 8, m * 8, 2)
            got[:, j] = col
        assert gemm_bound_ratio(got, "N", "N", m, n, k, 1.5, a, m, b, k, -0.5, c, m) <= 1.0
        # TRSM with B at ld = 2^30 + extra: X = L^-1 B for a lower triangle.
        t = np.tril(oracle.fill_uniform(m, m, 10)) + np.eye(m) * 3.0
        dt = H.DeviceArray.from_host(t)
        assert H.trsm_dev("L", "L", "N", "N", m, n, 1.0, dt, m, pc, ld, dtype=np.float64) == 0
        H.synchronize()
        x = np.zeros((m, n), order="F")
        for j in range(n):
            L.hwyblas_amd_memcpy(col.ctypes.data, pc + j * ld * 8, m * 8, 2)
            x[:, j] = col
        assert np.allclose(t @ x, got, rtol=1e-13, atol=1e-13)
    finally:
        L.hwyblas_amd_free(pa)
        L.hwyblas_amd_free(pc)


def poisoned(rows, cols, ld, seed, dtype):
    x = np.full((ld, cols), np.nan, dtype=dtype, order="F")
    x[:rows, :] = oracle.fill_uniform(rows, cols, seed, dtype=dtype)
    return x


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_differential_fuzz_with_poisoned_unreferenced_data(H, dtype):
    rng = np.random.default_rng(2026 if dtype == np.float64 else 2027)
    for case in range(120):
        seed = 1000 * case
        if case % 2 == 0:  # GEMM
            ta, tb = "NT"[rng.integers(2)], "NT"[rng.integers(2)]
            m, n, k = (int(x) for x in rng.integers(0, 97, 3))
            alpha = float(rng.choice([0.0, 1.0, -1.5, 0.25]))
            beta = float(rng.choice([0.0, 1.0, -0.5, 2.0]))
            ra, ca = (m, k) if ta == "N" else (k, m)
            rb, cb = (k, n) if tb == "N" else (n, k)
            lda, ldb, ldc = (r + int(rng.integers(0, 4)) for r in (max(ra, 1), max(rb, 1),
                                                                   max(m, 1)))
            a = poisoned(ra, ca, lda, seed, dtype)
            b = poisoned(rb, cb, ldb, seed + 1, dtype)
            if alpha == 0.0:  # A and B must not be read
                a[:] = np.nan
                b[:] = np.nan
            c = poisoned(m, n, ldc, seed + 2, dtype)
            if beta == 0.0:  # C must not be read
                c[:m] = np.nan
            a0, b0, c0 = a.copy(order="F"), b.copy(order="F"), c.copy(order="F")
            H.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dtype=dtype)
            assert np.array_equal(a, a0, equal_nan=True) and np.array_equal(b, b0,
                                                                             equal_nan=True)
            assert np.array_equal(c[m:], c0[m:], equal_nan=True), "padding of C written"
            if m and n:
                assert not np.isnan(c[:m]).any(), (case, "an unreferenced NaN reached C")
                if alpha != 0.0:
                    cref = c0.copy(order="F")
                    if beta == 0.0:
                        cref[:m] = 0.0
                    r = gemm_bound_ratio(c, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, cref,
                                         ldc)
                    assert r <= 1.0, (case, ta, tb, m, n, k, r)
                else:
                    want = (beta * c0[:m, :n]) if beta != 0.0 else np.zeros((m, n), dtype)
                    if beta == 1.0:
                        want = c0[:m, :n]
                    assert np.array_equal(c[:m], want.astype(dtype)), case
        else:  # TRSM
            side, uplo = "LR"[rng.integers(2)], "LU"[rng.integers(2)]
            ta, diag = "NT"[rng.integers(2)], "NU"[rng.integers(2)]
            m, n = (int(x) for x in rng.integers(0, 97, 2))
            alpha = float(rng.choice([0.0, 1.0, -2.0]))
            ka = m if side == "L" else n
            lda = max(ka, 1) + int(rng.integers(0, 4))
            ldb = max(m, 1) + int(rng.integers(0, 4))
            a = np.full((lda, max(ka, 1)), np.nan, dtype=dtype, order="F")
            if ka:
                blk = oracle.fill_uniform(ka, ka, seed, dtype=dtype) / ka
                blk[np.arange(ka), np.arange(ka)] = (1.5 + rng.uniform(0, 0.5, ka)) * \
                    rng.choice([-1.0, 1.0], ka)
                keep = np.tril(np.ones((ka, ka), bool)) if uplo == "L" else \
                    np.triu(np.ones((ka, ka), bool))
                if diag == "U":
                    keep[np.arange(ka), np.arange(ka)] = False
                a[:ka, :ka] = np.where(keep, blk, np.nan)
            if alpha == 0.0:
                a[:] = np.nan
            b = poisoned(m, n, ldb, seed + 1, dtype)
            a0, b0 = a.copy(order="F"), b.copy(order="F")
            H.trsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, dtype=dtype)
            assert np.array_equal(a, a0, equal_nan=True)
            assert np.array_equal(b[m:], b0[m:], equal_nan=True), "padding of B written"
            if m and n:
                assert not np.isnan(b[:m]).any(), (case, "an unreferenced NaN reached X")
                if alpha == 0.0:
                    assert np.all(b[:m] == 0)
                else:
                    r = trsm_residual_ratio(side, uplo, ta, diag, m, n, alpha, a, lda, b0, ldb,
                                            b)
                    assert r <= 1.0, (case, side, uplo, ta, diag, m, n, r)
