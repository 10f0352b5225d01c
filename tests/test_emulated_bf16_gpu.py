"""Emulated DGEMM with BF16 slices (Ozaki scheme I) on the GPU through the C-ABI.

Pinned two ways: (1) bit for bit against the oracle's CPU restatement of the
same algorithm (digits and chunk sums are exact, the FP64 accumulation runs
in the same order), which checks the digit planes, the BF16 MFMA operand
layout, the chunked FP32-to-FP64 accumulation, the pair order and the fused
epilogue; (2) against the double-double reference with the native DGEMM
componentwise bound 4 (k + 2) u, with the slice count chosen automatically,
on uniform operands, rows/columns scaled by 10^s (s in [-8, 8]) and entries
u 2^e (e in [-30, 30]).
"""
import itertools

import numpy as np
import pytest

import oracle
from helpers import gemm_bound_ratio, gemm_operands

pytestmark = pytest.mark.gpu
TRANS = list(itertools.product("NT", "NT"))


def run_bf16(H, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, slices):
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    info, used = H.dgemm_emulated_bf16_dev(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc,
                                           ldc, slices)
    assert info == 0
    H.synchronize()
    return dc.to_host(), used


@pytest.mark.parametrize("ta,tb", TRANS)
def test_bitwise_equal_to_restatement(H, ta, tb):
    # Shapes cross the 256 x 128 tiles and the 1024-long k chunks (ragged
    # ends), with padded leading dimensions; slice counts 1..16.
    for (m, n, k, pad, s) in ((37, 29, 100, 1, 7), (300, 257, 2600, 0, 5), (1, 513, 65, 2, 3),
                              (256, 128, 1024, 0, 16), (70, 3, 1, 1, 1), (260, 131, 1300, 0, 9)):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, pad, np.float64, m + 2 * k)
        got, used = run_bf16(H, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.5, c, ldc, s)
        assert used == s
        want = c.copy(order="F")
        oracle.dgemm_ozaki_bf16(ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.5, want, ldc, s)
        assert np.array_equal(got, want), (m, n, k, s)


def test_exact_on_two_digit_integers(H):
    """Integers below 2^12 (two digits each) with 3 slices: C = A B exactly."""
    rng = np.random.default_rng(11)
    m, n, k = 150, 90, 2300
    a = rng.integers(-(2 ** 12) + 1, 2 ** 12, (m, k)).astype(float)
    b = rng.integers(-(2 ** 12) + 1, 2 ** 12, (k, n)).astype(float)
    a[:, 0] = 2 ** 12 - 1
    b[0, :] = -(2 ** 12) + 1
    a, b = np.asfortranarray(a), np.asfortranarray(b)
    c = np.zeros((m, n), order="F")
    got, used = run_bf16(H, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, 3)
    assert used == 3
    assert np.array_equal(got, (a.astype(np.int64) @ b.astype(np.int64)).astype(float))


@pytest.mark.parametrize("ta,tb", TRANS)
def test_auto_slices_native_level_error(H, ta, tb):
    rng = np.random.default_rng(4)
    for (m, n, k) in ((200, 150, 300), (129, 65, 2000)):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 0, np.float64, 3 * k)
        for scaled in (False, True):
            if scaled:
                ra = 10.0 ** rng.uniform(-8, 8, m)
                cb = 10.0 ** rng.uniform(-8, 8, n)
                a = np.asfortranarray(a * (ra[:, None] if ta == "N" else ra[None, :]))
                b = np.asfortranarray(b * (cb[None, :] if tb == "N" else cb[:, None]))
            got, used = run_bf16(H, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, c, ldc, 0)
            want_count = oracle.emulated_choose("bf16", ta, tb, m, n, k, a, lda, b, ldb)[0]
            assert used == want_count and 3 <= used <= 16, (used, want_count)
            r = gemm_bound_ratio(got, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, c, ldc)
            assert r <= 1.0, (m, n, k, scaled, used, r)
            want = c.copy(order="F")
            oracle.dgemm_ozaki_bf16(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, want, ldc, used)
            assert np.array_equal(got, want)


def test_auto_slices_wide_exponent_range(H):
    """Entries u 2^e with e uniform in [-30, 30]: the statistics ask for more
    slices than uniform operands do, and the error stays within the bound.
    (At k = 512 some pair's large entries do not meet at all, the coarse
    product cannot certify it and the native kernel runs; k = 2048 leaves
    every pair dozens of meeting large entries.)"""
    m, n, k = 160, 96, 2048
    a = oracle.fill_pow2(m, k, 31)
    b = oracle.fill_pow2(k, n, 32)
    c = np.zeros((m, n), order="F")
    _, used_uniform = run_bf16(H, "N", "N", m, n, k, 1.0, oracle.fill_uniform(m, k, 1), m,
                               oracle.fill_uniform(k, n, 2), k, 0.0, c, m, 0)
    got, used = run_bf16(H, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, 0)
    assert used > used_uniform
    assert used == oracle.emulated_choose("bf16", "N", "N", m, n, k, a, m, b, k)[0]
    r = gemm_bound_ratio(got, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    assert r <= 1.0, (used, r)


@pytest.mark.parametrize("meets", [False, True])
def test_auto_row_with_a_dominant_entry(H, meets):
    """As for the INT8 method: a row with one entry 2^60 above the rest,
    against a column of op(B) that is zero there (no certified count: the
    native kernel runs) or not (a count is certified and the bound holds)."""
    m, n, k = 100, 90, 513
    a = oracle.fill_uniform(m, k, 41)
    b = oracle.fill_uniform(k, n, 42)
    c = oracle.fill_uniform(m, n, 43)
    a[37, 200] = 2.0 ** 60
    if meets:
        b[200, :] = 0.75
    else:
        b[200, 41] = 0.0
    got, used = run_bf16(H, "N", "N", m, n, k, 1.0, a, m, b, k, 0.5, c, m, 0)
    assert used == oracle.emulated_choose("bf16", "N", "N", m, n, k, a, m, b, k)[0]
    assert (used > 0) == meets
    r = gemm_bound_ratio(got, "N", "N", m, n, k, 1.0, a, m, b, k, 0.5, c, m)
    assert r <= 1.0, (used, r)
    if not meets:
        nat = c.copy(order="F")
        H.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.5, nat, m)
        assert np.array_equal(got, nat)


def test_special_values_fall_back_to_native(H):
    m, n, k = 50, 40, 30
    a, lda, b, ldb, c, ldc = gemm_operands("N", "N", m, n, k, 0, np.float64, 8)
    a[3, 4] = np.nan
    b[5, 6] = -np.inf
    got, used = run_bf16(H, "N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc, 7)
    assert used == 0
    nat = c.copy(order="F")
    H.gemm("N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, nat, ldc)
    assert np.array_equal(np.isnan(got), np.isnan(nat))
    assert np.array_equal(got[np.isfinite(nat)], nat[np.isfinite(nat)])


def test_beta_zero_does_not_read_c_and_quick_returns(H):
    m, n, k = 33, 17, 90
    a, lda, b, ldb, c, ldc = gemm_operands("T", "N", m, n, k, 0, np.float64, 12)
    c[:] = np.nan
    got, used = run_bf16(H, "T", "N", m, n, k, 2.0, a, lda, b, ldb, 0.0, c, ldc, 7)
    assert np.isfinite(got).all() and used == 7
    c2 = np.ones((m, n), order="F")
    anan = np.full_like(a, np.nan)
    got, used = run_bf16(H, "T", "N", m, n, k, 0.0, anan, lda, anan, ldb, 3.0, c2, ldc, 7)
    assert np.all(got == 3.0) and used == 0


def test_invalid_arguments_reported(H):
    with H.ErrorCapture() as cap:
        info, _ = H.dgemm_emulated_bf16_dev("N", "X", 4, 4, 4, 1.0, None, 4, None, 4, 0.0,
                                            None, 4)
    assert info == 2 and cap.calls == [("DGEMM", 2)]
