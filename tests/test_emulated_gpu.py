"""Emulated DGEMM (INT8 Ozaki scheme II) on the GPU through the C-ABI.

Pinned two ways: (1) bit for bit against the oracle's CPU restatement of the
same algorithm (every step but the last FP64 conversions is exact integer
arithmetic, and those conversions are the same operations), which checks the
slicing, the INT8 MFMA operand layout, the epilogue reduction and the CRT
reconstruction; (2) against the double-double reference with the native
DGEMM componentwise bound 4 (k + 2) u, with the moduli chosen automatically.
"""
import itertools

import numpy as np
import pytest

import oracle
from helpers import gemm_bound_ratio, gemm_operands

pytestmark = pytest.mark.gpu
TRANS = list(itertools.product("NT", "NT"))


def run_emulated(H, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmod):
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    info, used = H.dgemm_emulated_dev(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc, ldc,
                                      nmod)
    assert info == 0
    H.synchronize()
    return dc.to_host(), used


@pytest.mark.parametrize("ta,tb", TRANS)
def test_bitwise_equal_to_restatement(H, ta, tb):
    for (m, n, k, pad, nmod) in ((37, 29, 100, 1, 14), (300, 257, 70, 0, 16), (1, 513, 65, 2, 12),
                                 (256, 256, 256, 0, 9), (70, 3, 1, 1, 20)):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, pad, np.float64, m + k)
        got, used = run_emulated(H, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.5, c, ldc, nmod)
        assert used == nmod
        want = c.copy(order="F")
        oracle.dgemm_ozaki(ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.5, want, ldc, nmod)
        assert np.array_equal(got, want), (m, n, k, nmod)


@pytest.mark.parametrize("ta,tb", TRANS)
def test_auto_moduli_native_level_error(H, ta, tb):
    """With the moduli chosen from the operands, the error is within the
    native DGEMM bound, also for rows/columns scaled by 10^s (configs
    distribution (c))."""
    rng = np.random.default_rng(4)
    for (m, n, k) in ((200, 150, 300), (129, 65, 2000)):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 0, np.float64, 3 * k)
        for scaled in (False, True):
            if scaled:
                ra = 10.0 ** rng.uniform(-8, 8, m)
                cb = 10.0 ** rng.uniform(-8, 8, n)
                if ta == "N":
                    a = np.asfortranarray(a * ra[:, None])
                else:
                    a = np.asfortranarray(a * ra[None, :])
                if tb == "N":
                    b = np.asfortranarray(b * cb[None, :])
                else:
                    b = np.asfortranarray(b * cb[:, None])
            got, used = run_emulated(H, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, c, ldc, 0)
            assert 8 <= used <= 24
            r = gemm_bound_ratio(got, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, c, ldc)
            assert r <= 1.0, (m, n, k, scaled, used, r)


def test_special_values_fall_back_to_native(H):
    m, n, k = 50, 40, 30
    a, lda, b, ldb, c, ldc = gemm_operands("N", "N", m, n, k, 0, np.float64, 8)
    a[3, 4] = np.nan
    b[5, 6] = np.inf
    got, used = run_emulated(H, "N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc, 14)
    assert used == 0
    nat = c.copy(order="F")
    H.gemm("N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, nat, ldc)
    assert np.array_equal(np.isnan(got), np.isnan(nat))
    assert np.array_equal(got[np.isfinite(nat)], nat[np.isfinite(nat)])


def test_beta_zero_does_not_read_c_and_quick_returns(H):
    m, n, k = 33, 17, 90
    a, lda, b, ldb, c, ldc = gemm_operands("T", "N", m, n, k, 0, np.float64, 12)
    c[:] = np.nan
    got, used = run_emulated(H, "T", "N", m, n, k, 2.0, a, lda, b, ldb, 0.0, c, ldc, 14)
    assert np.isfinite(got).all() and used == 14
    # alpha == 0: C := beta C without reading A or B.
    c2 = np.ones((m, n), order="F")
    anan = np.full_like(a, np.nan)
    got, used = run_emulated(H, "T", "N", m, n, k, 0.0, anan, lda, anan, ldb, 3.0, c2, ldc, 14)
    assert np.all(got == 3.0) and used == 0


def test_invalid_arguments_reported(H):
    a = np.ones((4, 4), order="F")
    with H.ErrorCapture() as cap:
        info, _ = H.dgemm_emulated_dev("N", "N", 4, 4, 4, 1.0, None, 3, None, 4, 0.0, None, 4)
    assert info == 8 and cap.calls == [("DGEMM", 8)]
