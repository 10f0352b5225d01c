"""Emulated DGEMM (INT8 Ozaki scheme II) on the GPU through the C-ABI.

Pinned three ways: (1) bit for bit against the oracle's CPU restatement of
the same algorithm (every step but the last FP64 conversions is exact integer
arithmetic, and those conversions are the same operations), which checks the
slicing, the INT8 MFMA operand layout, the epilogue reduction and the CRT
reconstruction; (2) the count chosen on the device equals the oracle's
restatement of the choice (exact statistics of every pair, the coarse
product, the rule), including its native fallback on operands whose bound
cannot be certified; (3) against the double-double reference with the native
DGEMM componentwise bound 4 (k + 2) u, on every entry, with the count chosen
automatically, for uniform operands, rows/columns scaled by 10^s, entries
u 2^e (e in [-30, 30]) and rows with one entry 2^60 above the rest.
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
            want_count = oracle.emulated_choose("int8", ta, tb, m, n, k, a, lda, b, ldb)[0]
            assert used == want_count and 8 <= used <= 24, (used, want_count)
            r = gemm_bound_ratio(got, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, c, ldc)
            assert r <= 1.0, (m, n, k, scaled, used, r)
            want = c.copy(order="F")
            oracle.dgemm_ozaki(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.5, want, ldc, used)
            assert np.array_equal(got, want)


def native_result(H, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    assert H.gemm_dev(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc, ldc,
                      dtype=np.float64) == 0
    H.synchronize()
    return dc.to_host()


def check_auto(H, ta, tb, m, n, k, a, lda, b, ldb, c, ldc, alpha=1.0, beta=0.5):
    """Auto count: equal to the oracle's choice, the bits of the oracle's
    restatement with that count (or of the native kernel when it is 0), and
    within the native bound on every entry."""
    got, used = run_emulated(H, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, 0)
    want_count = oracle.emulated_choose("int8", ta, tb, m, n, k, a, lda, b, ldb)[0]
    assert used == want_count, (used, want_count)
    if used > 0:
        want = c.copy(order="F")
        oracle.dgemm_ozaki(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, want, ldc, used)
    else:
        want = native_result(H, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert np.array_equal(got, want)
    r = gemm_bound_ratio(got, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    assert r <= 1.0, (m, n, k, used, r)
    return used


@pytest.mark.parametrize("ta,tb", TRANS)
def test_auto_wide_exponent_range(H, ta, tb):
    """Entries +-u 2^e, e uniform in [-30, 30] (the wide-range distribution)."""
    m, n, k = 192, 160, 1024
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    a = oracle.fill_pow2(ra, ca, 31)
    b = oracle.fill_pow2(rb, cb, 32)
    c = oracle.fill_uniform(m, n, 33)
    used = check_auto(H, ta, tb, m, n, k, a, ra, b, rb, c, m)
    assert used >= 14


@pytest.mark.parametrize("meets", [False, True])
def test_auto_row_with_a_dominant_entry(H, meets):
    """A row of op(A) with one entry 2^60 above the rest (rows 0..31 and the
    columns of op(B) on any sampling lattice are left alone: the row is 37).
    meets=False: the column of op(B) is zero where the large entry sits, so
    C(37, 41) is small while the row's fixed-point grid is 2^60 coarser: no
    count can be certified and the native kernel runs. meets=True: that
    entry of B is non-zero, the pair's |A||B| is dominated by the large
    product and a count is certified."""
    m, n, k = 100, 90, 513
    a = oracle.fill_uniform(m, k, 41)
    b = oracle.fill_uniform(k, n, 42)
    c = oracle.fill_uniform(m, n, 43)
    a[37, 200] = 2.0 ** 60
    if meets:
        b[200, :] = 0.75
    else:
        b[200, 41] = 0.0
    used = check_auto(H, "N", "N", m, n, k, a, m, b, k, c, m)
    assert (used > 0) == meets


def test_auto_block_sparse_falls_back_to_native(H):
    """Disjoint supports: |A||B|(i, j) = 0 for some pairs, which the coarse
    product cannot certify, so the native kernel runs (count 0)."""
    m, n, k = 128, 96, 256
    a = oracle.fill_uniform(m, k, 51)
    b = oracle.fill_uniform(k, n, 52)
    a[:64, 128:] = 0.0
    a[64:, :128] = 0.0
    b[:128, 48:] = 0.0
    b[128:, :48] = 0.0
    c = oracle.fill_uniform(m, n, 53)
    assert check_auto(H, "N", "N", m, n, k, a, m, b, k, c, m) == 0


def test_auto_zero_operand_is_exact(H):
    m, n, k = 64, 48, 80
    a = np.zeros((m, k), order="F")
    b = oracle.fill_uniform(k, n, 61)
    c = oracle.fill_uniform(m, n, 62)
    got, used = run_emulated(H, "N", "N", m, n, k, 1.0, a, m, b, k, 0.5, c, m, 0)
    assert used == 2 and np.array_equal(got, 0.5 * c)


def test_asynchronous_call_and_graph_replays(H):
    """Without a count pointer the call does not synchronize; captured in a
    HIP graph (after an eager call has grown the workspace) it replays with
    the decision made again on the device for the operands of each replay:
    uniform operands (14 moduli) and then a NaN (the native kernel)."""
    m, n, k = 300, 200, 700
    a = oracle.fill_uniform(m, k, 71)
    b = oracle.fill_uniform(k, n, 72)
    c = oracle.fill_uniform(m, n, 73)
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    dc0 = H.DeviceArray.from_host(c)
    L = H.lib()
    nbytes = c.nbytes
    info, used = H.dgemm_emulated_dev("N", "N", m, n, k, 1.5, da, m, db, k, -0.5, dc, m, 0)
    H.synchronize()
    assert info == 0 and used == 14
    eager = dc.to_host()
    want = c.copy(order="F")
    oracle.dgemm_ozaki("N", "N", m, n, k, 1.5, a, m, b, k, -0.5, want, m, 14)
    assert np.array_equal(eager, want)
    stream = L.hwyblas_amd_stream_create()
    try:
        L.hwyblas_amd_memcpy(dc.ptr, dc0.ptr, nbytes, 3)
        H.synchronize()
        info, used = H.dgemm_emulated_dev("N", "N", m, n, k, 1.5, da, m, db, k, -0.5, dc, m, 0,
                                          stream=stream, report=False)
        assert info == 0 and used is None
        L.hwyblas_amd_stream_synchronize(stream)
        assert np.array_equal(dc.to_host(), eager)
        L.hwyblas_amd_capture_begin(stream)
        H.dgemm_emulated_dev("N", "N", m, n, k, 1.5, da, m, db, k, -0.5, dc, m, 0,
                             stream=stream, report=False)
        graph = L.hwyblas_amd_capture_end(stream)
        try:
            for replay in range(3):
                L.hwyblas_amd_memcpy(dc.ptr, dc0.ptr, nbytes, 3)
                scratch = H.DeviceArray(512, 8)
                scratch.fill_uniform(512, 8, 90 + replay)
                H.synchronize()
                L.hwyblas_amd_graph_launch(graph, stream)
                L.hwyblas_amd_stream_synchronize(stream)
                assert np.array_equal(dc.to_host(), eager), replay
            # New operand contents, same pointers: a NaN sends the replay to
            # the native kernel.
            a2 = a.copy(order="F")
            a2[5, 6] = np.nan
            L.hwyblas_amd_memcpy(da.ptr, a2.ctypes.data, a2.nbytes, 1)
            L.hwyblas_amd_memcpy(dc.ptr, dc0.ptr, nbytes, 3)
            H.synchronize()
            L.hwyblas_amd_graph_launch(graph, stream)
            L.hwyblas_amd_stream_synchronize(stream)
            nat = native_result(H, "N", "N", m, n, k, 1.5, a2, m, b, k, -0.5, c, m)
            got = dc.to_host()
            assert np.array_equal(np.isnan(got), np.isnan(nat))
            assert np.array_equal(got[~np.isnan(nat)], nat[~np.isnan(nat)])
        finally:
            L.hwyblas_amd_graph_destroy(graph)
    finally:
        L.hwyblas_amd_stream_destroy(stream)


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


def test_dgemm_mode_routes_large_products_to_the_emulation(H):
    """hwyblas_set_dgemm_mode(1): dgemm_ on host operands and
    hwyblas_dgemm_dev give the emulated DGEMM's bits (count chosen on the
    device) for m, n >= 2048, k >= 1024, and the native kernel's below."""
    L = H.lib()
    prev = L.hwyblas_set_dgemm_mode(1)
    try:
        for (m, n, k, emulated) in ((2048, 2304, 1024, True), (2048, 2048, 512, False)):
            a = oracle.fill_uniform(m, k, 81)
            b = oracle.fill_uniform(k, n, 82)
            c = oracle.fill_uniform(m, n, 83)
            host = c.copy(order="F")
            H.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.5, host, m)
            da, db, dc = (H.DeviceArray.from_host(x) for x in (a, b, c))
            assert H.gemm_dev("N", "N", m, n, k, 1.0, da, m, db, k, 0.5, dc, m,
                              dtype=np.float64) == 0
            H.synchronize()
            dev = dc.to_host()
            ref = H.DeviceArray.from_host(c)
            if emulated:
                info, used = H.dgemm_emulated_dev("N", "N", m, n, k, 1.0, da, m, db, k, 0.5,
                                                  ref, m)
                assert info == 0 and used == 14
            else:
                L.hwyblas_set_dgemm_mode(0)
                assert H.gemm_dev("N", "N", m, n, k, 1.0, da, m, db, k, 0.5, ref, m,
                                  dtype=np.float64) == 0
                L.hwyblas_set_dgemm_mode(1)
            H.synchronize()
            want = ref.to_host()
            assert np.array_equal(host, want) and np.array_equal(dev, want), (m, n, k)
    finally:
        L.hwyblas_set_dgemm_mode(prev)
