"""GEMM parity on the GPU, through the C-ABI (dgemm_/sgemm_, cblas_?gemm,
hwyblas_?gemm_dev), against the oracle's double-double reference with
hwyBLAS's componentwise bound. Restates tests/gemm_test.cc:126-212,
tests/semantics_test.cc:93-139, :311-359, :436-514 and
tests/cblas_test.cc:838-889 of the reference.
"""
import itertools

import numpy as np
import pytest

import oracle
from helpers import gemm_bound_ratio, gemm_operands, op_shape

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
TRANS = list(itertools.product("NT", "NT"))


def run_gemm(H, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    c = c.copy(order="F")
    H.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    return c


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ta,tb", TRANS)
def test_small_shapes(H, dtype, ta, tb):
    """Every (ta, tb); m across tile edges; n and k small (gemm_test.cc:126-143)."""
    seed = 11
    worst = 0.0
    for m in (1, 2, 15, 16, 17, 33, 63, 64, 65, 129):
        for n in (1, 7, 16, 25):
            for k in (1, 2, 7, 16, 17):
                a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 0, dtype, seed)
                seed += 3
                got = run_gemm(H, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.5, c, ldc)
                r = gemm_bound_ratio(got, ta, tb, m, n, k, 1.5, a, lda, b, ldb, -0.5, c, ldc)
                assert r <= 1.0, (ta, tb, m, n, k, r)
                worst = max(worst, r)
    print(f"worst bound ratio {worst:.3g}")


BLOCKED = [(64, 64, 64, 0), (150, 35, 300, 3), (333, 17, 513, 1), (31, 4100, 9, 2),
           (257, 129, 1, 0), (1, 300, 600, 5), (300, 300, 300, 7), (520, 260, 1030, 1)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ta,tb", TRANS)
def test_blocked_shapes_padded_ld(H, dtype, ta, tb):
    """Shapes crossing tile and block edges with padded ld (gemm_test.cc:146-166);
    padding rows of C stay untouched (:105-113)."""
    for i, (m, n, k, pad) in enumerate(BLOCKED):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, pad, dtype, 100 + i)
        got = run_gemm(H, ta, tb, m, n, k, 0.75, a, lda, b, ldb, 0.25, c, ldc)
        r = gemm_bound_ratio(got, ta, tb, m, n, k, 0.75, a, lda, b, ldb, 0.25, c, ldc)
        assert r <= 1.0, (m, n, k, r)
        if pad:
            assert np.array_equal(got[m:, :], c[m:, :]), "padding of C modified"


@pytest.mark.parametrize("dtype", DTYPES)
def test_alpha_beta_special_cases(H, dtype):
    """alpha, beta in {0, 1, -1, 0.5}, k in {0, 5, 300}; beta == 0 never reads C
    (C poisoned with NaN) (gemm_test.cc:169-188)."""
    m, n = 37, 29
    for k in (0, 5, 300):
        for alpha in (0.0, 1.0, -1.0, 0.5):
            for beta in (0.0, 1.0, -1.0, 0.5):
                a, lda, b, ldb, c, ldc = gemm_operands("N", "T", m, n, k, 2, dtype, 7 + k)
                if beta == 0.0:
                    c[:m, :] = np.nan
                got = run_gemm(H, "N", "T", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
                assert not np.isnan(got[:m, :n]).any() or beta != 0.0
                r = gemm_bound_ratio(got, "N", "T", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
                assert r <= 1.0, (k, alpha, beta, r)
                assert np.array_equal(got[m:, :], c[m:, :])
                if (k == 0 or alpha == 0.0) and beta == 1.0:
                    assert np.array_equal(got, c)


def test_empty_dimensions_leave_c_untouched(H):
    c = np.full((4, 3), 3.0, order="F")
    a = np.ones((4, 4), order="F")
    for m, n in ((0, 3), (4, 0)):
        got = c.copy(order="F")
        H.gemm("N", "N", m, n, 4, 1.0, a, 4, a, 4, 0.0, got, 4)
        assert np.array_equal(got, c)


@pytest.mark.parametrize("dtype", DTYPES)
def test_nan_inf_reach_one_row_or_column(H, dtype):
    """A NaN/Inf in op(A)(i0, l) reaches row i0 only; in op(B)(l, j0) column j0
    only; in C (beta = 1) one element (semantics_test.cc:93-139)."""
    for size in (1, 7, 33, 130):
        m, n, k = size, size + 3, size + 1
        i0, j0, l0 = m // 2, n - 1, k // 3
        for ta, tb in TRANS:
            for where in range(3):
                for value in (np.nan, np.inf):
                    a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 0, dtype, 5 * size)
                    if where == 0:
                        if ta == "N":
                            a[i0, l0] = value
                        else:
                            a[l0, i0] = value
                    elif where == 1:
                        if tb == "N":
                            b[l0, j0] = value
                        else:
                            b[j0, l0] = value
                    else:
                        c[i0, j0] = value
                    beta = 1.0 if where == 2 else 0.0
                    got = run_gemm(H, ta, tb, m, n, k, 0.5, a, lda, b, ldb, beta, c, ldc)
                    hit = np.zeros((m, n), bool)
                    if where == 0:
                        hit[i0, :] = True
                    elif where == 1:
                        hit[:, j0] = True
                    else:
                        hit[i0, j0] = True
                    g = got[:m, :n]
                    special = np.isnan(g) if np.isnan(value) else ~np.isfinite(g)
                    assert special[hit].all(), (size, ta, tb, where, value)
                    assert np.isfinite(g[~hit]).all(), (size, ta, tb, where, value)


@pytest.mark.parametrize("dtype", DTYPES)
def test_subnormals_are_exact(H, dtype):
    """gemm of s = 2^-530 (double) or 2^-70 (float) values gives exactly n s^2,
    a subnormal (semantics_test.cc:311-342)."""
    s = np.ldexp(1.0, -530 if dtype == np.float64 else -70)
    s = dtype(s)
    for n in (1, 5, 64, 200):
        expected = dtype(n) * (s * s)
        assert expected != 0 and abs(expected) < np.finfo(dtype).tiny
        a = np.full((n, n), s, dtype=dtype, order="F")
        c = np.ones((n, n), dtype=dtype, order="F")
        H.gemm("N", "N", n, n, n, 1.0, a, n, a, n, 0.0, c, n)
        assert np.all(c == expected), (n, c.ravel()[:4], expected)


@pytest.mark.parametrize("dtype", DTYPES)
def test_integer_valued_operands_exact(H, dtype):
    """Small integers: every partial sum is exact, so any order gives the exact
    product (a closed form independent of the oracle)."""
    rng = np.random.default_rng(3)
    for ta, tb in TRANS:
        m, n, k = 97, 71, 300
        a = np.asfortranarray(rng.integers(-8, 9, op_shape(ta, m, k)).astype(dtype))
        b = np.asfortranarray(rng.integers(-8, 9, op_shape(tb, k, n)).astype(dtype))
        c = np.asfortranarray(rng.integers(-50, 50, (m, n)).astype(dtype))
        A = a if ta == "N" else a.T
        B = b if tb == "N" else b.T
        want = 2.0 * (A.astype(np.float64) @ B) + 3.0 * c
        got = run_gemm(H, ta, tb, m, n, k, 2.0, a, a.shape[0], b, b.shape[0], 3.0, c, m)
        assert np.array_equal(got, want.astype(dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_reproducible_repeated_and_device_offsets(H, dtype):
    """Bitwise identical results when repeated and with operands at every
    element offset of a 64-byte line (semantics_test.cc:436-514)."""
    item = np.dtype(dtype).itemsize
    for n in (37, 150):
        a, lda, b, ldb, c, ldc = gemm_operands("N", "T", n, n, n, 0, dtype, 77)
        ref = run_gemm(H, "N", "T", n, n, n, 0.75, a, lda, b, ldb, -0.5, c, ldc)
        assert np.array_equal(ref, run_gemm(H, "N", "T", n, n, n, 0.75, a, lda, b, ldb, -0.5, c, ldc))
        for off in range(1, 64 // item):
            bufs = []
            for x in (a, b, c):
                d = H.DeviceArray(x.shape[0] * x.shape[1] + off, 1, dtype)
                flat = np.concatenate([np.zeros(off, dtype), x.ravel(order="F")])
                H.lib().hwyblas_amd_memcpy(d.ptr, flat.ctypes.data, flat.nbytes, 1)
                bufs.append(d)
            pa, pb, pc = (d.ptr + off * item for d in bufs)
            rc = H.lib().hwyblas_dgemm_dev if dtype == np.float64 else H.lib().hwyblas_sgemm_dev
            assert rc(b"N", b"T", n, n, n, 0.75, pa, lda, pb, ldb, -0.5, pc, ldc, None) == 0
            H.synchronize()
            out = bufs[2].to_host().ravel(order="F")[off:].reshape((ldc, n), order="F")
            assert np.array_equal(out, ref), off


@pytest.mark.parametrize("dtype", DTYPES)
def test_cblas_layouts(H, dtype):
    """Column-major CBLAS is identical to the Fortran routine; row-major is the
    transposed problem (cblas_test.cc:838-889)."""
    m, n, k = 45, 67, 83
    for ta, tb in TRANS:
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 1, dtype, 31)
        f = run_gemm(H, ta, tb, m, n, k, 1.25, a, lda, b, ldb, 0.5, c, ldc)
        cc = c.copy(order="F")
        T = {"N": H.NoTrans, "T": H.Trans}
        H.cblas_gemm(H.ColMajor, T[ta], T[tb], m, n, k, 1.25, a, lda, b, ldb, 0.5, cc, ldc)
        assert np.array_equal(cc, f)
        # Row-major: the same matrices stored row-major are the transposes.
        ar = np.ascontiguousarray(a[: op_shape(ta, m, k)[0], :])
        br = np.ascontiguousarray(b[: op_shape(tb, k, n)[0], :])
        cr = np.ascontiguousarray(c[:m, :])
        # Row-major arrays with ld = row length; op() of the row-major matrix
        # equals op() of the column-major one stored in ar/br.
        cr2 = cr.copy()
        H.cblas_gemm(H.RowMajor, T[ta], T[tb], m, n, k, 1.25, ar, ar.shape[1], br, br.shape[1],
                     0.5, cr2, n)
        A = ar if ta == "N" else ar.T
        B = br if tb == "N" else br.T
        want = 1.25 * (A.astype(np.float64) @ B) + 0.5 * cr
        scale = 1.25 * (np.abs(A).astype(np.float64) @ np.abs(B)) + 0.5 * np.abs(cr)
        tol = 4 * (k + 2) * np.finfo(dtype).eps
        assert np.all(np.abs(cr2 - want) <= tol * scale)


@pytest.mark.parametrize("dtype", DTYPES)
def test_device_entry_and_fill_match_oracle_generator(H, dtype):
    """The device fill produces the oracle generator's values bit for bit, and
    the device entry point agrees with the host entry point."""
    m, n, k = 300, 200, 250
    d = H.DeviceArray(m + 3, k, dtype)
    d.fill_uniform(m, k, 1234)
    H.synchronize()
    host = oracle.fill_uniform(m, k, 1234, dtype=dtype, ld=m + 3)
    assert np.array_equal(d.to_host()[:m], host[:m])
    a, lda, b, ldb, c, ldc = gemm_operands("T", "N", m, n, k, 0, dtype, 9)
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    assert H.gemm_dev("T", "N", m, n, k, -1.5, da, lda, db, ldb, 2.0, dc, ldc, dtype=dtype) == 0
    H.synchronize()
    host_res = run_gemm(H, "T", "N", m, n, k, -1.5, a, lda, b, ldb, 2.0, c, ldc)
    assert np.array_equal(dc.to_host(), host_res)
    # dgemm_ on device pointers runs in place too.
    dc2 = H.DeviceArray.from_host(c)
    H.gemm("T", "N", m, n, k, -1.5, da, lda, db, ldb, 2.0, dc2, ldc, dtype=dtype)
    assert np.array_equal(dc2.to_host(), host_res)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ta,tb", TRANS)
def test_large_square_sampled(H, dtype, ta, tb):
    """m = n = k = 2048, operands generated on the device, 512 sampled entries
    checked against the double-double reference."""
    n = 2048
    a = H.DeviceArray(n, n, dtype)
    b = H.DeviceArray(n, n, dtype)
    c = H.DeviceArray(n, n, dtype)
    a.fill_uniform(n, n, 1)
    b.fill_uniform(n, n, 2)
    c.fill_uniform(n, n, 3)
    assert H.gemm_dev(ta, tb, n, n, n, 1.0, a, n, b, n, 0.5, c, n, dtype=dtype) == 0
    got = c.to_host()
    ha = oracle.fill_uniform(n, n, 1, dtype=dtype)
    hb = oracle.fill_uniform(n, n, 2, dtype=dtype)
    hc = oracle.fill_uniform(n, n, 3, dtype=dtype)
    rng = np.random.default_rng(0)
    rows = rng.integers(0, n, 512)
    cols = rng.integers(0, n, 512)
    exp, bnd = oracle.gemm_ref_entries(ta, tb, n, 1.0, ha, n, hb, n, 0.5, hc, n, rows, cols)
    err = np.abs(got[rows, cols] - exp) / bnd
    tol = 4 * (n + 2) * np.finfo(dtype).eps
    assert np.max(err) <= tol, np.max(err) / tol


@pytest.mark.parametrize("dtype", DTYPES)
def test_every_tile_config_gives_identical_bits(H, dtype):
    """Every tile configuration runs the same k order, so all of them must
    return bitwise the same C, on shapes with edge tiles in m, n and k."""
    nconf = len(H.gemm_configs())
    for (ta, tb), (m, n, k) in zip(TRANS * 2, [(129, 67, 33), (300, 257, 70), (64, 64, 16),
                                               (1, 513, 260), (520, 3, 9), (200, 200, 300),
                                               (17, 130, 1), (256, 128, 64)]):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 1, dtype, m + n + k)
        da, db = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b)
        ref = None
        for cfg in range(nconf):
            dc = H.DeviceArray.from_host(c)
            assert H.gemm_dev_cfg(cfg, ta, tb, m, n, k, 0.5, da, lda, db, ldb, -1.0, dc, ldc,
                                  dtype=dtype) == 0
            got = dc.to_host()
            if ref is None:
                ref = got
                r = gemm_bound_ratio(got, ta, tb, m, n, k, 0.5, a, lda, b, ldb, -1.0, c, ldc)
                assert r <= 1.0
            assert np.array_equal(got, ref), (cfg, ta, tb, m, n, k)
            assert np.array_equal(got[m:], c[m:])


@pytest.mark.parametrize("dtype", DTYPES)
def test_dma_configs_cover_full_tiles_strips_and_fallbacks(H, dtype):
    """The DMA-fed configurations take the full tiles of calls with k a
    multiple of 16, 16-byte aligned operands and leading dimensions of whole
    16-byte chunks, and hand the bottom and right strips (and every other
    call) to the plain kernel. Each case equals a plain configuration bit for
    bit and stays within the GEMM bound; C's padding rows are untouched."""
    L = H.lib()
    dma = [c for c in range(L.hwyblas_gemm_config_count()) if L.hwyblas_gemm_config_stages(c) > 0]
    assert dma and L.hwyblas_gemm_config_stages(0) == 0
    assert L.hwyblas_gemm_config_stages(L.hwyblas_gemm_config_count()) == -1
    item = np.dtype(dtype).itemsize
    cases = [  # (m, n, k, pad, byte offset of A's first element)
        (300, 520, 48, 0, 0),    # full tiles plus both strips
        (256, 256, 64, 4, 0),    # whole tiles only, padded leading dimensions
        (1000, 600, 160, 0, 0),  # several tiles and strips
        (300, 520, 40, 0, 0),    # k not a multiple of 16: the plain kernel
        (300, 520, 48, 1, 0),    # odd leading dimensions: the plain kernel
        (300, 520, 48, 0, item),  # A one element off 16 bytes: the plain kernel
        (100, 90, 32, 0, 0),     # smaller than one tile: the plain kernel
    ]
    for (ta, tb), (m, n, k, pad, off) in itertools.product(TRANS, cases):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, pad, dtype, m + 3 * n + k)
        abig = np.zeros(a.size + 4, dtype=dtype)
        da = H.DeviceArray.from_host(abig.reshape(-1, 1))
        H.lib().hwyblas_amd_memcpy(da.ptr + off, a.ctypes.data, a.nbytes, 1)
        H.synchronize()
        db = H.DeviceArray.from_host(b)
        ref = None
        for cfg in [0] + dma:
            dc = H.DeviceArray.from_host(c)
            assert H.gemm_dev_cfg(cfg, ta, tb, m, n, k, 0.5, da.ptr + off, lda, db, ldb, -1.0,
                                  dc, ldc, dtype=dtype) == 0
            got = dc.to_host()
            if ref is None:
                ref = got
                r = gemm_bound_ratio(got, ta, tb, m, n, k, 0.5, a, lda, b, ldb, -1.0, c, ldc)
                assert r <= 1.0, (ta, tb, m, n, k, pad, off, r)
            assert np.array_equal(got, ref), (cfg, ta, tb, m, n, k, pad, off)
            assert np.array_equal(got[m:], c[m:])


@pytest.mark.parametrize("ta,tb,beta,pad", [("N", "N", 0.0, 0), ("T", "T", 0.5, 3),
                                            ("N", "T", 0.5, 0), ("T", "N", 0.0, 2)])
def test_host_operand_pipeline_equals_device_call(H, ta, tb, beta, pad):
    """dgemm_ on host operands large enough for the pipelined path (blocks of
    C's columns copied in, multiplied and copied out on three streams) gives
    bit for bit the device-operand call's C, and leaves C's padding rows
    alone."""
    m, n, k = 4096, 4608, 4096
    ra, ca = op_shape(ta, m, k)
    rb, cb = op_shape(tb, k, n)
    lda, ldb, ldc = ra + pad, rb + pad, m + pad
    arrs = []
    for rows, cols, ld, seed in ((ra, ca, lda, 1), (rb, cb, ldb, 2), (m, n, ldc, 3)):
        d = H.DeviceArray(ld, cols)
        d.fill_uniform(ld, cols, seed)
        arrs.append(d)
    H.synchronize()
    a, b, c0 = (d.to_host() for d in arrs)
    c0[m:] = 7.25
    dc = H.DeviceArray.from_host(c0)
    H.gemm(ta, tb, m, n, k, 1.0, arrs[0], lda, arrs[1], ldb, beta, dc, ldc)
    H.synchronize()
    want = dc.to_host()
    c = c0.copy(order="F")
    H.gemm(ta, tb, m, n, k, 1.0, a, lda, b, ldb, beta, c, ldc)
    assert np.array_equal(c, want)
    assert np.all(c[m:] == 7.25)


def test_default_config_takes_the_dma_kernel_for_large_products(H):
    L = H.lib()
    for dtype_bytes in (8, 4):
        cfg = L.hwyblas_gemm_default_config(dtype_bytes, b"N", b"N", 8192, 8192, 8192)
        assert L.hwyblas_gemm_config_stages(cfg) > 0
        small = L.hwyblas_gemm_default_config(dtype_bytes, b"N", b"N", 256, 256, 256)
        assert L.hwyblas_gemm_config_stages(small) == 0
    # FP64 by the per-CU time model: 2048^3 (256 tiles, one per CU) takes the
    # DMA kernel; 3072^3 (576 tiles, three on the busiest CU) and a k = 256
    # update of 2048^2 take the finer plain tiles.
    stages = lambda *mnk: L.hwyblas_gemm_config_stages(
        L.hwyblas_gemm_default_config(8, b"N", b"N", *mnk))
    assert stages(2048, 2048, 2048) > 0 and stages(512, 8192, 4096) > 0
    assert stages(3072, 3072, 3072) == 0 and stages(2048, 2048, 256) == 0


def test_device_pow2_fill_and_scaling_match_oracle(H):
    """The wide-exponent test data (+-u 2^e) and the row/column scaling the
    benchmark uses are bit for bit those the oracle regenerates on the host."""
    m, n = 257, 130
    d = H.DeviceArray(m + 1, n)
    H.lib().hwyblas_amd_fill_pow2(d.ptr, m, n, m + 1, 77, -30, 30, None)
    rs = np.power(10.0, np.random.default_rng(1).uniform(-8, 8, m))
    cs = np.power(10.0, np.random.default_rng(2).uniform(-8, 8, n))
    drs, dcs = H.DeviceArray.from_host(rs.reshape(-1, 1)), H.DeviceArray.from_host(cs.reshape(-1, 1))
    H.synchronize()
    host = oracle.fill_pow2(m, n, 77)
    got = d.to_host()[:m]
    assert np.array_equal(got, host)
    e = np.frexp(np.abs(host))[1] - 1
    assert e.min() == -30 and e.max() == 30 and (host < 0).any() and (host > 0).any()
    H.lib().hwyblas_amd_scale_rows_cols(d.ptr, m, n, m + 1, drs.ptr, dcs.ptr, None)
    H.synchronize()
    assert np.array_equal(d.to_host()[:m], (host * rs[:, None]) * cs[None, :])


# 32 ranges of the 32 x 32 tiles; 4 of the DMA-fed 128 x 128 tiles; 9 ranges,
# the last one shorter; 5 ranges of one row.
SPLIT_SHAPES = [(64, 64, 32768), (128, 256, 4096), (100, 37, 5000), (1, 300, 3000)]


@pytest.mark.parametrize("dtype", DTYPES)
def test_skinny_long_k_products_split_k(H, dtype):
    """Products with too few tiles of C to fill the device and k >= 1024 run
    split over k (the ranges at once, the partials added in order, then
    hwyBLAS's Store rules): within the componentwise GEMM bound, for every
    transposition and beta form, padding rows untouched; the host entry, the
    device entry and a second call give the same bits."""
    for (m, n, k), (ta, tb), beta in zip(SPLIT_SHAPES, TRANS, (0.0, 1.0, -0.5, 2.0)):
        a, lda, b, ldb, c, ldc = gemm_operands(ta, tb, m, n, k, 3, dtype, m + n + k)
        got = run_gemm(H, ta, tb, m, n, k, 0.75, a, lda, b, ldb, beta, c, ldc)
        r = gemm_bound_ratio(got, ta, tb, m, n, k, 0.75, a, lda, b, ldb, beta, c, ldc)
        assert r <= 1.0, (m, n, k, ta, tb, beta, r)
        assert np.array_equal(got[m:], c[m:])
        assert np.array_equal(run_gemm(H, ta, tb, m, n, k, 0.75, a, lda, b, ldb, beta, c, ldc),
                              got)
        da, db, dc = (H.DeviceArray.from_host(x) for x in (a, b, c))
        assert H.gemm_dev(ta, tb, m, n, k, 0.75, da, lda, db, ldb, beta, dc, ldc,
                          dtype=dtype) == 0
        H.synchronize()
        assert np.array_equal(dc.to_host(), got)


def test_split_k_partials_shared_across_streams(H):
    """One thread's device calls on two streams share its split-k partials:
    the second stream waits for the first call's use of them, so both
    results equal the synchronous calls'."""
    m, n, k = 64, 96, 20000
    L = H.lib()
    s1, s2 = L.hwyblas_amd_stream_create(), L.hwyblas_amd_stream_create()
    try:
        want, dcs = [], []
        for seed in (1, 2):
            a, lda, b, ldb, c, ldc = gemm_operands("N", "T", m, n, k, 0, np.float64, seed)
            want.append(run_gemm(H, "N", "T", m, n, k, 1.0, a, lda, b, ldb, 0.0, c, ldc))
            dcs.append((H.DeviceArray.from_host(a), H.DeviceArray.from_host(b),
                        H.DeviceArray.from_host(c), lda, ldb, ldc))
        H.synchronize()
        for (da, db, dc, lda, ldb, ldc), s in zip(dcs, (s1, s2)):
            assert H.gemm_dev("N", "T", m, n, k, 1.0, da, lda, db, ldb, 0.0, dc, ldc,
                              stream=s) == 0
        L.hwyblas_amd_stream_synchronize(s1)
        L.hwyblas_amd_stream_synchronize(s2)
        for (_, _, dc, *_), w in zip(dcs, want):
            assert np.array_equal(dc.to_host(), w)
    finally:
        L.hwyblas_amd_stream_destroy(s1)
        L.hwyblas_amd_stream_destroy(s2)
