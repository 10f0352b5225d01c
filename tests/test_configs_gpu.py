"""Every benchmark workload at its stated size, on the GPU through the C-ABI.

The workloads (m = n = k unless said otherwise, column-major unless said
otherwise, entries uniform[-1, 1] from the generator shared by the device
fill and the oracle):
 [0] DGEMM NN 512^3, alpha 1.5, beta 0.5 (A, B, C from seeds 0, 1, 2), and
     DTRSM L/L/N/N 512 x 512 with diagonal 512 + U[0, 1], alpha 1; checked
     in full.
 [1] SGEMM and DGEMM at 4096, 8192 and 16384, every transa/transb, column-
     and row-major (CBLAS), alpha 1, beta 0: 4 full rows, 4 full columns and
     512 random entries against the double-double reference with hwyBLAS's
     componentwise bound 4 (k + 2) u (tests/gemm_test.cc:72-101). DTRSM
     8192 x 8192 for every side/uplo/trans with diagonal n + U[0, 1] and the
     other triangle NaN: the residual bound 8 (ka + 2) u
     (tests/level3_test.cc:482-520) on 16 right-hand sides.
 [2] The emulated DGEMM at 16384^3 on uniform operands (a), +-u 2^e with
     e in [-30, 30] (b), and rows of A / columns of B scaled by 10^s,
     s in [-8, 8] (c), INT8 moduli and BF16 slices chosen on the device: the
     same entries within the bound, and the INT8 method's componentwise error
     within 2x the native kernel's on them.
 [3] 10,000 uneven GEMMs (FP64 and FP32; m, n, k log-uniform in [8, 512],
     random transpositions) and 10,000 DTRSMs (L/L/N/N, m in [8, 256], n in
     [1, 64], diagonal m + U[0, 1]), one graph launch each: every problem
     checked against the double-double reference.
 [4] The LU solve at n = 16384 with 16 right-hand sides, A = Q1 diag(sigma)
     Q2^T (sigma geometric from 1 to 10^kappa, kappa in {2, 6, 10}; Q1, Q2
     random orthogonal): with emulated trailing updates, gerfs's berr below
     4 (n + 1) u (tests/lapack/linear_test.cc:498-560), and the componentwise
     and normwise backward errors recomputed on the host in double-double.
"""
import itertools
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from helpers import gemm_bound_ratio

pytestmark = pytest.mark.gpu
TRANS = list(itertools.product("NT", "NT"))
U64 = 2.0 ** -53
CBLAS = {"row": 101, "col": 102, "N": 111, "T": 112}


def eps_of(dtype):
    return float(np.finfo(dtype).eps)


def check_entries(n):
    rng = np.random.default_rng(n)
    return (rng.choice(n, 4, replace=False), rng.choice(n, 4, replace=False),
            rng.integers(0, n, 512), rng.integers(0, n, 512))


def entries_of(c, R, Cc, sr, sc):
    return np.concatenate([c[R, :].ravel(), c[:, Cc].T.ravel(), c[sr, sc]])


def reference(n, x, xt, tx, y, yt, ty, R, Cc, sr, sc, alpha=1.0):
    """Double-double value and |op(X)||op(Y)| scale of the checked entries of
    op(X) op(Y) (n x n x n), through contiguous rows of op(X) (x stored, xt
    its transpose) and contiguous columns of op(Y)."""
    rows_src = xt if tx == "N" else x  # op(X)(i, p) = rows_src[p + i n] (op "T")
    cols_src = y if ty == "N" else yt  # op(Y)(p, j) = cols_src[p + j n] (op "N")
    ri = np.concatenate([np.repeat(np.asarray(R), n), np.tile(np.arange(n), len(Cc)), sr])
    cj = np.concatenate([np.tile(np.arange(n), len(R)), np.repeat(np.asarray(Cc), n), sc])
    return oracle.gemm_ref_entries("T", "N", n, alpha, rows_src, n, cols_src, n, 0.0, None, n,
                                   ri, cj)


# ---------------------------------------------------------------------------
# [0]

def test_config0_dgemm_and_dtrsm_512(H):
    n = 512
    a = oracle.fill_uniform(n, n, 0)
    b = oracle.fill_uniform(n, n, 1)
    c = oracle.fill_uniform(n, n, 2)
    da, db, dc = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b), H.DeviceArray.from_host(c)
    H.gemm("N", "N", n, n, n, 1.5, da, n, db, n, 0.5, dc, n)
    got = dc.to_host()
    assert gemm_bound_ratio(got, "N", "N", n, n, n, 1.5, a, n, b, n, 0.5, c, n) <= 1.0
    t = oracle.fill_uniform(n, n, 3)
    t[np.arange(n), np.arange(n)] = n + (oracle.fill_uniform(n, 1, 4)[:, 0] + 1.0) / 2.0
    t[np.triu_indices(n, 1)] = np.nan  # the other triangle is never read
    rhs = oracle.fill_uniform(n, n, 5)
    dt, dx = H.DeviceArray.from_host(t), H.DeviceArray.from_host(rhs)
    H.trsm("L", "L", "N", "N", n, n, 1.0, dt, n, dx, n)
    x = dx.to_host()
    r = oracle.trsm_residual("L", "L", "N", "N", n, n, 1.0, t, n, rhs, n, x, n, np.arange(n))
    assert float(np.max(r)) <= 8 * (n + 2) * U64 * 2, float(np.max(r))


# ---------------------------------------------------------------------------
# [1]

@pytest.mark.parametrize("n", [4096, 8192, 16384])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_config1_gemm_every_op_and_layout(H, dtype, n):
    a = oracle.fill_uniform(n, n, 11, dtype=dtype)
    b = oracle.fill_uniform(n, n, 12, dtype=dtype)
    da, db, dc = H.DeviceArray(n, n, dtype), H.DeviceArray(n, n, dtype), H.DeviceArray(n, n, dtype)
    da.fill_uniform(n, n, 11)
    db.fill_uniform(n, n, 12)
    at, bt = np.asfortranarray(a.T), np.asfortranarray(b.T)
    R, Cc, sr, sc = check_entries(n)
    tol = 4.0 * (n + 2) * eps_of(dtype)
    worst = 0.0
    for layout in ("col", "row"):
        for ta, tb in TRANS:
            H.lib().hwyblas_amd_memset(dc.ptr, 0xFF, n * n * np.dtype(dtype).itemsize)  # NaN
            if layout == "col":
                H.gemm(ta, tb, n, n, n, 1.0, da, n, db, n, 0.0, dc, n, dtype=dtype)
                ops = (a, at, ta, b, bt, tb)
            else:
                H.cblas_gemm(CBLAS["row"], CBLAS[ta], CBLAS[tb], n, n, n, 1.0, da, n, db, n,
                             0.0, dc, n, dtype=dtype)
                # Row-major C = op(A_rm) op(B_rm) is, in column-major storage,
                # op_tb(B) op_ta(A).
                ops = (b, bt, tb, a, at, ta)
            c = dc.to_host()
            got = entries_of(c, R, Cc, sr, sc).astype(np.float64)
            del c
            exp, bnd = reference(n, *ops, R, Cc, sr, sc)
            ratio = float(np.max(np.abs(got - exp) / (tol * bnd)))
            assert ratio <= 1.0, (layout, ta, tb, ratio)
            worst = max(worst, ratio)
    print(f"{np.dtype(dtype).name} {n}: worst bound ratio {worst:.3g}")


@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "LU")))
def test_config1_dtrsm_8192(H, side, uplo):
    n = 8192
    t = oracle.fill_uniform(n, n, 21)
    t[np.arange(n), np.arange(n)] = n + (oracle.fill_uniform(n, 1, 22)[:, 0] + 1.0) / 2.0
    if uplo == "L":
        t[np.triu_indices(n, 1)] = np.nan
    else:
        t[np.tril_indices(n, -1)] = np.nan
    rhs = oracle.fill_uniform(n, n, 23)
    dt = H.DeviceArray.from_host(t)
    which = np.random.default_rng(7).choice(n, 16, replace=False)
    for ta in "NT":
        dx = H.DeviceArray.from_host(rhs)
        H.trsm(side, uplo, ta, "N", n, n, 1.0, dt, n, dx, n)
        x = dx.to_host()
        del dx
        r = oracle.trsm_residual(side, uplo, ta, "N", n, n, 1.0, t, n, rhs, n, x, n, which)
        assert float(np.max(r)) <= 8 * (n + 2) * U64 * 2, (ta, float(np.max(r)))


# ---------------------------------------------------------------------------
# [2]

@pytest.mark.parametrize("dist", ["a", "b", "c"])
def test_config2_emulated_16384(H, dist):
    n = 16384
    L = H.lib()
    da, db, dc = H.DeviceArray(n, n), H.DeviceArray(n, n), H.DeviceArray(n, n)
    if dist == "b":
        L.hwyblas_amd_fill_pow2(da.ptr, n, n, n, 11, -30, 30, None)
        L.hwyblas_amd_fill_pow2(db.ptr, n, n, n, 12, -30, 30, None)
        a, b = oracle.fill_pow2(n, n, 11), oracle.fill_pow2(n, n, 12)
    else:
        da.fill_uniform(n, n, 1)
        db.fill_uniform(n, n, 2)
        a, b = oracle.fill_uniform(n, n, 1), oracle.fill_uniform(n, n, 2)
        if dist == "c":
            rs = np.power(10.0, np.random.default_rng(21).uniform(-8, 8, n))
            cs = np.power(10.0, np.random.default_rng(22).uniform(-8, 8, n))
            drs, dcs = H.DeviceArray.from_host(rs.reshape(-1, 1)), H.DeviceArray.from_host(
                cs.reshape(-1, 1))
            L.hwyblas_amd_scale_rows_cols(da.ptr, n, n, n, drs.ptr, None, None)
            L.hwyblas_amd_scale_rows_cols(db.ptr, n, n, n, None, dcs.ptr, None)
            a *= rs[:, None]
            b *= cs[None, :]
    H.synchronize()
    R, Cc, sr, sc = check_entries(n)
    at = np.asfortranarray(a.T)
    exp, bnd = reference(n, a, at, "N", b, None, "N", R, Cc, sr, sc)
    del a, at, b
    tol = 4.0 * (n + 2) * U64

    def run(kind):
        if kind == "native":
            assert H.gemm_dev("N", "N", n, n, n, 1.0, da, n, db, n, 0.0, dc, n,
                              dtype=np.float64) == 0
            used = None
        elif kind == "int8":
            info, used = H.dgemm_emulated_dev("N", "N", n, n, n, 1.0, da, n, db, n, 0.0, dc, n)
            assert info == 0
        else:
            info, used = H.dgemm_emulated_bf16_dev("N", "N", n, n, n, 1.0, da, n, db, n, 0.0,
                                                   dc, n)
            assert info == 0
        H.synchronize()
        c = dc.to_host()
        got = entries_of(c, R, Cc, sr, sc)
        del c
        err = np.abs(got - exp) / bnd
        return used, float(np.max(err))

    _, e_nat = run("native")
    used8, e8 = run("int8")
    used16, e16 = run("bf16")
    print(f"dist {dist}: native {e_nat:.3g}, int8 ({used8} moduli) {e8:.3g}, "
          f"bf16 ({used16} slices) {e16:.3g}")
    assert e_nat <= tol and e8 <= tol and e16 <= tol
    assert used8 >= 14 and used16 >= 7
    if dist == "a":
        assert used8 == 14 and used16 == 7
    assert e8 <= 2.0 * e_nat, (e8, e_nat)


# ---------------------------------------------------------------------------
# [3]

def log_uniform_ints(rng, lo, hi, size):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size))).astype(int).clip(lo, hi)


def arena_layout(shapes):
    sizes = np.array([max(1, r * c) for r, c in shapes], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum((sizes + 63) // 64 * 64)])
    return off[:-1], int(off[-1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_config3_ten_thousand_gemms(H, dtype):
    count = 10000
    rng = np.random.default_rng(7)
    m, n, k = (log_uniform_ints(rng, 8, 512, count) for _ in range(3))
    ta = rng.integers(0, 2, count)
    tb = rng.integers(0, 2, count)
    shapes = [[(k[t], m[t]) if ta[t] else (m[t], k[t]) for t in range(count)],
              [(n[t], k[t]) if tb[t] else (k[t], n[t]) for t in range(count)],
              [(m[t], n[t]) for t in range(count)]]
    isz = np.dtype(dtype).itemsize
    dev, host, offs = [], [], []
    for s, sh in enumerate(shapes):
        off, total = arena_layout(sh)
        d = H.DeviceArray(total, 1, dtype)
        d.fill_uniform(total, 1, 100 + s)
        dev.append(d)
        host.append(oracle.fill_uniform(total, 1, 100 + s, dtype=dtype)[:, 0])
        offs.append(off)
    H.synchronize()
    probs = [H.GemmProblem(b"T" if ta[t] else b"N", b"T" if tb[t] else b"N", int(m[t]),
                           int(n[t]), int(k[t]), 1.0, 0.0, dev[0].ptr + int(offs[0][t]) * isz,
                           int(shapes[0][t][0]), dev[1].ptr + int(offs[1][t]) * isz,
                           int(shapes[1][t][0]), dev[2].ptr + int(offs[2][t]) * isz, int(m[t]))
             for t in range(count)]
    plan = H.Plan("gemm", dtype, probs)
    plan.launch(graph=True)
    H.synchronize()
    cflat = np.empty(dev[2].ld, dtype=dtype)
    H.lib().hwyblas_amd_memcpy(cflat.ctypes.data, dev[2].ptr, cflat.nbytes, 2)
    plan.destroy()

    def block(which, t):
        r, c = shapes[which][t]
        o = int(offs[which][t])
        return host[which][o:o + r * c].reshape((r, c), order="F")

    def check(t):
        mm, nn, kk = int(m[t]), int(n[t]), int(k[t])
        a, b = block(0, t), block(1, t)
        got = cflat[int(offs[2][t]):int(offs[2][t]) + mm * nn].reshape((mm, nn), order="F")
        return gemm_bound_ratio(got, "T" if ta[t] else "N", "T" if tb[t] else "N", mm, nn, kk,
                                1.0, a, a.shape[0], b, b.shape[0], 0.0, None, mm)

    with ThreadPoolExecutor(16) as ex:
        ratios = np.array(list(ex.map(check, range(count))))
    assert np.all(ratios <= 1.0), (int(np.argmax(ratios)), float(ratios.max()))


def test_config3_ten_thousand_dtrsms(H):
    count = 10000
    rng = np.random.default_rng(11)
    m = log_uniform_ints(rng, 8, 256, count)
    n = log_uniform_ints(rng, 1, 64, count)
    offa, tota = arena_layout([(x, x) for x in m])
    offb, totb = arena_layout([(m[t], n[t]) for t in range(count)])
    ha = oracle.fill_uniform(tota, 1, 200)[:, 0]
    hb = oracle.fill_uniform(totb, 1, 201)[:, 0]
    for t in range(count):
        o, mm = int(offa[t]), int(m[t])
        idx = o + np.arange(mm) * (mm + 1)
        ha[idx] = mm + (ha[idx] + 1.0) * 0.5  # diagonal m + U[0, 1]
    da = H.DeviceArray.from_host(ha.reshape(-1, 1))
    db = H.DeviceArray.from_host(hb.reshape(-1, 1))
    probs = [H.TrsmProblem(b"L", b"L", b"N", b"N", int(m[t]), int(n[t]), 1.0,
                           da.ptr + int(offa[t]) * 8, int(m[t]), db.ptr + int(offb[t]) * 8,
                           int(m[t])) for t in range(count)]
    plan = H.Plan("trsm", np.float64, probs)
    plan.launch(graph=True)
    H.synchronize()
    x = db.to_host()[:, 0]
    plan.destroy()

    def check(t):
        mm, nn = int(m[t]), int(n[t])
        a = ha[int(offa[t]):int(offa[t]) + mm * mm].reshape((mm, mm), order="F")
        b0 = hb[int(offb[t]):int(offb[t]) + mm * nn].reshape((mm, nn), order="F")
        xt = x[int(offb[t]):int(offb[t]) + mm * nn].reshape((mm, nn), order="F")
        r = oracle.trsm_residual("L", "L", "N", "N", mm, nn, 1.0, a, mm, b0, mm, xt, mm,
                                 np.arange(nn))
        return float(np.max(r)) / (8 * (mm + 2) * U64 * 2)

    with ThreadPoolExecutor(16) as ex:
        ratios = np.array(list(ex.map(check, range(count))))
    assert np.all(ratios <= 1.0), (int(np.argmax(ratios)), float(ratios.max()))


# ---------------------------------------------------------------------------
# [4]

@pytest.fixture(scope="module")
def orthogonal_factors(H):
    """Q1, Q2 (n = 16384): Stewart's random orthogonal matrices (the Q of a
    Householder QR of a Gaussian matrix), applied to the identity on the
    device through the library's DGEMM (bench.apply_orthogonal)."""
    import bench
    n = 16384
    stream = H.lib().hwyblas_amd_stream_create()
    qs = []
    for seed in (31, 32):
        q = H.DeviceArray(n, n)
        for j in range(0, n, 2048):  # the identity, 2048 columns at a time
            blk = np.zeros((n, min(2048, n - j)), order="F")
            blk[j + np.arange(blk.shape[1]), np.arange(blk.shape[1])] = 1.0
            H.lib().hwyblas_amd_memcpy(q.ptr + j * n * 8, blk.ctypes.data, blk.nbytes, 1)
        bench.apply_orthogonal(H, q, n, seed, "L", stream)
        qs.append(q)
    H.lib().hwyblas_amd_stream_synchronize(stream)
    H.lib().hwyblas_amd_stream_destroy(stream)
    yield qs
    for q in qs:
        q.free()


@pytest.mark.parametrize("kappa", [2, 6, 10])
def test_config4_lu_solve_16384(H, orthogonal_factors, kappa):
    n, nrhs = 16384, 16
    q1, q2 = orthogonal_factors
    sig = np.power(10.0, kappa * np.arange(n) / (n - 1))
    dsig = H.DeviceArray.from_host(sig.reshape(-1, 1))
    qs = H.DeviceArray(n, n)
    H.lib().hwyblas_amd_memcpy(qs.ptr, q1.ptr, n * n * 8, 3)
    H.lib().hwyblas_amd_scale_rows_cols(qs.ptr, n, n, n, None, dsig.ptr, None)
    da = H.DeviceArray(n, n)
    assert H.gemm_dev("N", "T", n, n, n, 1.0, qs, n, q2, n, 0.0, da, n, dtype=np.float64) == 0
    qs.free()
    db = H.DeviceArray(n, nrhs)
    db.fill_uniform(n, nrhs, 33)
    H.synchronize()
    x, berr, nbe, iters, info = H.gesv_ir_dev(da, db, update=1, itmax=5)
    assert info == 0
    bound = 4 * (n + 1) * U64
    assert float(berr.max()) < bound, (kappa, berr.max(), iters)
    # Independent check on the host: r = A x - b and |A||x| + |b| in
    # double-double, every row of every right-hand side.
    a = da.to_host()
    at = np.asfortranarray(a.T)
    anorm = float(np.abs(a).sum(axis=1).max())
    del a
    xh = x.to_host()
    b = db.to_host()
    ri = np.tile(np.arange(n), nrhs)
    cj = np.repeat(np.arange(nrhs), n)
    r, scale = oracle.gemm_ref_entries("T", "N", n, 1.0, at, n, xh, n, -1.0, b, n, ri, cj)
    del at
    berr_host = float(np.max(np.abs(r) / scale))
    nbe_host = float(max(np.max(np.abs(r[j * n:(j + 1) * n])) /
                         (anorm * np.max(np.abs(xh[:, j]))) for j in range(nrhs)))
    print(f"kappa 1e{kappa}: berr {berr.max():.3g} (host {berr_host:.3g}), normwise "
          f"{nbe.max():.3g} (host {nbe_host:.3g}), corrections {iters.min()}-{iters.max()}")
    assert berr_host < bound, berr_host
    assert nbe_host < bound, nbe_host
    assert abs(berr_host - float(berr.max())) <= 0.05 * berr_host + 4 * U64
