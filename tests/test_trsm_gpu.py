"""TRSM parity on the GPU through the C-ABI (dtrsm_/strsm_, cblas_?trsm,
hwyblas_?trsm_dev), checked through the residual op(A) X - alpha B with
hwyBLAS's bound (tests/level3_test.cc:482-555) and the special-value and
subnormal semantics (tests/semantics_test.cc:189-230, :344-356).
"""
import itertools

import numpy as np
import pytest

import oracle
from helpers import trsm_operands, trsm_residual_ratio

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
COMBOS = list(itertools.product("LR", "LU", "NT", "NU"))  # side, uplo, trans, diag


def run_trsm(H, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    x = b.copy(order="F")
    H.trsm(side, uplo, ta, diag, m, n, alpha, a, lda, x, ldb)
    return x


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side,uplo,ta,diag", COMBOS)
def test_all_combinations(H, dtype, side, uplo, ta, diag):
    """All 16 side/uplo/trans/diag combinations, m across the base and
    recursion edges (level3_test.cc:522-538)."""
    seed = 1
    for m in (1, 2, 5, 31, 32, 33, 63, 64, 65, 100, 129, 300):
        for (mm, nn, alpha) in ((m, 17, 1.5), (13, m, 1.0)):
            a, lda, b, ldb = trsm_operands(side, uplo, diag, mm, nn, dtype, seed)
            seed += 2
            x = run_trsm(H, side, uplo, ta, diag, mm, nn, alpha, a, lda, b, ldb)
            r = trsm_residual_ratio(side, uplo, ta, diag, mm, nn, alpha, a, lda, b, ldb, x)
            assert r <= 1.0, (mm, nn, r)
            assert np.array_equal(x[mm:, :], b[mm:, :]), "padding of B modified"
    # Leading dimension a multiple of 1 KiB.
    unit = 1024 // np.dtype(dtype).itemsize
    ldb = (300 + unit - 1) // unit * unit
    a, lda, b, ldb = trsm_operands(side, uplo, diag, 300, 45, dtype, 999, ldb=ldb)
    x = run_trsm(H, side, uplo, ta, diag, 300, 45, 0.5, a, lda, b, ldb)
    assert trsm_residual_ratio(side, uplo, ta, diag, 300, 45, 0.5, a, lda, b, ldb, x) <= 1.0


@pytest.mark.parametrize("dtype", DTYPES)
def test_alpha_zero_does_not_read_a(H, dtype):
    """alpha == 0 sets B to zero without reading A (all NaN)
    (level3_test.cc:540-555)."""
    a = np.full((3, 3), np.nan, dtype=dtype, order="F")
    b = np.ones((3, 2), dtype=dtype, order="F")
    H.trsm("L", "L", "N", "N", 3, 2, 0.0, a, 3, b, 3)
    assert np.all(b == 0)


@pytest.mark.parametrize("dtype", DTYPES)
def test_nan_reaches_dependent_elements_only(H, dtype):
    """A NaN/Inf in B(i0, j0) reaches exactly the elements that depend on it
    (semantics_test.cc:189-230)."""
    m, n = 70, 9
    i0, j0 = 33, 4
    for side, uplo, ta in itertools.product("LR", "LU", "NT"):
        ka = m if side == "L" else n
        rng = np.random.default_rng(ka)
        a = np.zeros((ka, ka), dtype=dtype, order="F")
        for j in range(ka):
            for i in range(ka):
                if i == j:
                    a[i, j] = 2 + abs(rng.uniform(-1, 1))
                elif (i < j) if uplo == "U" else (i > j):
                    a[i, j] = rng.uniform(-1, 1) / ka
        for value in (np.nan, np.inf):
            b = np.asfortranarray(rng.uniform(-1, 1, (m, n)).astype(dtype))
            b[i0, j0] = value
            H.trsm(side, uplo, ta, "N", m, n, 1.0, a, ka, b, m)
            op_lower = (uplo == "L") == (ta == "N")
            hit = np.zeros((m, n), bool)
            for j in range(n):
                for i in range(m):
                    if side == "L":
                        hit[i, j] = j == j0 and (i >= i0 if op_lower else i <= i0)
                    else:
                        hit[i, j] = i == i0 and (j <= j0 if op_lower else j >= j0)
            special = ~np.isfinite(b)
            assert special[hit].all(), (side, uplo, ta, value)
            assert not special[~hit].any(), (side, uplo, ta, value)


@pytest.mark.parametrize("dtype", DTYPES)
def test_identity_preserves_subnormals(H, dtype):
    """trsm with the identity keeps subnormal right-hand sides exactly
    (semantics_test.cc:344-356)."""
    s2 = np.ldexp(dtype(1), -1060 if dtype == np.float64 else -140)
    assert s2 != 0
    for n in (1, 5, 64, 200):
        eye = np.eye(n, dtype=dtype, order="F")
        b = np.full((n, 3), s2, dtype=dtype, order="F")
        H.trsm("L", "L", "N", "N", n, 3, 1.0, eye, n, b, n)
        assert np.all(b == s2)


@pytest.mark.parametrize("dtype", DTYPES)
def test_cblas_trsm_layouts(H, dtype):
    """Column-major CBLAS equals the Fortran routine; row-major solves the
    transposed problem (cblas_test.cc:1024-1076)."""
    m, n = 90, 37
    S = {"L": H.Left, "R": H.Right}
    U = {"L": H.Lower, "U": H.Upper}
    T = {"N": H.NoTrans, "T": H.Trans}
    for side, uplo, ta in itertools.product("LR", "LU", "NT"):
        a, lda, b, ldb = trsm_operands(side, uplo, "N", m, n, dtype, 5)
        f = run_trsm(H, side, uplo, ta, "N", m, n, 0.5, a, lda, b, ldb)
        x = b.copy(order="F")
        H.cblas_trsm(H.ColMajor, S[side], U[uplo], T[ta], H.NonUnit, m, n, 0.5, a, lda, x, ldb)
        assert np.array_equal(x, f)
        # Row-major: B (m x n) row-major; A stored row-major.
        ka = m if side == "L" else n
        ar = np.ascontiguousarray(np.nan_to_num(a[:ka, :ka], nan=0.0))
        br = np.ascontiguousarray(b[:m, :n])
        xr = br.copy()
        H.cblas_trsm(H.RowMajor, S[side], U[uplo], T[ta], H.NonUnit, m, n, 0.5, ar, ka, xr, n)
        A = ar if ta == "N" else ar.T
        A = np.tril(A) if (uplo == "L") == (ta == "N") else np.triu(A)
        lhs = A.astype(np.float64) @ xr if side == "L" else xr.astype(np.float64) @ A
        scale = (np.abs(A) @ np.abs(xr) if side == "L" else np.abs(xr) @ np.abs(A)) + 0.5 * np.abs(br)
        tol = 8 * (ka + 2) * np.finfo(dtype).eps
        assert np.all(np.abs(lhs - 0.5 * br) <= tol * scale)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side,uplo,ta", list(itertools.product("LR", "LU", "NT")))
def test_large_device_sampled(H, dtype, side, uplo, ta):
    """m = n = 1536 on device pointers (diagonal n + U[0,1] as in the
    benchmark's TRSM workload), 16 sampled right-hand sides checked."""
    n = 1536
    rng = np.random.default_rng(1)
    a = oracle.fill_uniform(n, n, 40, dtype=dtype)
    a[np.arange(n), np.arange(n)] = n + rng.uniform(0, 1, n)
    b = oracle.fill_uniform(n, n, 41, dtype=dtype)
    da, db = H.DeviceArray.from_host(a), H.DeviceArray.from_host(b)
    assert H.trsm_dev(side, uplo, ta, "N", n, n, 1.0, da, n, db, n, dtype=dtype) == 0
    x = db.to_host()
    which = rng.integers(0, n, 16)
    r = trsm_residual_ratio(side, uplo, ta, "N", n, n, 1.0, a, n, b, n, x, which)
    assert r <= 1.0, r


FUSED_CASES = [(side, m, n, uplo, ta, diag, dt)
               for side in "LR"
               for (m, n) in ((128, 17), (256, 40), (300, 33), (1000, 1), (2048, 300))
               for uplo, ta, diag in (("L", "N", "N"), ("U", "N", "U"), ("L", "T", "N"),
                                      ("U", "T", "N"))
               for dt in ("f8", "f4")]


def _fused_shape(side, m, n):
    # The order of the triangle is the first size; the right side swaps them.
    return (m, n) if side == "L" else (n, m)


def solve_fused_cases(H):
    """Solves whose recursion reaches diagonal blocks of order 128 and 256
    (the fused block kernel's cases), both sides."""
    out = []
    for i, (side, m, n, uplo, ta, diag, dt) in enumerate(FUSED_CASES):
        mm, nn = _fused_shape(side, m, n)
        a, lda, b, ldb = trsm_operands(side, uplo, diag, mm, nn, np.dtype(dt).type, 7000 + 3 * i)
        out.append(run_trsm(H, side, uplo, ta, diag, mm, nn, 0.75, a, lda, b, ldb))
    return out


def test_fused_block_solve_equals_unfused_recursion(H, tmp_path):
    """A diagonal block of order 128 or 256 is solved by one launch that
    runs the recursion's own steps; its bits equal the unfused
    recursion's (small solves and GEMM updates launched one by one, taken
    with HWYBLAS_AMD_TRSM_FUSED=0 in a child process)."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    out = tmp_path / "unfused.npz"
    code = ("import sys, numpy as np\n"
            f"sys.path[:0] = [{root!r}, {here!r}]\n"
            "import hwyblas_amd, test_trsm_gpu as T\n"
            f"np.savez({str(out)!r}, *T.solve_fused_cases(hwyblas_amd))\n")
    env = dict(os.environ, HWYBLAS_AMD_TRSM_FUSED="0")
    subprocess.run([sys.executable, "-c", code], env=env, cwd=root, check=True, timeout=240)
    got = solve_fused_cases(H)
    with np.load(out) as ref:
        for i, case in enumerate(FUSED_CASES):
            want = ref[f"arr_{i}"]
            assert got[i].tobytes() == want.tobytes(), case
    # And each is a solution within the residual bound.
    for i, (side, m, n, uplo, ta, diag, dt) in enumerate(FUSED_CASES):
        mm, nn = _fused_shape(side, m, n)
        a, lda, b, ldb = trsm_operands(side, uplo, diag, mm, nn, np.dtype(dt).type, 7000 + 3 * i)
        r = trsm_residual_ratio(side, uplo, ta, diag, mm, nn, 0.75, a, lda, b, ldb, got[i])
        assert r <= 1.0, (side, m, n, uplo, ta, diag, dt, r)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side,uplo,ta", [("L", "L", "N"), ("L", "U", "T"), ("R", "U", "N"),
                                          ("R", "L", "T")])
def test_few_right_hand_sides_split_the_updates(H, dtype, side, uplo, ta):
    """A 4096 x 4096 triangle against 8 right-hand sides: the recursion's
    updates (8 columns or rows of B, k up to 2048) split k; the residual
    stays within the bound and B's padding is untouched."""
    mm, nn = (4096, 8) if side == "L" else (8, 4096)
    a, lda, b, ldb = trsm_operands(side, uplo, "N", mm, nn, dtype, 77)
    x = run_trsm(H, side, uplo, ta, "N", mm, nn, 1.0, a, lda, b, ldb)
    r = trsm_residual_ratio(side, uplo, ta, "N", mm, nn, 1.0, a, lda, b, ldb, x)
    assert r <= 1.0, r
    assert np.array_equal(x[mm:, :], b[mm:, :])
