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
    assert lu_residual_ratio(a, lu, piv, x, b, trans) <= 1.0


def lu_residual_ratio(a, lu, piv, x, b, trans):
    """max |op(A) x - b| / bound, the residual in double-double. For LU with
    substitution, (A + dA) x = b with |dA| <= gamma_3n P^T |L||U| (Higham,
    Accuracy and Stability, Thm 9.4). The GPU solve applies the inverses of
    the 64 x 64 diagonal blocks of L and U instead of substituting within
    them, which amplifies a block step's error by at most about
    || |T^-1| |T| ||inf of that block (Skeel's condition number): the bound
    is gamma_3n times the largest such factor over the blocks, times
    P^T |L||U| |x|."""
    n, nrhs = x.shape
    u = 2.0 ** -53
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    perm = np.arange(n)
    for i, p in enumerate(piv):
        perm[[i, p - 1]] = perm[[p - 1, i]]
    M = np.empty((n, n))
    M[perm, :] = np.abs(L) @ np.abs(U)  # P A = L U, so A = P^T L U
    skeel = 1.0
    for t in (L, U):
        for j in range(0, n, 64):
            blk = t[j:j + 64, j:j + 64]
            skeel = max(skeel, float(np.max(np.abs(np.linalg.inv(blk)) @ np.abs(blk) @ np.ones(len(blk)))))
    r, _ = oracle.gemm_ref(trans, "N", n, nrhs, n, 1.0, a, n, x, n, -1.0, b, n)
    opM = M if trans == "N" else M.T
    gamma = 3 * n * u / (1 - 3 * n * u)
    bound = gamma * skeel * (opM @ np.abs(x)) * (1 + 2 * n * u) + np.finfo(float).tiny
    return float(np.max(np.abs(r) / bound))


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
    # Singular: column 17 (0-based) stays exactly zero through the
    # elimination (row swaps and updates only combine its zeros), so the
    # first zero pivot is U(18, 18), 1-based.
    s = a.copy(order="F")
    s[:, 17] = 0.0
    ps = np.zeros(n, dtype=np.int32)
    info = H.lapacke_lu("getrf", H.ColMajor, n, n, s, n, ps)
    assert info == 18


def well_conditioned(n, nrhs, seed):
    """hwyBLAS's WellConditioned (tests/lapack/linear_test.cc:33-42):
    uniform[-1, 1] entries with n / 2 added to the diagonal."""
    a = oracle.fill_uniform(n, n, seed) + np.eye(n) * (n / 2)
    b = oracle.fill_uniform(n, nrhs, seed + 1)
    return np.asfortranarray(a), np.asfortranarray(b)


def perturbed_solution(a, b, trans, delta, seed):
    """The exact solution of op(A) X = B (numpy, then one refinement step in
    double-double is not needed at this conditioning) times (1 + delta r)."""
    op = a if trans == "N" else a.T
    xt = np.linalg.solve(op, b)
    rng = np.random.default_rng(seed)
    return xt, np.asfortranarray(xt * (1 + delta * rng.uniform(-1, 1, xt.shape)))


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("nrhs", [2, 20])
def test_dgerfs_refines_like_xgerfs(trans, nrhs):
    """hwyBLAS's refinement test (tests/lapack/linear_test.cc:498-560) on the
    GPU dgerfs_: solutions perturbed by 1e-6 come back within 100 n eps,
    ferr bounds the remaining error (>= 0.3 err / |x|) and stays below
    1e3 n eps, berr < 4 (n + 1) eps; ferr matches the oracle's xGERFS for the
    same X and factors within 5 %. 20 right-hand sides cross the GPU path's
    groups of 16."""
    n = 300
    a, b = well_conditioned(n, nrhs, 31)
    lu = a.copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    assert H.getrf(n, n, lu, n, piv) == 0
    xt, x0 = perturbed_solution(a, b, trans, 1e-6, 7)
    x = x0.copy(order="F")
    info, ferr, berr = H.gerfs(trans, n, nrhs, a, n, lu, n, piv, b, n, x, n)
    assert info == 0
    eps = 2.0 ** -53
    # The oracle's xGERFS on the GPU's refined X: its bound for the same X.
    _, ferr_o, berr_o = oracle.gerfs(trans, a, lu, piv, b, x)
    for j in range(nrhs):
        err = np.max(np.abs(x[:, j] - xt[:, j]))
        xn = np.max(np.abs(x[:, j]))
        assert err / xn < 100 * n * eps, (j, err / xn)
        assert ferr[j] >= 0.3 * err / xn
        assert ferr[j] < 1e3 * n * eps
        assert berr[j] < 4 * (n + 1) * eps
        assert abs(ferr[j] - ferr_o[j]) <= 0.05 * ferr_o[j], (j, ferr[j], ferr_o[j])
    assert float(np.max(berr)) <= 4 * (n + 1) * eps and float(np.max(berr_o)) <= 4 * (n + 1) * eps


def test_dgerfs_device_operands_and_underflow_scale():
    """Device operands give the host call's bits; with B scaled near underflow,
    SAFE1 dominates and berr equals the formula's value recomputed on the host
    (linear_test.cc:536-554) within 1 %."""
    n, nrhs = 200, 3
    a, b = well_conditioned(n, nrhs, 41)
    lu = a.copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    assert H.getrf(n, n, lu, n, piv) == 0
    _, x0 = perturbed_solution(a, b, "N", 1e-6, 9)
    x_host = x0.copy(order="F")
    info, ferr, berr = H.gerfs("N", n, nrhs, a, n, lu, n, piv, b, n, x_host, n)
    assert info == 0
    da, dlu, db, dx = (H.DeviceArray.from_host(v) for v in (a, lu, b, x0))
    dpiv = H.DeviceArray.from_host(piv.reshape(-1, 1))
    info2, ferr2, berr2 = H.gerfs("N", n, nrhs, da, n, dlu, n, dpiv, db, n, dx, n)
    assert info2 == 0
    assert np.array_equal(dx.to_host(), x_host)
    assert np.array_equal(ferr, ferr2) and np.array_equal(berr, berr2)
    # Near underflow.
    scale = np.finfo(float).tiny * 2.0 ** 40
    bs = np.asfortranarray(b * scale)
    _, xs0 = perturbed_solution(a, bs, "N", 1e-6, 11)
    xs = xs0.copy(order="F")
    info, ferr, berr = H.gerfs("N", n, nrhs, a, n, lu, n, piv, bs, n, xs, n)
    assert info == 0
    safe1 = (n + 1) * np.finfo(float).tiny
    safe2 = safe1 / 2.0 ** -53
    r, scl = oracle.gemm_ref("N", "N", n, nrhs, n, 1.0, a, n, xs, n, -1.0, bs, n)
    for j in range(nrhs):
        num, den = np.abs(r[:, j]), scl[:, j]
        terms = np.where(den > safe2, num / den, (num + safe1) / (den + safe1))
        want = float(np.max(terms))
        assert abs(berr[j] - want) <= 0.01 * want, (j, berr[j], want)


def test_dgetrf_reports_a_failed_panel_barrier(monkeypatch):
    """HWYBLAS_AMD_LU_FAIL_BARRIER=1 raises the cooperative panel kernel's
    failure flag from the host before the first panel: every bounded wait
    gives up at its first poll and every workgroup returns, and dgetrf_
    returns -1000, reported to the error handler as ("DGETRF", -1000). A
    normal call afterwards factors correctly."""
    n = 512
    a = oracle.fill_uniform(n, n, 3)
    monkeypatch.setenv("HWYBLAS_AMD_LU_FAIL_BARRIER", "1")
    lu = a.copy(order="F")
    piv = np.zeros(n, dtype=np.int32)
    with H.ErrorCapture() as cap:
        info = H.getrf(n, n, lu, n, piv)
    assert info == -1000 and cap.calls == [("DGETRF", -1000)]
    monkeypatch.delenv("HWYBLAS_AMD_LU_FAIL_BARRIER")
    lu = a.copy(order="F")
    assert H.getrf(n, n, lu, n, piv) == 0
    _, want_piv, info = oracle.getrf(a)
    assert info == 0 and np.array_equal(piv, want_piv)
