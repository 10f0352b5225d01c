"""The LU consumer on the GPU (hwyblas_dgetrf_dev / dgetrs_dev /
dgesv_ir_dev) through the C-ABI, against the oracle's restatement of
hwyBLAS's dgetrf / dgetrs (pivots, factors) and the criteria of hwyBLAS's
tests: ||P L U - A||_1 / (n ||A||_1 eps) < 30 (tests/getrf_test.cc:42-96)
and the refined backward error (tests/linear_test.cc:498-560)."""
import numpy as np
import pytest

import oracle
from helpers import forward_bound, permuted_triangular_system

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def lu_ratio(a, lu, ipiv):
    P, L, U = oracle.lu_factors(lu, ipiv)
    n = max(a.shape)
    return np.abs(P @ L @ U - a).sum(0).max() / (n * np.abs(a).sum(0).max() * EPS)


def gpu_getrf(H, a, update=0):
    d = H.DeviceArray.from_host(np.asfortranarray(a))
    ipiv, info = H.getrf_dev(d, update)
    lu = d.to_host()
    piv = ipiv.to_host().ravel()[:min(a.shape)]
    return lu, piv, info


@pytest.mark.parametrize("n", [1, 2, 7, 63, 64, 65, 130, 300, 513, 1000])
def test_getrf_matches_oracle(H, n):
    """Same pivots as the oracle (hwyBLAS's algorithm), factors equal to
    rounding, and the residual criterion."""
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    lu, piv, info = gpu_getrf(H, a)
    elu, epiv, einfo = oracle.getrf(a)
    assert info == einfo == 0
    assert np.array_equal(piv, epiv)
    assert np.max(np.abs(lu - elu)) <= 1e-9 * max(1.0, np.max(np.abs(elu)))
    assert lu_ratio(a, lu, piv) < 30


@pytest.mark.parametrize("m,n", [(1000, 1000), (2100, 700), (5000, 300), (90, 300), (129, 129)])
def test_cooperative_panel_equals_column_steps(H, m, n, monkeypatch):
    """The cooperative panel kernel (default) and the one-launch-per-column
    panel give the same factors and pivots bit for bit, also with a zero
    column, an exact tie between candidate rows of different workgroups and
    a NaN."""
    rng = np.random.default_rng(m + 3 * n)
    a = rng.standard_normal((m, n))
    a[:, min(n - 1, 40)] = 0.0
    a[5, 3] = a[m - 2, 3] = 50.0  # tie: the smaller row wins
    outs = []
    for mode in (None, "steps"):
        if mode:
            monkeypatch.setenv("HWYBLAS_AMD_LU_PANEL", mode)
        else:
            monkeypatch.delenv("HWYBLAS_AMD_LU_PANEL", raising=False)
        outs.append(gpu_getrf(H, a))
    (lu0, p0, i0), (lu1, p1, i1) = outs
    assert i0 == i1 and np.array_equal(p0, p1)
    assert np.array_equal(lu0, lu1)
    b = a.copy()
    b[m // 2, 7] = np.nan
    outs = []
    for mode in (None, "steps"):
        if mode:
            monkeypatch.setenv("HWYBLAS_AMD_LU_PANEL", mode)
        else:
            monkeypatch.delenv("HWYBLAS_AMD_LU_PANEL", raising=False)
        outs.append(gpu_getrf(H, b))
    (lu0, p0, i0), (lu1, p1, i1) = outs
    assert i0 == i1 and np.array_equal(p0, p1)
    assert np.array_equal(lu0, lu1, equal_nan=True)


@pytest.mark.parametrize("m,n", [(300, 100), (100, 300), (700, 129), (65, 64)])
def test_getrf_rectangular(H, m, n):
    rng = np.random.default_rng(m * 7 + n)
    a = rng.standard_normal((m, n))
    lu, piv, info = gpu_getrf(H, a)
    elu, epiv, _ = oracle.getrf(a)
    assert info == 0 and np.array_equal(piv, epiv)
    assert lu_ratio(a, lu, piv) < 30


def test_getrf_zero_column_reports_info(H):
    """An exactly zero pivot column: info is its 1-based index (the first
    one), the factorization still completes (src/lapack/getrf.cc:107-110)."""
    rng = np.random.default_rng(4)
    a = rng.standard_normal((200, 200))
    a[:, 77] = 0.0
    a[:, 150] = 0.0
    lu, piv, info = gpu_getrf(H, a)
    elu, epiv, einfo = oracle.getrf(a)
    assert info == einfo == 78
    assert np.array_equal(piv, epiv)


def test_getrf_nan_first_row_pivot(H):
    """xIAMAX semantics: a NaN in the first candidate row wins the pivot."""
    a = np.random.default_rng(5).standard_normal((40, 40))
    a[0, 0] = np.nan
    _, piv, _ = gpu_getrf(H, a)
    _, epiv, _ = oracle.getrf(a)
    assert piv[0] == epiv[0] == 1


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("n,nrhs", [(400, 5), (333, 17), (200, 70), (1, 3), (64, 64)])
@pytest.mark.parametrize("mode", [None, "trsm"])
def test_getrs_matches_oracle(H, trans, n, nrhs, mode, monkeypatch):
    """Blocked solves with the diagonal blocks' inverses (nrhs <= 64, ragged
    last block at n = 333, several right-hand-side chunks at 17 and 64), the
    blocked TRSM solves (HWYBLAS_AMD_GETRS=trsm) and the recursive TRSMs
    (nrhs = 70), both transpositions."""
    if mode:
        monkeypatch.setenv("HWYBLAS_AMD_GETRS", mode)
    else:
        monkeypatch.delenv("HWYBLAS_AMD_GETRS", raising=False)
    rng = np.random.default_rng(11)
    a = rng.standard_normal((n, n))
    b = rng.uniform(-1, 1, (n, nrhs))
    da = H.DeviceArray.from_host(np.asfortranarray(a))
    ipiv, info = H.getrf_dev(da)
    assert info == 0
    db = H.DeviceArray.from_host(np.asfortranarray(b))
    assert H.lib().hwyblas_dgetrs_dev(trans.encode(), n, nrhs, da.ptr, n, ipiv.ptr, db.ptr, n,
                                      None) == 0
    H.synchronize()
    x = db.to_host()
    elu, epiv, _ = oracle.getrf(a)
    ex = oracle.getrs(trans, elu, epiv, b)
    assert np.max(np.abs(x - ex)) <= 1e-9 * max(1.0, np.max(np.abs(ex)))


def test_getrf_update_precisions(H):
    """n = 2304: the top-level update has k = 1152, so update 1 runs the
    emulated GEMM (FP64-accurate: the criterion holds) and update 2 runs FP32
    (backward error at FP32 level, well above the FP64 criterion)."""
    rng = np.random.default_rng(21)
    n = 2304
    a = rng.standard_normal((n, n))
    lu0, piv0, _ = gpu_getrf(H, a, 0)
    lu1, piv1, _ = gpu_getrf(H, a, 1)
    lu2, piv2, _ = gpu_getrf(H, a, 2)
    r0, r1, r2 = lu_ratio(a, lu0, piv0), lu_ratio(a, lu1, piv1), lu_ratio(a, lu2, piv2)
    print(f"residual ratios: fp64 {r0:.3g}, emulated {r1:.3g}, fp32 {r2:.3g}")
    assert r0 < 30 and r1 < 30
    assert 1e3 < r2 < 1e9


def conditioned(n, kappa, seed):
    rng = np.random.default_rng(seed)
    q1, _ = np.linalg.qr(rng.standard_normal((n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.geomspace(1.0, 10.0 ** kappa, n)
    return (q1 * s) @ q2.T


@pytest.mark.parametrize("update,kappa", [(0, 6), (1, 6), (2, 2), (2, 6)])
def test_gesv_refinement(H, update, kappa):
    """A = Q1 diag(sigma) Q2^T: after refinement every column's berr is at
    the FP64 level (hwyBLAS's criterion berr < 4 (n + 1) eps) and agrees with
    the oracle's berr of the returned X."""
    n, nrhs = 1024, 16
    a = conditioned(n, kappa, 100 + kappa)
    b = np.random.default_rng(7).uniform(-1, 1, (n, nrhs))
    da = H.DeviceArray.from_host(np.asfortranarray(a))
    db = H.DeviceArray.from_host(np.asfortranarray(b))
    x, berr, nbe, iters, info = H.gesv_ir_dev(da, db, update=update)
    assert info == 0
    xh = x.to_host()
    ob = oracle.gerfs_berr(a, b, xh)
    print(f"update {update} kappa 1e{kappa}: berr {berr.max():.3g} (oracle {ob.max():.3g}), "
          f"nbe {nbe.max():.3g}, iterations {iters.min()}-{iters.max()}")
    assert np.all(berr < 4 * (n + 1) * EPS / 2)
    assert np.all(ob < 4 * (n + 1) * EPS / 2)
    assert np.all(nbe < 10 * EPS)
    if update == 2:
        assert iters.min() >= 1  # FP32 factors need corrections


@pytest.mark.parametrize("update", [0, 1, 2])
def test_gesv_forward_bound_closed_form(H, update):
    """Permuted triangular integer systems are solved exactly (R = 0), and
    inv(A) >= 0 makes the estimated matrix nonnegative, so the estimate is
    exact: ferr equals the bound's defining formula, evaluated with an
    explicit inverse. n = 2100 runs its top-level update (k = 1050) on the
    emulated GEMM for update 1."""
    for n, nrhs in ((1, 1), (7, 3), (300, 16), (2100, 16)):
        a, xt, b = permuted_triangular_system(n, nrhs, n + update)
        da = H.DeviceArray.from_host(a)
        db = H.DeviceArray.from_host(b)
        x, berr, nbe, iters, info, ferr = H.gesv_ir_dev(da, db, update=update, with_ferr=True)
        assert info == 0
        assert np.array_equal(x.to_host(), xt), n
        assert np.all(berr == 0) and np.all(iters == 0)
        want = forward_bound(a, xt, b)
        assert np.allclose(ferr, want, rtol=1e-12, atol=0), (n, ferr / want - 1)


@pytest.mark.parametrize("kappa", [2, 6, 10])
def test_gesv_forward_bound_matches_oracle(H, kappa):
    """A = Q1 diag(sigma) Q2^T: the GPU's ferr for its refined X against the
    oracle's xGERFS on the same X (its own LU and solves), within 5%: the
    weights differ only in |R| at the rounding level, a few percent of the
    (n + 1) eps term at n = 512. The bound grows with the condition number."""
    n, nrhs = 512, 16
    a = conditioned(n, kappa, 200 + kappa)
    b = np.random.default_rng(8).uniform(-1, 1, (n, nrhs))
    da = H.DeviceArray.from_host(np.asfortranarray(a))
    db = H.DeviceArray.from_host(np.asfortranarray(b))
    x, berr, nbe, iters, info, ferr = H.gesv_ir_dev(da, db, with_ferr=True)
    assert info == 0
    xh = x.to_host()
    lu, piv, _ = oracle.getrf(a)
    _, oferr, oberr = oracle.gerfs("N", a, lu, piv, b, xh)
    print(f"kappa 1e{kappa}: ferr {ferr.max():.3g} (oracle {oferr.max():.3g})")
    assert np.allclose(ferr, oferr, rtol=0.05, atol=0)
    assert np.all(ferr < 10.0 ** kappa * 10 * n * EPS)


def test_gesv_forward_bound_near_underflow(H):
    """B scaled by 2^-982 (hwyBLAS's near-underflow case): safe1 enters the
    weights and ferr >= 0.3 safe1 / (||A||inf ||X||inf)
    (tests/lapack/linear_test.cc:541-558), and matches the oracle's."""
    n, nrhs = 300, 4
    a = conditioned(n, 3, 5)
    b = np.random.default_rng(9).uniform(-1, 1, (n, nrhs)) * 2.0 ** -982
    da = H.DeviceArray.from_host(np.asfortranarray(a))
    db = H.DeviceArray.from_host(np.asfortranarray(b))
    x, berr, nbe, iters, info, ferr = H.gesv_ir_dev(da, db, with_ferr=True)
    assert info == 0
    xh = x.to_host()
    safe1 = (n + 1) * np.finfo(np.float64).tiny
    xn = np.abs(xh).max(0)
    assert np.all(ferr >= 0.3 * safe1 / (np.abs(a).sum(1).max() * xn))
    lu, piv, _ = oracle.getrf(a)
    _, oferr, _ = oracle.gerfs("N", a, lu, piv, b, xh)
    assert np.allclose(ferr, oferr, rtol=0.05, atol=0)


def test_gesv_without_ferr_leaves_the_rest_unchanged(H):
    """ferr = NULL skips the estimate; X, berr and nbe are the same."""
    n, nrhs = 257, 5
    a = conditioned(n, 4, 6)
    b = np.random.default_rng(10).uniform(-1, 1, (n, nrhs))
    da = H.DeviceArray.from_host(np.asfortranarray(a))
    db = H.DeviceArray.from_host(np.asfortranarray(b))
    x0, berr0, nbe0, it0, info0 = H.gesv_ir_dev(da, db)
    x1, berr1, nbe1, it1, info1, ferr = H.gesv_ir_dev(da, db, with_ferr=True)
    assert info0 == info1 == 0
    assert np.array_equal(x0.to_host(), x1.to_host())
    assert np.array_equal(berr0, berr1) and np.array_equal(nbe0, nbe1)
    assert np.all(ferr > 0)
