"""Batched (grouped) GEMM and TRSM plans on the GPU, through the C-ABI.

Every problem of an uneven batch is checked against the oracle's
double-double reference with hwyBLAS's componentwise bounds (the GEMM bound
of tests/gemm_test.cc:72-101, the TRSM residual of
tests/level3_test.cc:482-520). The grouped GEMM tiles sum k in the same
order as the single-call kernel, so each problem's result is also bitwise
equal to hwyblas_?gemm_dev on the same operands, under every tile
configuration. hipGraph replays are checked one by one against the eager
launch, with eager work in between.
"""
import numpy as np
import pytest

import oracle
from helpers import gemm_bound_ratio, op_shape, trsm_operands, trsm_residual_ratio

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def log_uniform(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size)).astype(int).clip(lo, hi)


def gemm_batch(rng, count, dtype, lo=1, hi=160, zeros=True):
    """Host operands of `count` uneven GEMMs (shapes log-uniform in [lo, hi])."""
    shapes = np.stack([log_uniform(rng, lo, hi, count) for _ in range(3)], axis=1)
    if zeros:
        for t in range(0, count, 37):
            shapes[t, t % 3] = 0
    probs = []
    for t, (m, n, k) in enumerate(shapes):
        m, n, k = int(m), int(n), int(k)
        ta, tb = "NT"[rng.integers(2)], "NT"[rng.integers(2)]
        pad = int(rng.integers(0, 3))
        ra, ca = op_shape(ta, m, k)
        rb, cb = op_shape(tb, k, n)
        lda, ldb, ldc = max(1, ra + pad), max(1, rb + pad), max(1, m + pad)
        seed = 1000 + 3 * t
        a = oracle.fill_uniform(ra, ca, seed, dtype=dtype, ld=lda)
        b = oracle.fill_uniform(rb, cb, seed + 1, dtype=dtype, ld=ldb)
        c = oracle.fill_uniform(m, n, seed + 2, dtype=dtype, ld=ldc)
        if pad:
            c[m:, :] = 7.25
        alpha = float(rng.choice([1.0, -1.0, 0.5, 0.0, 1.25]))
        beta = float(rng.choice([0.0, 1.0, -0.5, 2.0]))
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, a=a, lda=lda,
                          b=b, ldb=ldb, c=c, ldc=ldc))
    return probs


class DeviceGemmBatch:
    def __init__(self, H, probs, dtype):
        self.H, self.probs, self.dtype = H, probs, np.dtype(dtype)
        self.A = H.Arena([p["a"] for p in probs], dtype)
        self.B = H.Arena([p["b"] for p in probs], dtype)
        self.C = H.Arena([p["c"] for p in probs], dtype)

    def problems(self, C=None):
        C = C or self.C
        out = []
        for t, p in enumerate(self.probs):
            out.append(self.H.GemmProblem(p["ta"].encode(), p["tb"].encode(), p["m"], p["n"],
                                          p["k"], p["alpha"], p["beta"], self.A.ptr(t), p["lda"],
                                          self.B.ptr(t), p["ldb"], C.ptr(t), p["ldc"]))
        return out

    def results(self, C=None):
        C = C or self.C
        flat = C.download()
        return [C.block(flat, t, p["c"].shape) for t, p in enumerate(self.probs)]

    def reset_c(self):
        self.C = self.H.Arena([p["c"] for p in self.probs], self.dtype)

    def single_call_results(self):
        """Each problem through hwyblas_?gemm_dev, one call at a time."""
        H = self.H
        C = H.Arena([p["c"] for p in self.probs], self.dtype)
        for t, p in enumerate(self.probs):
            info = H.gemm_dev(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"],
                              self.A.ptr(t), p["lda"], self.B.ptr(t), p["ldb"], p["beta"],
                              C.ptr(t), p["ldc"], dtype=self.dtype)
            assert info == 0
        H.synchronize()
        return self.results(C)


def check_gemm_results(probs, got):
    worst = 0.0
    for t, (p, g) in enumerate(zip(probs, got)):
        m, n = p["m"], p["n"]
        r = gemm_bound_ratio(g, p["ta"], p["tb"], m, n, p["k"], p["alpha"], p["a"], p["lda"],
                             p["b"], p["ldb"], p["beta"], p["c"], p["ldc"])
        assert r <= 1.0, (t, m, n, p["k"], p["ta"], p["tb"], p["alpha"], p["beta"], r)
        worst = max(worst, r)
        assert np.array_equal(g[m:, :], p["c"][m:, :]), f"padding of C[{t}] modified"
        if (p["k"] == 0 or p["alpha"] == 0.0) and p["beta"] == 1.0:
            assert np.array_equal(g, p["c"], equal_nan=True)
    return worst


@pytest.mark.parametrize("dtype", DTYPES)
def test_grouped_gemm_uneven_batch(H, dtype):
    """400 uneven problems (log-uniform shapes up to 160, random trans, padded
    ld, zero dimensions, alpha/beta special values) in one launch."""
    rng = np.random.default_rng(5)
    probs = gemm_batch(rng, 400, dtype)
    batch = DeviceGemmBatch(H, probs, dtype)
    plan = H.Plan("gemm", dtype, batch.problems())
    plan.launch()
    H.synchronize()
    got = batch.results()
    worst = check_gemm_results(probs, got)
    ref = batch.single_call_results()
    for t, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(g, r), f"problem {t} differs from the single-call kernel"
    for t, p in enumerate(probs):
        cfg = plan.problem_config(t)
        if p["m"] == 0 or p["n"] == 0 or ((p["k"] == 0 or p["alpha"] == 0) and p["beta"] == 1):
            assert cfg == -1
        else:
            # alpha == 0 is planned as k == 0 (A and B are not read).
            k_eff = 0 if p["alpha"] == 0.0 else p["k"]
            assert cfg == H.lib().hwyblas_grouped_default_config(np.dtype(dtype).itemsize, p["m"],
                                                                 p["n"], k_eff)
    print(f"{plan.work_items} work items, worst bound ratio {worst:.3g}")


@pytest.mark.parametrize("dtype", DTYPES)
def test_grouped_gemm_every_config_bitwise(H, dtype):
    """Each tile configuration, forced on every problem, gives the same bits."""
    rng = np.random.default_rng(9)
    probs = gemm_batch(rng, 60, dtype, hi=140)
    batch = DeviceGemmBatch(H, probs, dtype)
    ref = batch.single_call_results()
    check_gemm_results(probs, ref)
    for cfg in range(len(H.grouped_configs())):
        batch.reset_c()
        plan = H.Plan("gemm", dtype, batch.problems(), configs=[cfg] * len(probs))
        plan.launch()
        H.synchronize()
        for t, (g, r) in enumerate(zip(batch.results(), ref)):
            assert np.array_equal(g, r), (cfg, t)
        plan.destroy()


def test_grouped_gemm_mixed_configs_and_edges(H):
    """Per-problem configurations (random), tile-edge shapes, k = 1, m or n = 1,
    NaN in C with beta = 0 (not read) and NaN in A (propagates)."""
    rng = np.random.default_rng(21)
    dtype = np.float64
    shapes = [(1, 1, 1), (16, 16, 4), (17, 15, 5), (64, 64, 64), (65, 63, 129), (1, 300, 33),
              (300, 1, 33), (128, 128, 1), (33, 97, 256), (200, 190, 180)]
    probs = []
    for t, (m, n, k) in enumerate(shapes):
        a = oracle.fill_uniform(m, k, 50 + t, dtype=dtype)
        b = oracle.fill_uniform(k, n, 70 + t, dtype=dtype)
        c = oracle.fill_uniform(m, n, 90 + t, dtype=dtype)
        beta = 0.0 if t % 2 == 0 else 0.5
        if beta == 0.0:
            c[:] = np.nan
        probs.append(dict(ta="N", tb="N", m=m, n=n, k=k, alpha=1.0, beta=beta, a=a, lda=m, b=b,
                          ldb=k, c=c, ldc=m))
    probs[3]["a"][5, 7] = np.nan  # row 5 of C[3] becomes NaN
    batch = DeviceGemmBatch(H, probs, dtype)
    cfgs = rng.integers(0, len(H.grouped_configs()), len(probs))
    plan = H.Plan("gemm", dtype, batch.problems(), configs=cfgs)
    plan.launch()
    H.synchronize()
    got = batch.results()
    assert np.isnan(got[3][5, :]).all()
    assert not np.isnan(np.delete(got[3], 5, axis=0)).any()
    probs[3]["a"][5, 7] = 0.0
    got[3][5, :] = 0.0
    for t, (p, g) in enumerate(zip(probs, got)):
        if t == 3:
            continue
        assert not np.isnan(g).any(), t
    check = [p for t, p in enumerate(probs) if t != 3]
    check_gemm_results(check, [g for t, g in enumerate(got) if t != 3])


@pytest.mark.parametrize("dtype", DTYPES)
def test_grouped_gemm_graph_replays(H, dtype):
    """The plan captured in a hipGraph: three replays, each after eager work
    that overwrites C, each bitwise equal to the eager launch."""
    rng = np.random.default_rng(17)
    probs = gemm_batch(rng, 150, dtype, hi=120)
    batch = DeviceGemmBatch(H, probs, dtype)
    plan = H.Plan("gemm", dtype, batch.problems())
    plan.launch()
    H.synchronize()
    eager = batch.results()
    check_gemm_results(probs, eager)
    C0 = H.Arena([p["c"] for p in probs], dtype)
    for replay in range(3):
        # Eager work between replays: restore C (beta != 0 reads it), then
        # scribble on an unrelated buffer with the fill kernel.
        H.lib().hwyblas_amd_memcpy(batch.C.dev.ptr, C0.dev.ptr,
                                   batch.C.total * np.dtype(dtype).itemsize, 3)
        scratch = H.DeviceArray(1024, 4, dtype)
        scratch.fill_uniform(1024, 4, 99 + replay)
        H.synchronize()
        plan.launch(graph=True)
        H.synchronize()
        for t, (g, e) in enumerate(zip(batch.results(), eager)):
            assert np.array_equal(g, e, equal_nan=True), (replay, t)


def trsm_batch(rng, count, dtype, hi=150):
    probs = []
    for t in range(count):
        side = "LR"[rng.integers(2)]
        uplo = "LU"[rng.integers(2)]
        ta = "NT"[rng.integers(2)]
        diag = "NU"[rng.integers(2)]
        m, n = (int(x) for x in log_uniform(rng, 1, hi, 2))
        if t % 41 == 0:
            m = 0
        if t % 43 == 0:
            n = 0
        alpha = float(rng.choice([1.0, -0.5, 2.0, 0.0]))
        if m == 0 or n == 0:
            ka = m if side == "L" else n
            lda = max(1, ka)
            a = np.zeros((lda, max(ka, 1)), dtype=dtype, order="F")
            ldb = max(1, m)
            b = np.zeros((ldb, max(n, 1)), dtype=dtype, order="F")
        else:
            a, lda, b, ldb = trsm_operands(side, uplo, diag, m, n, dtype, 3000 + 5 * t,
                                           lda_pad=int(rng.integers(0, 3)),
                                           ldb=m + int(rng.integers(0, 3)))
        probs.append(dict(side=side, uplo=uplo, ta=ta, diag=diag, m=m, n=n, alpha=alpha, a=a,
                          lda=lda, b=b, ldb=ldb))
    return probs


class DeviceTrsmBatch:
    def __init__(self, H, probs, dtype):
        self.H, self.probs, self.dtype = H, probs, np.dtype(dtype)
        self.A = H.Arena([p["a"] for p in probs], dtype)
        self.B = H.Arena([p["b"] for p in probs], dtype)

    def problems(self):
        return [self.H.TrsmProblem(p["side"].encode(), p["uplo"].encode(), p["ta"].encode(),
                                   p["diag"].encode(), p["m"], p["n"], p["alpha"],
                                   self.A.ptr(t), p["lda"], self.B.ptr(t), p["ldb"])
                for t, p in enumerate(self.probs)]

    def results(self):
        flat = self.B.download()
        return [self.B.block(flat, t, p["b"].shape) for t, p in enumerate(self.probs)]


def check_trsm_results(probs, got):
    worst = 0.0
    for t, (p, x) in enumerate(zip(probs, got)):
        m, n = p["m"], p["n"]
        if m == 0 or n == 0:
            assert np.array_equal(x, p["b"]), t
            continue
        if p["alpha"] == 0.0:
            assert not x[:m, :n].any(), t
        else:
            r = trsm_residual_ratio(p["side"], p["uplo"], p["ta"], p["diag"], m, n, p["alpha"],
                                    p["a"], p["lda"], p["b"], p["ldb"], x)
            assert r <= 1.0, (t, p["side"], p["uplo"], p["ta"], p["diag"], m, n, r)
            worst = max(worst, r)
        assert np.array_equal(x[m:, :], p["b"][m:, :]), f"padding of B[{t}] modified"
    return worst


@pytest.mark.parametrize("dtype", DTYPES)
def test_grouped_trsm_uneven_batch(H, dtype):
    """300 uneven triangular solves over every side/uplo/trans/diag, alpha in
    {1, -0.5, 2, 0}, padded ld, empty problems; the unreferenced triangle and
    a unit diagonal hold NaN, so reading them would show."""
    rng = np.random.default_rng(3)
    probs = trsm_batch(rng, 300, dtype)
    batch = DeviceTrsmBatch(H, probs, dtype)
    plan = H.Plan("trsm", dtype, batch.problems())
    plan.launch()
    H.synchronize()
    worst = check_trsm_results(probs, batch.results())
    print(f"{plan.work_items} work items, worst residual ratio {worst:.3g}")


def test_grouped_trsm_wide_rhs_and_graph(H):
    """Right-hand sides past one 64-lane chunk (n = 200, left; m = 130, right),
    systems past one 32-row block, replayed from a hipGraph after resetting B."""
    dtype = np.float64
    cases = [("L", "L", "N", "N", 100, 200), ("L", "U", "T", "U", 65, 130),
             ("R", "U", "N", "N", 130, 70), ("R", "L", "T", "N", 129, 33),
             ("L", "U", "N", "N", 256, 64), ("R", "L", "N", "U", 1, 256)]
    probs = []
    for t, (side, uplo, ta, diag, m, n) in enumerate(cases):
        a, lda, b, ldb = trsm_operands(side, uplo, diag, m, n, dtype, 400 + t)
        probs.append(dict(side=side, uplo=uplo, ta=ta, diag=diag, m=m, n=n, alpha=0.75, a=a,
                          lda=lda, b=b, ldb=ldb))
    batch = DeviceTrsmBatch(H, probs, dtype)
    plan = H.Plan("trsm", dtype, batch.problems())
    plan.launch()
    H.synchronize()
    eager = batch.results()
    check_trsm_results(probs, eager)
    B0 = H.Arena([p["b"] for p in probs], dtype)
    for replay in range(3):
        H.lib().hwyblas_amd_memcpy(batch.B.dev.ptr, B0.dev.ptr, batch.B.total * 8, 3)
        H.synchronize()
        plan.launch(graph=True)
        H.synchronize()
        for t, (g, e) in enumerate(zip(batch.results(), eager)):
            assert np.array_equal(g, e), (replay, t)


def test_empty_plan(H):
    plan = H.Plan("gemm", np.float64, [])
    assert plan.work_items == 0
    plan.launch()
    plan.launch(graph=True)
    H.synchronize()
