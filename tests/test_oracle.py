"""Pins the CPU oracle (CPU only; no GPU).

The oracle's restated hwyBLAS GEMM/TRSM and its double-double references are
checked against (1) golden fixtures computed by numpy/OpenBLAS and scipy
(tests/golden/blas_fixtures.npz), (2) exact rational arithmetic on tiny
inputs, (3) closed forms (integer-valued products, constant matrices,
subnormal known answers, exactly solvable triangles) and (4) the published
splitmix64 test vector. Every planted mistake in the oracle must make at
least one of these checks fail.
"""
import itertools
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import (forward_bound, gemm_bound_ratio, permuted_triangular_system,
                     trsm_operands, trsm_residual_ratio)

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "blas_fixtures.npz"))


@pytest.fixture(autouse=True)
def _no_planted_mistake():
    oracle.plant_mistake(0)
    yield
    oracle.plant_mistake(0)


def gemm_fixture_cases():
    for line in FIX["gemm_cases"]:
        key, dt, ta, tb, m, n, k, alpha, beta = str(line).split(",")
        yield (key, np.dtype(dt), ta, tb, int(m), int(n), int(k), float(alpha), float(beta))


def trsm_fixture_cases():
    for line in FIX["trsm_cases"]:
        key, dt, side, uplo, ta, diag, m, n, alpha = str(line).split(",")
        yield (key, np.dtype(dt), side, uplo, ta, diag, int(m), int(n), float(alpha))


def _f(x):
    return np.asfortranarray(x)


def oracle_matches_gemm_fixtures(square_only=False):
    """Largest bound ratio of the oracle's GEMM and of the double-double
    reference against numpy's results (both must satisfy the bound, so they
    are within twice the tolerance of each other). square_only: the m = n = k
    cases alone, where a wrongly transposed operand is still read in bounds."""
    worst = 0.0
    for key, dt, ta, tb, m, n, k, alpha, beta in gemm_fixture_cases():
        if square_only and not (m == n == k):
            continue
        a, b, c = _f(FIX[key + "_a"]), _f(FIX[key + "_b"]), _f(FIX[key + "_c"])
        out = FIX[key + "_out"]
        got = c.copy(order="F")
        oracle.gemm(ta, tb, m, n, k, alpha, a, a.shape[0], b, b.shape[0], beta, got, m)
        exp, bnd = oracle.gemm_ref(ta, tb, m, n, k, alpha, a, a.shape[0], b, b.shape[0], beta, c, m)
        tol = 4 * (k + 2) * np.finfo(dt).eps
        worst = max(worst,
                    float(np.max(np.abs(got - out) / (2 * tol * bnd))),
                    float(np.max(np.abs(exp - out) / (tol * bnd))))
    return worst


def test_gemm_matches_numpy_fixtures():
    assert oracle_matches_gemm_fixtures() <= 1.0


def oracle_matches_trsm_fixtures():
    worst = 0.0
    for key, dt, side, uplo, ta, diag, m, n, alpha in trsm_fixture_cases():
        a, b, out = _f(FIX[key + "_a"]), _f(FIX[key + "_b"]), FIX[key + "_out"]
        ka = a.shape[0]
        x = b.copy(order="F")
        oracle.trsm(side, uplo, ta, diag, m, n, alpha, a, ka, x, m)
        # Both solutions satisfy the residual bound; they agree to within the
        # forward error, which for these well-conditioned triangles is a few
        # ka * eps relative to |X|.
        tol = 8 * (ka + 2) * np.finfo(dt).eps
        worst = max(worst, float(np.max(np.abs(x - out) / (tol * (np.abs(out) + 1)))))
        r = trsm_residual_ratio(side, uplo, ta, diag, m, n, alpha, a, ka, b, m, x)
        worst = max(worst, r)
    return worst


def test_trsm_matches_scipy_fixtures():
    assert oracle_matches_trsm_fixtures() <= 1.0


def test_reference_is_exact_on_tiny_inputs():
    """The double-double reference equals the exact rational result rounded
    to double (within one ulp) on tiny inputs with heavy cancellation."""
    rng = np.random.default_rng(5)
    for trial in range(20):
        m, n, k = 3, 2, 6
        a = _f(rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-20, 20, (m, k))))
        b = _f(rng.uniform(-1, 1, (k, n)))
        # Force cancellation: the last term nearly cancels the rest.
        for i in range(m):
            for j in range(n):
                pass
        c = _f(rng.uniform(-1, 1, (m, n)))
        alpha, beta = 0.75, -1.25
        exp, _ = oracle.gemm_ref("N", "N", m, n, k, alpha, a, m, b, k, beta, c, m)
        for i in range(m):
            for j in range(n):
                s = sum(Fraction(a[i, p]) * Fraction(b[p, j]) for p in range(k))
                exact = Fraction(alpha) * s + Fraction(beta) * Fraction(c[i, j])
                ulp = np.spacing(abs(float(exact)))
                assert abs(Fraction(exp[i, j]) - exact) <= Fraction(ulp), (trial, i, j)


def test_cancellation_needs_the_low_word():
    """x + y - x with |y| << |x|: the exact answer is y; a reference without
    double-double accumulation returns 0."""
    a = _f(np.array([[1.0, 1e-20, -1.0]]))
    b = _f(np.array([[1.0], [1.0], [1.0]]))
    exp, _ = oracle.gemm_ref("N", "N", 1, 1, 3, 1.0, a, 1, b, 3, 0.0, None, 1)
    assert exp[0, 0] == 1e-20


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_closed_forms(dtype):
    rng = np.random.default_rng(2)
    for ta, tb in itertools.product("NT", "NT"):
        m, n, k = 41, 23, 517
        a = _f(rng.integers(-8, 9, (m, k) if ta == "N" else (k, m)).astype(dtype))
        b = _f(rng.integers(-8, 9, (k, n) if tb == "N" else (n, k)).astype(dtype))
        c = _f(rng.integers(-9, 9, (m, n)).astype(dtype))
        A = a if ta == "N" else a.T
        B = b if tb == "N" else b.T
        want = (2 * (A.astype(np.int64) @ B.astype(np.int64)) - 3 * c.astype(np.int64)).astype(dtype)
        got = c.copy(order="F")
        oracle.gemm(ta, tb, m, n, k, 2.0, a, a.shape[0], b, b.shape[0], -3.0, got, m)
        assert np.array_equal(got, want), (ta, tb)
    # Constant matrices: k * a * b.
    k = 1000
    a = np.full((5, k), 0.5, dtype=dtype, order="F")
    b = np.full((k, 4), 0.25, dtype=dtype, order="F")
    c = np.zeros((5, 4), dtype=dtype, order="F")
    oracle.gemm("N", "N", 5, 4, k, 1.0, a, 5, b, k, 0.0, c, 5)
    assert np.all(c == dtype(k * 0.125))
    # Subnormal products (semantics_test.cc:311-342).
    s = dtype(np.ldexp(1.0, -530 if dtype == np.float64 else -70))
    for n in (1, 5, 64, 200):
        x = np.full((n, n), s, dtype=dtype, order="F")
        c = np.ones((n, n), dtype=dtype, order="F")
        oracle.gemm("N", "N", n, n, n, 1.0, x, n, x, n, 0.0, c, n)
        assert np.all(c == dtype(n) * (s * s))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_trsm_exactly_solvable(dtype):
    """Triangles with power-of-two diagonals and integer entries: X integer,
    B = op(T) X exactly, so the solve must return X exactly."""
    rng = np.random.default_rng(9)
    for side, uplo, ta, diag in itertools.product("LR", "LU", "NT", "NU"):
        m, n = 70, 5
        ka = m if side == "L" else n
        t = rng.integers(-1, 2, (ka, ka)).astype(np.float64)
        t = np.tril(t, -1) if uplo == "L" else np.triu(t, 1)
        d = rng.choice([1.0, -1.0, 2.0, -2.0, 0.5], ka) if diag == "N" else np.ones(ka)
        t[np.arange(ka), np.arange(ka)] = d
        op = t if ta == "N" else t.T
        # Keep X small integers and B exactly representable: solve upward
        # from an integer X.
        x = rng.integers(-3, 4, (m, n)).astype(np.float64)
        bb = op @ x if side == "L" else x @ op
        if np.max(np.abs(bb)) > 2 ** 20:
            continue
        a = _f(t.astype(dtype))
        if diag == "U":
            a[np.arange(ka), np.arange(ka)] = np.nan
        b = _f(bb.astype(dtype))
        oracle.trsm(side, uplo, ta, diag, m, n, 1.0, a, ka, b, m)
        assert np.array_equal(b, x.astype(dtype)), (side, uplo, ta, diag)


def test_trsm_random_residuals():
    for side, uplo, ta, diag in itertools.product("LR", "LU", "NT", "NU"):
        for m, n in ((100, 17), (13, 100), (300, 45)):
            a, lda, b, ldb = trsm_operands(side, uplo, diag, m, n, np.float64, m + n)
            x = b.copy(order="F")
            oracle.trsm(side, uplo, ta, diag, m, n, 1.5, a, lda, x, ldb)
            assert trsm_residual_ratio(side, uplo, ta, diag, m, n, 1.5, a, lda, b, ldb, x) <= 1.0


def test_splitmix64_published_vector():
    """splitmix64 (Vigna) from state 0: first outputs e220a8397b1dcdaf,
    6e789e6aa1b965f4 (state advances by the golden gamma)."""
    assert oracle.lib().orc_splitmix64(0) == 0xE220A8397B1DCDAF
    assert oracle.lib().orc_splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_generator_is_layout_independent():
    x = oracle.fill_uniform(37, 11, 5)
    y = oracle.fill_uniform(37, 11, 5, ld=50)
    assert np.array_equal(x, y[:37])
    assert np.all((x >= -1) & (x < 1))
    assert abs(np.mean(oracle.fill_uniform(1000, 100, 6))) < 0.01


@pytest.mark.parametrize("mistake", [1, 2, 3])
def test_planted_gemm_mistakes_are_caught(mistake):
    oracle.plant_mistake(mistake)
    # A wrong transposition (2) would read past non-square operands.
    assert oracle_matches_gemm_fixtures(square_only=mistake == 2) > 1.0


def test_planted_trsm_mistake_is_caught():
    oracle.plant_mistake(4)
    assert oracle_matches_trsm_fixtures() > 1.0


def test_planted_reference_mistake_is_caught():
    oracle.plant_mistake(5)
    with pytest.raises(AssertionError):
        test_cancellation_needs_the_low_word()


def test_ozaki_crt_matches_python_integers():
    """The emulation's CRT step (oracle restatement) against exact Python
    integers: for random X with |X| < P/2 (as the scaling guarantees), the
    residues mod the first nmod moduli reconstruct X to within one rounding
    plus the fixed-point truncation P N 256 2^-160."""
    import random
    from math import prod
    rnd = random.Random(7)
    for nmod in (2, 5, 9, 14, 20, 24):
        mods = oracle.ozaki_moduli(nmod)
        assert all(np.gcd(a, b) == 1 for i, a in enumerate(mods) for b in mods[i + 1:])
        P = prod(mods)
        lim = P // 2 - 1 - (P >> 30)
        for t in range(300):
            X = rnd.randrange(-lim, lim)
            if t < 5:
                X = [0, 1, -1, lim, -lim][t]
            elif t < 120:
                X = rnd.randrange(-min(2 ** 40, lim), min(2 ** 40, lim))
            got = oracle.ozaki_crt(nmod, [X % p for p in mods])
            tol = 6e-16 * abs(X) + P * nmod * 256 * 2.0 ** -160
            assert abs(got - X) <= tol, (nmod, X, got)


def test_ozaki_bits_rule():
    """k 2^(bA+bB) < P / 2 with the largest such bit count."""
    from math import log2, prod
    for nmod in (3, 8, 14, 24):
        P = prod(oracle.ozaki_moduli(nmod))
        for k in (1, 100, 16384):
            bits = oracle.ozaki_total_bits(nmod, k)
            assert k * 2 ** bits < P // 2 and k * 2 ** (bits + 1) >= P // 2 * 0.999


def test_ozaki_emulation_error_at_native_level():
    """With 14 moduli the restated emulation's componentwise error is at the
    typical level of native FP64 accumulation, sqrt(k) u; with 12 it is far
    above it (the moduli matter). The native worst-case bound 4 (k + 2) u
    holds for both."""
    m, n, k = 64, 64, 1000
    a = oracle.fill_uniform(m, k, 1)
    b = oracle.fill_uniform(k, n, 2)
    c = oracle.fill_uniform(m, n, 3)
    exp, bnd = oracle.gemm_ref("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, c, m)
    eps = np.finfo(float).eps
    for nmod, ok in ((14, True), (12, False)):
        cc = c.copy(order="F")
        oracle.dgemm_ozaki("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, cc, m, nmod)
        err = np.max(np.abs(cc - exp) / bnd)
        assert err <= 4 * (k + 2) * eps
        assert (err <= np.sqrt(k) * eps) == ok, (nmod, err / eps)
    # Integer operands: the integer product is recovered exactly and only
    # the final conversion rounds (a few ulps).
    rng = np.random.default_rng(1)
    ai = np.asfortranarray(rng.integers(-1000, 1000, (m, k)).astype(float))
    bi = np.asfortranarray(rng.integers(-1000, 1000, (k, n)).astype(float))
    cc = np.zeros((m, n), order="F")
    oracle.dgemm_ozaki("N", "N", m, n, k, 1.0, ai, m, bi, k, 0.0, cc, m, 14)
    want = ai @ bi
    assert np.all(np.abs(cc - want) <= 4 * np.spacing(np.abs(want)))


def test_planted_crt_mistake_is_caught():
    oracle.plant_mistake(6)
    with pytest.raises(AssertionError):
        test_ozaki_crt_matches_python_integers()


# ---------------------------------------------------------------------------
# The BF16-slice emulation's restatement (oracle/emulated_bf16.cc)

def test_bf16_digits_known_answers():
    """Balanced base-256 digits (nearest, ties to even): worked by hand, and
    for random values of magnitude below 1/2 every digit is an integer of
    magnitude <= 128 and sum_t d_t 256^-t is within 256^-s / 2 of x (exact
    rationals)."""
    from fractions import Fraction
    assert list(oracle.bf16_digits(0.375, 0, 3)) == [96, 0, 0]
    assert list(oracle.bf16_digits(-0.375, 0, 2)) == [-96, 0]
    # x 2^-1 = 24 256^-1 + 2.5 256^-2 + 0.5 256^-3: 256 r = 24.0098 -> 24;
    # then 2.50195 -> 3 (not a tie), leaving -0.498; 256 (-0.498) = -127.5
    # -> -128 (tie to even), leaving 0.5; then 128.
    x = 3 * 2.0 ** -4 + 5 * 2.0 ** -16 + 2.0 ** -24
    assert list(oracle.bf16_digits(x, -1, 4)) == [24, 3, -128, 128]
    assert list(oracle.bf16_digits(x, -1, 5)) == [24, 3, -128, 128, 0]
    rng = np.random.default_rng(5)
    for v in rng.uniform(-0.5, 0.5, 200):
        for s in (3, 7, 16):
            d = oracle.bf16_digits(v, 0, s)
            assert np.all(np.abs(d) <= 128)
            total = sum(Fraction(int(dt), 256 ** (t + 1)) for t, dt in enumerate(d))
            assert abs(Fraction(v) - total) <= Fraction(1, 2 * 256 ** s)


def _two_digit_integers(rng, rows, cols):
    """Integers of magnitude < 2^12 with a full-range entry in every row and
    column: scaled by 2^-13 each is x / 8192 = d1 / 256 + d2 / 65536 with
    d1 = rint(x / 32) and an integer d2 (x - 32 d1 is an integer of
    magnitude <= 16): exactly two digits."""
    x = rng.integers(-(2 ** 12) + 1, 2 ** 12, (rows, cols)).astype(float)
    x[:, 0] = 2 ** 12 - 1
    x[0, :] = -(2 ** 12) + 1
    return np.asfortranarray(x)


def test_bf16_emulation_exact_on_two_digit_integers():
    """Two-digit entries need the digit pairs of levels 2..4 (t + u <= 4):
    with 3 slices every needed pair is present and C = A B exactly (every
    partial sum is a multiple of 2^-32 below 2^9); with 2 slices the pair
    (2, 2) of level 4 is left out and C is not exact."""
    rng = np.random.default_rng(11)
    m, n, k = 23, 19, 2600  # three k chunks, the last ragged
    a = _two_digit_integers(rng, m, k)
    b = _two_digit_integers(rng, k, n)
    want = (a.astype(np.int64) @ b.astype(np.int64)).astype(float)
    for s, exact in ((3, True), (5, True), (2, False)):
        c = np.zeros((m, n), order="F")
        oracle.dgemm_ozaki_bf16("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, s)
        assert np.array_equal(c, want) == exact, s
    # The transposed operand forms give the same product.
    at, bt = np.asfortranarray(a.T), np.asfortranarray(b.T)
    c = np.zeros((m, n), order="F")
    oracle.dgemm_ozaki_bf16("T", "T", m, n, k, 1.0, at, k, bt, n, 0.0, c, m, 3)
    assert np.array_equal(c, want)


def test_planted_bf16_level_mistake_is_caught():
    oracle.plant_mistake(8)
    with pytest.raises(AssertionError):
        test_bf16_emulation_exact_on_two_digit_integers()


def test_bf16_emulation_error_falls_with_slices():
    """Against the double-double reference on uniform operands: the
    componentwise error (relative to |alpha| |A| |B| + |beta| |C|) falls by
    about 2^-8 per slice until it reaches FP64 rounding; 7 slices are within
    native FP64's typical sqrt(k) u, 5 are not, and the native worst-case
    bound 4 (k + 2) u holds from 7 on."""
    m, n, k = 48, 40, 700
    a = oracle.fill_uniform(m, k, 21)
    b = oracle.fill_uniform(k, n, 22)
    c = oracle.fill_uniform(m, n, 23)
    exp, bnd = oracle.gemm_ref("N", "N", m, n, k, 1.5, a, m, b, k, -0.5, c, m)
    eps = np.finfo(float).eps
    errs = {}
    for s in range(3, 10):
        cc = c.copy(order="F")
        oracle.dgemm_ozaki_bf16("N", "N", m, n, k, 1.5, a, m, b, k, -0.5, cc, m, s)
        errs[s] = np.max(np.abs(cc - exp) / bnd)
    for s in range(3, 6):
        assert errs[s + 1] <= errs[s] * 2.0 ** -6, (s, errs)
    assert errs[7] <= np.sqrt(k) * eps and errs[5] > np.sqrt(k) * eps, errs
    for s in range(7, 10):
        assert errs[s] <= 4 * (k + 2) * eps


def test_bf16_emulation_quick_returns_and_beta():
    m, n, k = 5, 4, 3
    a = oracle.fill_uniform(m, k, 1)
    b = oracle.fill_uniform(k, n, 2)
    c = np.full((m, n), np.nan, order="F")
    oracle.dgemm_ozaki_bf16("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, 7)
    assert np.isfinite(c).all()  # beta = 0: C is not read
    c2 = np.full((m, n), 2.0, order="F")
    nan = np.full_like(a, np.nan)
    oracle.dgemm_ozaki_bf16("N", "N", m, n, k, 0.0, nan, m, nan, k, 3.0, c2, m, 7)
    assert np.all(c2 == 6.0)  # alpha = 0: C = beta C, A and B unread


# ---------------------------------------------------------------------------
# The cost model's CPU restatement (oracle/mlp.py)

def test_mlp_forward_by_hand():
    """A sparse parameter set whose output is worked out by hand."""
    from oracle import mlp
    outs = 4
    p = np.zeros(mlp.num_params(outs))
    w = {name: (off, shape) for name, off, shape in mlp.offsets(outs)}

    def put(name, idx, v):
        off, shape = w[name]
        p[off + np.ravel_multi_index(idx, shape)] = v
    put("W1", (0, 0), 2.0)     # h1[0] = relu(2 * 0.5) = 1
    put("W1", (1, 3), -1.0)    # h1[1] = relu(-1) = 0
    put("b1", (2,), 0.25)      # h1[2] = 0.25
    put("W2", (0, 0), 3.0)     # h2[0] = 3 * 1 + 4 * 0.25 = 4
    put("W2", (0, 2), 4.0)
    put("W2", (1, 1), 5.0)     # h2[1] = relu(5 * 0 - 1) = 0
    put("b2", (1,), -1.0)
    put("W3", (5, 0), 0.5)     # h3[5] = 2
    put("W3", (6, 1), 9.0)     # h3[6] = 0
    put("W4", (2, 5), 1.5)     # out[2] = 1.5 * 2 + 0.125
    put("b4", (2,), 0.125)
    put("b4", (3,), -7.0)
    x = np.array([[0.5, 0.0, 0.0, 1.0, 0.0, 1.0]])
    out, (_, h1, h2, h3) = mlp.forward(p, outs, x)
    assert np.array_equal(out[0], [0.0, 0.0, 3.125, -7.0])
    assert (h1[0, 0], h1[0, 1], h1[0, 2]) == (1.0, 0.0, 0.25)
    assert (h2[0, 0], h2[0, 1], h3[0, 5], h3[0, 6]) == (4.0, 0.0, 2.0, 0.0)


def _mlp_case(seed=4, outs=5, n=7):
    from oracle import mlp
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp.num_params(outs))
    x = rng.uniform(0, 1, (n, mlp.K_IN))
    cfg = rng.integers(0, outs, n)
    y = rng.normal(1.0, 0.5, n)
    return p, outs, x, cfg, y


def _fd_check(grad_fn, loss_fn, p, idx, h=1e-6):
    worst = 0.0
    g = grad_fn(p)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (loss_fn(p + e) - loss_fn(p - e)) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / (abs(fd) + abs(g[i]) + 1e-7))
    return worst


def test_mlp_gradient_matches_finite_differences():
    from oracle import mlp
    p, outs, x, cfg, y = _mlp_case()
    idx = np.random.default_rng(1).choice(len(p), 400, replace=False)
    # Every block is covered, including the biases of the last layer.
    idx = np.concatenate([idx, np.arange(len(p) - outs, len(p))])
    worst = _fd_check(lambda q: mlp.loss_and_grad(q, outs, x, cfg, y)[1],
                      lambda q: mlp.loss(q, outs, x, cfg, y), p, idx)
    assert worst < 2e-5, worst
    assert np.isclose(mlp.loss_and_grad(p, outs, x, cfg, y)[0], mlp.loss(p, outs, x, cfg, y))


def test_mlp_gradient_check_catches_a_dropped_relu_mask():
    """A backward pass that forgets one ReLU derivative fails the same check."""
    from oracle import mlp
    p, outs, x, cfg, y = _mlp_case()

    def wrong_grad(q):
        w = mlp.unpack(q, outs)
        out, (xx, h1, h2, h3) = mlp.forward(q, outs, x)
        n = len(cfg)
        d4 = np.zeros_like(out)
        d4[np.arange(n), cfg] = (out[np.arange(n), cfg] - y) / n
        d3 = (d4 @ w["W4"]) * (h3 > 0)
        d2 = d3 @ w["W3"]  # mask of layer 2 dropped
        d1 = (d2 @ w["W2"]) * (h1 > 0)
        g = {"W4": d4.T @ h3, "b4": d4.sum(0), "W3": d3.T @ h2, "b3": d3.sum(0),
             "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ xx, "b1": d1.sum(0)}
        return np.concatenate([g[name].ravel() for name, _, _ in mlp.offsets(outs)])
    names = {name: (off, shape) for name, off, shape in mlp.offsets(outs)}
    off, shape = names["b2"]
    worst = _fd_check(wrong_grad, lambda q: mlp.loss(q, outs, x, cfg, y), p,
                      np.arange(off, off + shape[0]))
    assert worst > 1e-3


def test_mlp_features_match_library():
    """The library's host feature function and the restatement agree."""
    import ctypes
    from oracle import mlp
    import hwyblas_amd as H
    buf = (ctypes.c_float * 6)()
    for args in [(8, 0, 1, 8, 512, 77), (4, 1, 0, 1, 300, 2), (8, 1, 1, 512, 512, 512)]:
        H.lib().hwyblas_grouped_features(*args, buf)
        assert np.allclose(np.array(buf[:]), mlp.features(*args), rtol=1e-6, atol=1e-7)


# ---------------------------------------------------------------------------
# LU with partial pivoting (oracle/lu.cc)

def lu_fixture_cases():
    for row in FIX["lu_cases"]:
        key, m, n = str(row).split(",")
        yield key, int(m), int(n)


def test_getrf_matches_lapack_fixtures():
    """Same pivots as LAPACK's dgetrf and the same factors to rounding; the
    zero column gives info = its index."""
    for key, m, n in lu_fixture_cases():
        a = _f(FIX[key + "_a"])
        lu, ipiv, info = oracle.getrf(a)
        assert np.array_equal(ipiv, FIX[key + "_ipiv"]), key
        ref = FIX[key + "_lu"]
        assert np.max(np.abs(lu - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), key
        assert info == (5 if key == "lu9" else 0), (key, info)


def test_getrs_matches_lapack_fixtures():
    for key, m, n in lu_fixture_cases():
        if key + "_b" not in FIX:
            continue
        lu, ipiv, _ = oracle.getrf(_f(FIX[key + "_a"]))
        for trans, name in (("N", "_x"), ("T", "_xt")):
            x = oracle.getrs(trans, lu, ipiv, FIX[key + "_b"])
            ref = FIX[key + name]
            assert np.max(np.abs(x - ref)) <= 1e-10 * max(1.0, np.max(np.abs(ref))), (key, trans)


def test_getrf_residual_criterion_and_planted_swap_mistake():
    """||P L U - A||_1 / (n ||A||_1 eps) < 30 (hwyBLAS tests/getrf_test.cc:42-96),
    at a size that runs the blocked path; a leaf that does not swap the
    columns left of the pivot column fails it."""
    rng = np.random.default_rng(8)
    n = 200
    a = _f(rng.standard_normal((n, n)))
    eps = np.finfo(np.float64).eps

    def ratio():
        lu, ipiv, _ = oracle.getrf(a)
        P, L, U = oracle.lu_factors(lu, ipiv)
        return np.abs(P @ L @ U - a).sum(0).max() / (n * np.abs(a).sum(0).max() * eps)
    assert ratio() < 30
    oracle.plant_mistake(7)
    try:
        assert ratio() > 1e6
    finally:
        oracle.plant_mistake(0)


def test_gerfs_berr_closed_forms():
    """For A = I: berr_j = max_i |b_ij - x_ij| / (|x_ij| + |b_ij|); an exact
    solution of an integer system has berr 0."""
    rng = np.random.default_rng(3)
    n = 9
    b = rng.uniform(-1, 1, (n, 2))
    x = b + rng.uniform(-1e-3, 1e-3, (n, 2))
    berr = oracle.gerfs_berr(np.eye(n), b, x)
    assert np.allclose(berr, np.max(np.abs(b - x) / (np.abs(x) + np.abs(b)), axis=0), rtol=1e-14)
    a = np.triu(rng.integers(1, 4, (n, n))).astype(float)
    xe = rng.integers(-5, 5, (n, 2)).astype(float)
    assert np.array_equal(oracle.gerfs_berr(a, a @ xe, xe), [0.0, 0.0])


# ---- xLACN2 and xGERFS's forward error bound --------------------------------

def _lacn2(m):
    return oracle.lacn2(m.shape[0], lambda x: m @ x, lambda x: m.T @ x)


def test_lacn2_known_answers():
    """n = 1: |m|. Nonnegative (nonsymmetric) M: the sign vector is all
    ones, M^T 1 holds the column sums, so the estimate is the exact 1-norm
    after M x0, M^T 1, M e_j and the alternating vector (4 products); a
    diagonal with mixed signs: exact max |d|."""
    est, v, calls = _lacn2(np.array([[-3.5]]))
    assert est == 3.5 and calls == 1 and v[0] == -3.5
    m = np.array([[1.0, 7, 0, 2], [0, 1, 5, 0], [3, 0, 1, 1], [0, 0, 4, 2]])
    est, v, calls = _lacn2(m)
    assert est == np.abs(m).sum(0).max() == 10.0 and calls == 4
    assert np.array_equal(v, m[:, 2])  # the column the estimate came from
    d = np.diag([2.0, -9.0, 4.0, -1.0, 3.0])
    assert _lacn2(d)[0] == 9.0


def test_lacn2_alternating_stage_by_hand():
    """M = [[-1,0,-4],[0,3,-1],[4,-3,2]]: M x0 = (-5/3, 2/3, 1) (est 10/3),
    M^T sign = (5, 0, 5) -> j = 1, M e_1 = (-1, 0, 4) (est 5, same signs:
    converged), then b = (1, -1.5, 2), M b = (-9, -6.5, 12.5) and
    2 |M b|_1 / (3 n) = 56/9 > 5. Without that stage the estimate is 5."""
    m = np.array([[-1.0, 0, -4], [0, 3, -1], [4, -3, 2]])
    est, v, calls = _lacn2(m)
    assert abs(est - 56 / 9) <= 1e-15 and calls == 4
    assert np.array_equal(v, [-9.0, -6.5, 12.5])
    oracle.plant_mistake(9)
    assert _lacn2(m)[0] == 5.0


def test_lacn2_lower_bound_on_random_matrices():
    """Every estimate is |M w|_1 / |w|_1 for some w, so it never exceeds
    ||M||_1; on Gaussian matrices it stays within a factor 2 of it."""
    rng = np.random.default_rng(5)
    ratios = []
    for _ in range(200):
        n = int(rng.integers(2, 40))
        m = rng.standard_normal((n, n))
        est, v, calls = _lacn2(m)
        assert calls <= 11
        ratios.append(est / np.abs(m).sum(0).max())
    ratios = np.array(ratios)
    assert ratios.max() <= 1 + 1e-14 and ratios.min() > 0.5


@pytest.mark.parametrize("trans", ["N", "T"])
def test_gerfs_forward_bound_closed_form(trans):
    """Exact solutions of permuted triangular integer systems: berr = 0 and
    ferr equals the bound's defining formula (inv(op(A)) >= 0, so the
    estimated matrix is nonnegative and the estimate exact)."""
    for n in (1, 2, 6, 40, 300):
        a, xt, b = permuted_triangular_system(n, 3, n, trans)
        lu, piv, info = oracle.getrf(a)
        assert info == 0
        x0 = oracle.getrs(trans, lu, piv, b)
        assert np.array_equal(x0, xt)
        x, ferr, berr = oracle.gerfs(trans, a, lu, piv, b, x0)
        assert np.array_equal(x, xt) and np.all(berr == 0)
        want = forward_bound(a if trans == "N" else a.T, xt, b)
        assert np.allclose(ferr, want, rtol=1e-12, atol=0), (n, ferr, want)


@pytest.mark.parametrize("mistake", [10, 11])
def test_planted_ferr_mistakes_are_caught(mistake):
    oracle.plant_mistake(mistake)
    off = []
    for trans in "NT":
        a, xt, b = permuted_triangular_system(40, 3, 40, trans)
        lu, piv, _ = oracle.getrf(a)
        _, ferr, _ = oracle.gerfs(trans, a, lu, piv, b, xt)
        off.append(np.abs(ferr / forward_bound(a if trans == "N" else a.T, xt, b) - 1).max())
    assert max(off) > 1e-3


@pytest.mark.parametrize("trans", ["N", "T"])
def test_gerfs_refinement_criteria(trans):
    """hwyBLAS's refinement test (tests/lapack/linear_test.cc:498-560), with
    exact integer solutions: X = X_true (1 + 1e-6 u) is refined to about
    the rounding error and its error bounded by ferr; at B scale 1,
    berr < 4 (n + 1) eps and ferr < 1e3 n eps; at 2^-982 safe1 dominates
    and ferr >= 0.3 safe1 / (||op(A)||inf ||X||inf)."""
    e = np.finfo(np.float64).eps / 2
    rng = np.random.default_rng(19)
    for n in (1, 6, 40):
        for bscale in (1.0, 2.0 ** -982):
            a = rng.integers(-5, 6, (n, n)).astype(float) + np.diag(np.full(n, 6.0 * n))
            opa = a if trans == "N" else a.T
            xt = rng.integers(-50, 51, (n, 2)).astype(float) * bscale
            b = opa @ xt
            lu, piv, _ = oracle.getrf(a)
            x0 = xt * (1 + 1e-6 * rng.uniform(-1, 1, xt.shape))
            x, ferr, berr = oracle.gerfs(trans, a, lu, piv, b, x0)
            for j in range(2):
                xn = np.abs(x[:, j]).max()
                err = np.abs(x[:, j] - xt[:, j]).max()
                assert err / xn < 100 * n * e
                assert ferr[j] >= 0.3 * err / xn
                if bscale == 1.0:
                    assert berr[j] < 4 * (n + 1) * e and ferr[j] < 1e3 * n * e
                else:
                    safe1 = (n + 1) * np.finfo(np.float64).tiny
                    assert ferr[j] >= 0.3 * safe1 / (np.abs(opa).sum(1).max() * xn)


def test_gerfs_quick_return_and_bad_trans():
    a = np.eye(3)
    lu, piv, _ = oracle.getrf(a)
    x, ferr, berr = oracle.gerfs("N", a, lu, piv, np.zeros((3, 0)), np.zeros((3, 0)))
    assert ferr.size == 0 and berr.size == 0
    with pytest.raises(ValueError):
        oracle.gerfs("X", a, lu, piv, np.ones((3, 1)), np.ones((3, 1)))


# ---- xPOTRF2 ------------------------------------------------------------------

def _unit_lower(n, seed):
    rng = np.random.default_rng(seed)
    low = np.tril(rng.integers(-3, 4, (n, n)).astype(float), -1) + np.eye(n)
    return low


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrf2_exact_on_unit_triangular_products(uplo):
    """A = L L^T with an integer unit lower L: every pivot is 1 and every
    step exact, so the factor is L itself (U = L^T for uplo U); the other
    triangle is untouched."""
    for n in (1, 2, 5, 17, 40):
        low = _unit_lower(n, n)
        a = low @ low.T
        f, info = oracle.potrf2(uplo, a)
        assert info == 0
        tri = np.tril if uplo == "L" else np.triu
        assert np.array_equal(tri(f), low if uplo == "L" else low.T)
        other = np.triu(np.ones((n, n), dtype=bool), 1) if uplo == "L" else \
            np.tril(np.ones((n, n), dtype=bool), -1)
        assert np.array_equal(f[other], a[other])


def test_potrf2_against_numpy_and_info():
    rng = np.random.default_rng(3)
    n = 60
    m = rng.standard_normal((n, n))
    a = m @ m.T + n * np.eye(n)
    f, info = oracle.potrf2("L", a)
    assert info == 0
    ref = np.linalg.cholesky(a)
    assert np.max(np.abs(np.tril(f) - ref)) <= 1e-12 * np.max(np.abs(ref))
    d = np.diag([4.0, 9.0, -1.0, 16.0])
    assert oracle.potrf2("L", d)[1] == 3
    d[2, 2] = np.nan
    assert oracle.potrf2("U", d)[1] == 3
    assert oracle.potrf2("L", np.zeros((0, 0)))[1] == 0


def test_planted_potrf_mistake_is_caught():
    oracle.plant_mistake(12)
    low = _unit_lower(17, 17)
    f, info = oracle.potrf2("L", low @ low.T)
    assert info != 0 or not np.array_equal(np.tril(f), low)


# ---- the checkers themselves, pinned in both directions -------------------

@pytest.mark.parametrize("side,uplo,ta", list(itertools.product("LR", "LU", "NT")))
def test_trsm_residual_rejects_a_perturbed_solution(side, uplo, ta):
    """The residual checker behind every TRSM parity test accepts the solved
    X and rejects the same X with one entry moved by 100 times the bound (in
    the row whose residual that entry drives through the diagonal)."""
    m, n = 60, 23
    a, lda, b, ldb = trsm_operands(side, uplo, "N", m, n, np.float64, 91)
    x = b.copy(order="F")
    oracle.trsm(side, uplo, ta, "N", m, n, 1.0, a, lda, x, ldb)
    assert trsm_residual_ratio(side, uplo, ta, "N", m, n, 1.0, a, lda, b, ldb, x) <= 1.0
    ka = m if side == "L" else n
    tol = 8.0 * (ka + 2) * np.finfo(np.float64).eps
    p, w = 17, 5  # X(p, w) for the left side, X(w, p) for the right
    dense = np.where(np.isnan(a[:ka, :ka]), 0.0, a[:ka, :ka])
    opa = dense if ta == "N" else dense.T
    if side == "L":
        scale = np.abs(opa[p]) @ np.abs(x[:m, w]) + abs(b[p, w])
        delta = 100 * tol * scale / abs(opa[p, p])
        y = x.copy(order="F")
        y[p, w] += delta
    else:
        scale = np.abs(x[w, :n]) @ np.abs(opa[:, p]) + abs(b[w, p])
        delta = 100 * tol * scale / abs(opa[p, p])
        y = x.copy(order="F")
        y[w, p] += delta
    assert trsm_residual_ratio(side, uplo, ta, "N", m, n, 1.0, a, lda, b, ldb, y) > 50.0


def test_trsm_residual_rejects_the_skipped_reciprocal():
    """Planted mistake 4 (the base kernel skips the diagonal's reciprocal)
    is caught by the residual checker alone, with no fixture involved."""
    for side, uplo, ta in itertools.product("LR", "LU", "NT"):
        m, n = 70, 19
        a, lda, b, ldb = trsm_operands(side, uplo, "N", m, n, np.float64, 92)
        x = b.copy(order="F")
        oracle.plant_mistake(4)
        try:
            oracle.trsm(side, uplo, ta, "N", m, n, 1.0, a, lda, x, ldb)
        finally:
            oracle.plant_mistake(0)
        assert trsm_residual_ratio(side, uplo, ta, "N", m, n, 1.0, a, lda, b, ldb, x) > 1.0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_entry_reference_equals_the_full_reference(ta, tb, dtype):
    """gemm_ref_entries (behind every large-shape test, smoke and the bench)
    gives, entry for entry, the same expected value and error scale as the
    full double-double reference gemm_ref."""
    m, n, k = 37, 29, 301
    a = oracle.fill_uniform(m if ta == "N" else k, k if ta == "N" else m, 93, dtype=dtype)
    b = oracle.fill_uniform(k if tb == "N" else n, n if tb == "N" else k, 94, dtype=dtype)
    c = oracle.fill_uniform(m, n, 95, dtype=dtype)
    for beta in (0.0, -0.75):
        exp, bnd = oracle.gemm_ref(ta, tb, m, n, k, 1.25, a, a.shape[0], b, b.shape[0], beta,
                                   c, m)
        rng = np.random.default_rng(96)
        rows = rng.integers(0, m, 200)
        cols = rng.integers(0, n, 200)
        e2, b2 = oracle.gemm_ref_entries(ta, tb, k, 1.25, a, a.shape[0], b, b.shape[0], beta, c,
                                         m, rows, cols)
        assert np.array_equal(e2, exp[rows, cols])
        assert np.array_equal(b2, bnd[rows, cols])


# ---------------------------------------------------------------------------
# The count choice of the emulated DGEMMs (oracle/emulated_choice.cc)
#
# The decision rests on three facts, each pinned here by exact arithmetic:
# (1) the coarse product bounds |A||B| from below and the exact sums bound
# the row/column sums from above; (2) the per-entry error of the emulation
# is within the analytic formula (for INT8 with bA + bB bits:
# 2^(E_i - bA - 1) sum_p |b_pj| + 2^(F_j - bB - 1) sum_p |a_ip|
# + k 2^(E_i + F_j - bA - bB - 2)), which adversarial operands attain within
# 1 %; (3) the decision's bound times the lower bound covers the formula on
# every pair. Planted mistakes 13 (a dropped sum term) and 14 (coarse
# magnitudes rounded to nearest) each break one of them.

def _exact_abs_product(a, b):
    m, k = a.shape
    n = b.shape[1]
    fa = [[Fraction(abs(float(a[i, p]))) for p in range(k)] for i in range(m)]
    fb = [[Fraction(abs(float(b[p, j]))) for j in range(n)] for p in range(k)]
    return [[sum(fa[i][p] * fb[p][j] for p in range(k)) for j in range(n)] for i in range(m)]


def _exact_product(a, b):
    m, k = a.shape
    n = b.shape[1]
    return [[sum(Fraction(float(a[i, p])) * Fraction(float(b[p, j])) for p in range(k))
             for j in range(n)] for i in range(m)]


def _choice_operands():
    rng = np.random.default_rng(21)
    m, n, k = 7, 6, 40
    yield m, n, k, oracle.fill_uniform(m, k, 1), oracle.fill_uniform(k, n, 2)
    yield m, n, k, oracle.fill_pow2(m, k, 3, -20, 20), oracle.fill_pow2(k, n, 4, -20, 20)
    a = oracle.fill_uniform(m, k, 5) * (10.0 ** rng.uniform(-8, 8, m))[:, None]
    b = oracle.fill_uniform(k, n, 6) * (10.0 ** rng.uniform(-8, 8, n))[None, :]
    a[2, :] = 0.0  # a zero row
    yield m, n, k, np.asfortranarray(a), np.asfortranarray(b)


def test_emulated_coarse_statistics_bound_exactly():
    for m, n, k, a, b in _choice_operands():
        lower, rs, cs = oracle.emulated_coarse("N", "N", m, n, k, a, m, b, k)
        exact = _exact_abs_product(a, b)
        for i in range(m):
            for j in range(n):
                assert Fraction(float(lower[i, j])) <= exact[i][j], (i, j)
        for i in range(m):
            s = sum(Fraction(abs(float(x))) for x in a[i])
            assert Fraction(float(rs[i])) >= s
            assert float(rs[i]) <= float(s) * (1 + 1e-9) + k * 2.0 ** -39 * float(np.max(np.abs(a[i])))
        for j in range(n):
            s = sum(Fraction(abs(float(x))) for x in b[:, j])
            assert Fraction(float(cs[j])) >= s
    # Uniform operands: the lower bound is close (7-bit magnitudes).
    m, n, k, a, b = next(_choice_operands())
    lower = oracle.emulated_coarse("N", "N", m, n, k, a, m, b, k)[0]
    exact = np.array([[float(x) for x in row] for row in _exact_abs_product(a, b)])
    assert np.all(lower >= 0.9 * exact)


def test_planted_coarse_rounding_mistake_is_caught():
    oracle.plant_mistake(14)
    with pytest.raises(AssertionError):
        test_emulated_coarse_statistics_bound_exactly()


def _int8_formula(a, b, nmod):
    """The per-entry error formula (exact sums), as floats, m x n."""
    m, k = a.shape
    n = b.shape[1]
    bits = max(oracle.lib().orc_ozaki_total_bits(nmod, k), 2)
    ba, bb = bits // 2, bits - bits // 2
    E = [np.frexp(np.max(np.abs(a[i])))[1] for i in range(m)]
    F = [np.frexp(np.max(np.abs(b[:, j])))[1] for j in range(n)]
    rows = [float(sum(Fraction(abs(float(x))) for x in a[i])) for i in range(m)]
    cols = [float(sum(Fraction(abs(float(x))) for x in b[:, j])) for j in range(n)]
    out = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            out[i, j] = (2.0 ** (E[i] - ba - 1) * cols[j] + 2.0 ** (F[j] - bb - 1) * rows[i]
                         + k * 2.0 ** (E[i] + F[j] - bits - 2))
    return out, ba, bb


def _emulated_error(a, b, nmod):
    m, k = a.shape
    n = b.shape[1]
    c = np.zeros((m, n), order="F")
    oracle.dgemm_ozaki("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, nmod)
    exact = _exact_product(a, b)
    return np.array([[abs(float(Fraction(float(c[i, j])) - exact[i][j])) for j in range(n)]
                     for i in range(m)])


def _half_integer_operands(k, rough_a):
    """With 5 moduli at k = 64 (bA = bB = 16): entries whose scaled values sit
    just below x.5 on one side (every rounding error about -1/2, aligned) and
    exactly representable values on the other, so the formula's active term
    is attained."""
    c = 0.75 + 2.0 ** -17 - 2.0 ** -36  # 0.75 2^16 + 1/2 - 2^-20
    m, n = 3, 2
    if rough_a:
        return np.full((m, k), c, order="F"), np.ones((k, n), order="F")
    return np.ones((m, k), order="F"), np.full((k, n), c, order="F")


def test_emulated_error_formula_holds_and_is_attained():
    u = 2.0 ** -53
    for m, n, k, a, b in _choice_operands():
        for nmod in (3, 5, 8):
            formula, _, _ = _int8_formula(a, b, nmod)
            err = _emulated_error(a, b, nmod)
            absab = np.array([[float(x) for x in row] for row in _exact_abs_product(a, b)])
            assert np.all(err <= formula + 32 * u * absab), nmod
    k = 64
    assert oracle.lib().orc_ozaki_total_bits(5, k) == 32
    for rough_a in (True, False):
        a, b = _half_integer_operands(k, rough_a)
        formula, _, _ = _int8_formula(a, b, 5)
        err = _emulated_error(a, b, 5)
        assert np.all(err <= formula)
        # The rounding side's term: 2^(E - bA - 1) k (rough A: E = 0, B = 1)
        # or 2^(F - bB - 1) k (rough B: F = 0, A = 1); both 2^-17 k.
        term = 2.0 ** -17 * k
        assert np.all(err >= 0.99 * term), (rough_a, err.max(), term)


def test_emulated_decision_bound_covers_the_formula():
    """worst(N) x lower_ij >= the per-entry formula on every pair, for the
    statistics X, gmin the choice computed."""
    for m, n, k, a, b in _choice_operands():
        count, x, gmin = oracle.emulated_choose("int8", "N", "N", m, n, k, a, m, b, k)
        lower = oracle.emulated_coarse("N", "N", m, n, k, a, m, b, k)[0]
        for nmod in (3, 8, 14, 20):
            formula, _, _ = _int8_formula(a, b, nmod)
            worst = oracle.emulated_bound("int8", nmod, k, x, gmin)[1]
            if not np.isfinite(worst):  # some pair cannot be certified: native
                assert count == 0 and not np.all(lower > 0)
                continue
            nz = (np.abs(a).sum(axis=1)[:, None] > 0) & (np.abs(b).sum(axis=0)[None, :] > 0)
            assert np.all(formula[nz] <= worst * lower[nz] * (1 + 1e-12)), nmod
        # The rule: the smallest count meeting the targets (restated here).
        u = 2.0 ** -53
        t_typ, t_bound = 2 * np.sqrt(k) * u, (4 * (k + 2) - 24) * u
        pick = 0
        for c in range(2, 25):
            typ, worst = oracle.emulated_bound("int8", c, k, x, gmin)
            if typ <= t_typ and worst <= t_bound:
                pick = c
                break
        if pick == 0:
            for c in range(2, 25):
                if oracle.emulated_bound("int8", c, k, x, gmin)[1] <= t_bound:
                    pick = c
                    break
        assert count == pick


def test_planted_dropped_sum_term_mistake_is_caught():
    oracle.plant_mistake(13)
    with pytest.raises(AssertionError):
        test_emulated_decision_bound_covers_the_formula()


def test_emulated_choice_special_cases():
    m, n, k = 5, 4, 30
    a = oracle.fill_uniform(m, k, 1)
    b = oracle.fill_uniform(k, n, 2)
    # Forced counts are clamped, NaN/Inf give 0.
    assert oracle.emulated_choose("int8", "N", "N", m, n, k, a, m, b, k, forced=30)[0] == 24
    assert oracle.emulated_choose("bf16", "N", "N", m, n, k, a, m, b, k, forced=30)[0] == 16
    a2 = a.copy(order="F")
    a2[1, 1] = np.inf
    assert oracle.emulated_choose("int8", "N", "N", m, n, k, a2, m, b, k)[0] == 0
    assert oracle.emulated_choose("int8", "N", "N", m, n, k, a2, m, b, k, forced=9)[0] == 0
    # All-zero op(A): every pair is exact, the smallest count.
    z = np.zeros((m, k), order="F")
    assert oracle.emulated_choose("int8", "N", "N", m, n, k, z, m, b, k)[0] == 2
    # Disjoint supports: |A||B| = 0 where the rows and columns are not zero.
    a3 = a.copy(order="F")
    b3 = b.copy(order="F")
    a3[:, 15:] = 0.0
    b3[:15, :] = 0.0
    assert oracle.emulated_choose("int8", "N", "N", m, n, k, a3, m, b3, k)[0] == 0
    # k too small for the guarantee to leave room: native.
    assert oracle.emulated_choose("int8", "N", "N", m, n, 3, a, m, b, 3)[0] == 0


# ---------------------------------------------------------------------------
# Householder QR (oracle/qr.cc)

def qr_fixture_cases():
    for line in FIX["qr_cases"]:
        key, m, n = str(line).split(",")
        yield key, int(m), int(n)


def explicit_q(h, tau):
    """Q = H(0) H(1) ... H(k-1) formed densely from the stored reflectors."""
    m, n = h.shape
    q = np.eye(m)
    for i in reversed(range(min(m, n))):
        v = np.zeros(m)
        v[i] = 1.0
        v[i + 1:] = h[i + 1:, i]
        q -= tau[i] * np.outer(v, v @ q)
    return q


def test_qr_matches_lapack_fixtures():
    """Against LAPACK dgeqrf (scipy): the same R, reflectors and tau up to
    rounding (same signs: beta = -sign(alpha) |(alpha, x)|), blocked
    (k > 128) and unblocked, with a zero column and a column already zero
    below the diagonal (tau = 0 there)."""
    eps = np.finfo(float).eps
    for key, m, n in qr_fixture_cases():
        a, hs, ts = FIX[key + "_a"], FIX[key + "_h"], FIX[key + "_tau"]
        h, tau = oracle.geqrf(a)
        scale = max(1.0, float(np.max(np.abs(hs))))
        assert np.max(np.abs(h - hs)) <= 64 * max(m, n) * eps * scale, key
        assert np.max(np.abs(tau - ts), initial=0.0) <= 64 * max(m, n) * eps, key
    h, tau = oracle.geqrf(FIX["qr9_a"])
    assert tau[0] == 0.0


def test_qr_orthogonality_and_backward_error():
    eps = np.finfo(float).eps
    rng = np.random.default_rng(3)
    for (m, n) in ((300, 300), (400, 170), (170, 400)):
        a = rng.standard_normal((m, n))
        h, tau = oracle.geqrf(a)
        q = explicit_q(h, tau)
        r = np.triu(h[:min(m, n), :])
        assert np.max(np.abs(q.T @ q - np.eye(m))) <= 8 * m * eps
        resid = q[:, :min(m, n)] @ r - a
        assert np.linalg.norm(resid, 1) <= 8 * m * eps * np.linalg.norm(a, 1)
        # Blocked and unblocked agree to rounding.
        h2, tau2 = oracle.geqrf(a, blocked=False)
        assert np.max(np.abs(h - h2)) <= 64 * m * eps * np.max(np.abs(h))


def test_larft_forms_the_block_reflector():
    """I - V T V^T equals H(0) ... H(k-1) formed one reflector at a time."""
    rng = np.random.default_rng(4)
    for (n, k) in ((50, 1), (50, 7), (80, 32), (33, 33)):
        h, tau = oracle.geqrf(rng.standard_normal((n, k)), blocked=False)
        t = oracle.larft(h, tau)
        v = np.tril(h, -1)[:, :k] + np.eye(n, k)
        assert np.allclose(np.triu(t), t)
        assert np.max(np.abs((np.eye(n) - v @ t @ v.T) - explicit_q(h, tau))) <= 1e-13


def test_planted_qr_sign_mistake_is_caught():
    oracle.plant_mistake(15)
    with pytest.raises(AssertionError):
        test_qr_matches_lapack_fixtures()
