"""The grouped-GEMM cost model on the GPU: the hand-written forward,
backward and update kernels against the CPU restatement (oracle/mlp.py),
training, and the configuration choice that uses the model."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T(H):
    from hwyblas_amd import tune
    return tune


def dataset(rng, n, outs):
    x = rng.uniform(0, 1, (n, omlp.K_IN)).astype(np.float32)
    x[:, 3:] = rng.integers(0, 2, (n, 3))
    cfg = rng.integers(0, outs, n).astype(np.int32)
    y = (3 * x[:, 0] + x[:, 1] - 2 * x[:, 2] + 0.5 * x[:, 3] + 0.25 * cfg - 1).astype(np.float32)
    return x, cfg, y


@pytest.mark.parametrize("n", [1, 37, 300])
def test_gradients_match_oracle(T, n):
    """Loss and every parameter's gradient from the HIP kernels agree with
    the float64 restatement at the same FP32 parameters."""
    outs = 14
    model = T.Model(outs, seed=3)
    p = model.params().astype(np.float64)
    rng = np.random.default_rng(n)
    x, cfg, y = dataset(rng, n, outs)
    loss, g = model.gradients(x, cfg, y)
    eloss, eg = omlp.loss_and_grad(p, outs, x.astype(np.float64), cfg, y.astype(np.float64))
    assert np.isclose(loss, eloss, rtol=1e-4, atol=1e-6)
    for name, off, shape in omlp.offsets(outs):
        size = int(np.prod(shape))
        got, exp = g[off:off + size], eg[off:off + size]
        scale = max(np.abs(exp).max(), 1e-6)
        assert np.abs(got - exp).max() <= 2e-5 * scale, name


def test_predict_matches_oracle(T):
    outs = 14
    model = T.Model(outs, seed=5)
    p = model.params().astype(np.float64)
    x, _, _ = dataset(np.random.default_rng(2), 1000, outs)
    got = model.predict(x)
    exp, _ = omlp.forward(p, outs, x.astype(np.float64))
    assert np.abs(got - exp).max() <= 1e-5 * max(1.0, np.abs(exp).max())


def test_set_params_round_trip_and_training_reduces_loss(T):
    outs = 14
    model = T.Model(outs, seed=9)
    p0 = model.params()
    model.set_params(p0)
    assert np.array_equal(model.params(), p0)
    x, cfg, y = dataset(np.random.default_rng(4), 4000, outs)
    start = omlp.loss(p0.astype(np.float64), outs, x, cfg, y)
    losses = model.train(x, cfg, y, epochs=60, batch=128, lr=3e-3)
    end = omlp.loss(model.params().astype(np.float64), outs, x, cfg, y)
    assert losses[-1] < losses[0]
    assert end < 0.05 * start, (start, end)
    # Training is deterministic for a fixed seed.
    m2 = T.Model(outs, seed=9)
    m2.train(x, cfg, y, epochs=3, batch=128, lr=3e-3)
    m3 = T.Model(outs, seed=9)
    m3.train(x, cfg, y, epochs=3, batch=128, lr=3e-3)
    assert np.array_equal(m2.params(), m3.params())


def test_policy_one_takes_the_models_argmin(H, T):
    """Policy 1 picks, for every problem with work, the configuration the
    installed model predicts fastest; policy 0 is the built-in rule."""
    outs = H.lib().hwyblas_grouped_config_count()
    model = T.Model(outs, seed=11)
    rng = np.random.default_rng(6)
    shapes = T.sample_shapes(rng, 200)
    probs = [H.GemmProblem(b"T" if ta else b"N", b"T" if tb else b"N", int(m), int(n), int(k),
                           1.0, 0.0, None, int(k if ta else m), None, int(n if tb else k), None,
                           int(m)) for ta, tb, m, n, k in shapes]
    arr = (H.GemmProblem * len(probs))(*probs)
    c0 = np.zeros(len(probs), dtype=np.int32)
    c1 = np.zeros(len(probs), dtype=np.int32)
    assert H.lib().hwyblas_grouped_choose_configs(8, len(probs), arr, 0, c0.ctypes.data,
                                                   None) == 0
    assert np.array_equal(c0, T.default_choice(8, shapes))
    model.install()
    try:
        assert H.lib().hwyblas_grouped_choose_configs(8, len(probs), arr, 1, c1.ctypes.data,
                                                       None) == 0
    finally:
        H.lib().hwyblas_grouped_set_model(None)
    pred = model.predict(T.features(8, shapes))
    assert np.array_equal(c1, pred.argmin(axis=1))


def test_exhaustive_policy_uses_the_tuner(H, T):
    """Without a model, policy 1 tunes each distinct shape; the choice is
    the fastest configuration of a tuning run, and a repeated shape is
    tuned once."""
    shapes = np.array([[0, 0, 40, 48, 64], [1, 0, 200, 16, 300], [0, 0, 40, 48, 64]],
                      dtype=np.int32)
    probs = [H.GemmProblem(b"T" if ta else b"N", b"N", int(m), int(n), int(k), 1.0, 0.0, None,
                           int(k if ta else m), None, int(k), None, int(m))
             for ta, tb, m, n, k in shapes]
    arr = (H.GemmProblem * len(probs))(*probs)
    cfg = np.zeros(len(probs), dtype=np.int32)
    H.lib().hwyblas_grouped_set_model(None)
    assert H.lib().hwyblas_grouped_choose_configs(4, len(probs), arr, 1, cfg.ctypes.data,
                                                   None) == 0
    ncfg = H.lib().hwyblas_grouped_config_count()
    assert cfg[0] == cfg[2] and all(0 <= c < ncfg for c in cfg)
    us = T.tune(4, shapes[:2], reps=2)
    assert us.shape == (2, ncfg) and (us > 0).all()


def test_trained_model_speeds_up_an_uneven_batch(H, T):
    """The committed model (trained on the autotuner's samples) picks
    configurations for an unseen uneven batch that run faster than the
    built-in rule's; both plans give the same bits (every configuration sums
    k in the same order)."""
    import time
    rng = np.random.default_rng(99)
    count = 3000
    shapes = T.sample_shapes(rng, count)
    sa = [(k, m) if ta else (m, k) for ta, tb, m, n, k in shapes]
    sb = [(n, k) if tb else (k, n) for ta, tb, m, n, k in shapes]
    sizes = [(int(r) * int(c) + 63) // 64 * 64 for r, c in sa + sb]
    offa = np.concatenate([[0], np.cumsum(sizes[:count])])
    offb = np.concatenate([[0], np.cumsum(sizes[count:])])
    offc = np.concatenate([[0], np.cumsum([(int(m) * int(n) + 63) // 64 * 64
                                           for _, _, m, n, _ in shapes])])
    A, B = H.DeviceArray(int(offa[-1]), 1), H.DeviceArray(int(offb[-1]), 1)
    C0, C1 = H.DeviceArray(int(offc[-1]), 1), H.DeviceArray(int(offc[-1]), 1)
    A.fill_uniform(int(offa[-1]), 1, 5)
    B.fill_uniform(int(offb[-1]), 1, 6)

    def problems(C):
        return [H.GemmProblem(b"T" if ta else b"N", b"T" if tb else b"N", int(m), int(n),
                              int(k), 1.0, 0.0, A.ptr + 8 * int(offa[t]), int(sa[t][0]),
                              B.ptr + 8 * int(offb[t]), int(sb[t][0]), C.ptr + 8 * int(offc[t]),
                              int(m)) for t, (ta, tb, m, n, k) in enumerate(shapes)]
    arr = (H.GemmProblem * count)(*problems(C0))
    model = T.load_model()
    try:
        cfg = {}
        for policy in (0, 1):
            c = np.zeros(count, dtype=np.int32)
            assert H.lib().hwyblas_grouped_choose_configs(8, count, arr, policy, c.ctypes.data,
                                                           None) == 0
            cfg[policy] = c
    finally:
        H.lib().hwyblas_grouped_set_model(None)
    assert (cfg[0] != cfg[1]).any()
    plans = {0: H.Plan("gemm", np.float64, problems(C0), configs=cfg[0]),
             1: H.Plan("gemm", np.float64, problems(C1), configs=cfg[1])}

    def best_ms(plan):
        plan.launch(graph=True)
        H.synchronize()
        best = np.inf
        for _ in range(7):
            t0 = time.perf_counter()
            for _ in range(5):
                plan.launch(graph=True)
            H.synchronize()
            best = min(best, (time.perf_counter() - t0) / 5 * 1e3)
        return best
    t_rule, t_model = best_ms(plans[0]), best_ms(plans[1])
    print(f"{count} uneven DGEMMs: built-in rule {t_rule:.3f} ms, model {t_model:.3f} ms, "
          f"gain {t_rule / t_model:.3f}x")
    assert np.array_equal(C0.to_host(), C1.to_host())
    assert t_model < 0.99 * t_rule
