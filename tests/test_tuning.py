"""The committed autotuner samples and trained cost model (CPU): the data
cover >= 20,000 timed (shape, configuration) samples, the model evaluated by
the float64 restatement (oracle/mlp.py) picks configurations whose measured
time on the held-out shapes is closer to the fastest than the built-in rule's,
and the batch assembled from the held-out shapes runs faster with them."""
import numpy as np
import pytest

import hwyblas_amd as H
from hwyblas_amd import tune
from oracle import mlp as omlp


@pytest.fixture(scope="module")
def data():
    with np.load(tune.DEFAULT_SAMPLES, allow_pickle=False) as z:
        samples = {k: z[k] for k in z.files}
    with np.load(tune.DEFAULT_MODEL, allow_pickle=False) as z:
        model = {k: z[k] for k in z.files}
    return samples, model


def test_samples_cover_the_batched_distribution(data):
    samples, model = data
    us, shapes = samples["us"], samples["shapes"]
    ncfg = H.lib().hwyblas_grouped_config_count()
    assert us.shape[1] == ncfg == int(model["outputs"])
    assert us.size >= 20000 and np.all(us > 0)
    assert shapes[:, 2:].min() >= 8 and shapes[:, 2:].max() <= 512
    assert set(np.unique(samples["dtype_bytes"])) == {4, 8}
    # The stored configuration table is the library's.
    assert [tuple(c) for c in samples["configs"]] == H.grouped_configs()
    assert int(model["train_samples"]) == int((samples["split"] == 0).sum()) * ncfg


def test_model_beats_the_built_in_rule_on_held_out_shapes(data):
    samples, model = data
    ncfg = int(model["outputs"])
    for b in (8, 4):
        sel = (samples["dtype_bytes"] == b) & (samples["split"] == 1)
        sh, us = samples["shapes"][sel], samples["us"][sel]
        x = np.stack([omlp.features(b, *s) for s in sh])
        pred, _ = omlp.forward(model["params"].astype(np.float64), ncfg, x)
        mlp_pick = pred.argmin(axis=1)
        rule = tune.default_choice(b, sh)
        rows = np.arange(len(us))
        r_mlp = (us[rows, mlp_pick] / us.min(axis=1)).mean()
        r_rule = (us[rows, rule] / us.min(axis=1)).mean()
        t_mlp, t_rule = us[rows, mlp_pick].sum(), us[rows, rule].sum()
        print(f"dtype {b}: {len(us)} held-out shapes; mean regret model {r_mlp:.3f} rule "
              f"{r_rule:.3f}; batch time model {t_mlp:.1f} us rule {t_rule:.1f} us")
        assert r_mlp < r_rule and r_mlp < 1.10
        assert t_mlp < 0.95 * t_rule


def test_library_prediction_matches_restatement(data):
    """The features the library computes give the restatement's inputs."""
    samples, model = data
    sh = samples["shapes"][:50]
    x = tune.features(8, sh)
    ref = np.stack([omlp.features(8, *s) for s in sh])
    assert np.allclose(x, ref, rtol=1e-6, atol=1e-7)
