"""Autotuner driver and cost-model trainer for the grouped (batched) GEMM.

    python -m hwyblas_amd.tune --shapes 1000 --out DIR

1. samples shapes per precision (m, n, k log-uniform integers in [8, 512],
   random transa / transb), the distribution of the uneven batches;
2. times every tile configuration on every shape with the library's
   autotuner (hwyblas_grouped_tune): one (shape, configuration) sample each;
3. trains the MLP cost model on the GPU (hwyblas_mlp_train) on the shapes of
   the training split, target log2(microseconds per problem);
4. reports, on the held-out shapes, how far each way of choosing a
   configuration is from the fastest one measured;
5. writes DIR/grouped_samples.npz (the samples) and DIR/grouped_mlp.npz (the
   trained parameters).

`load_model()` installs a trained model for policy 1 of
hwyblas_grouped_choose_configs.
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import grouped_configs, lib

TUNING_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
DEFAULT_MODEL = os.path.join(TUNING_DIR, "grouped_mlp.npz")
DEFAULT_SAMPLES = os.path.join(TUNING_DIR, "grouped_samples.npz")


def log(*a):
    print("[tune]", *a, file=sys.stderr, flush=True)


def log_uniform_ints(rng, lo, hi, size):
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size))).astype(int).clip(lo, hi)


def sample_shapes(rng, count, lo=8, hi=512):
    m, n, k = (log_uniform_ints(rng, lo, hi, count) for _ in range(3))
    ta = rng.integers(0, 2, count)
    tb = rng.integers(0, 2, count)
    return np.stack([ta, tb, m, n, k], axis=1).astype(np.int32)


def tune(dtype_bytes, shapes, reps=2, stream=None, progress=50):
    """Microseconds per problem, shapes x configurations."""
    ncfg = lib().hwyblas_grouped_config_count()
    us = np.zeros((len(shapes), ncfg), dtype=np.float32)
    t0 = time.time()
    for i0 in range(0, len(shapes), progress):
        chunk = np.ascontiguousarray(shapes[i0:i0 + progress], dtype=np.int32)
        out = np.zeros((len(chunk), ncfg), dtype=np.float32)
        rc = lib().hwyblas_grouped_tune(dtype_bytes, len(chunk), chunk.ctypes.data, reps,
                                        out.ctypes.data, stream)
        if rc != 0:
            raise ValueError("hwyblas_grouped_tune rejected the shapes")
        us[i0:i0 + len(chunk)] = out
        log(f"dtype {dtype_bytes}: {i0 + len(chunk)}/{len(shapes)} shapes, "
            f"{time.time() - t0:.1f} s")
    return us


def features(dtype_bytes, shapes):
    buf = (ctypes.c_float * 6)()
    out = np.zeros((len(shapes), 6), dtype=np.float32)
    for i, (ta, tb, m, n, k) in enumerate(shapes):
        lib().hwyblas_grouped_features(dtype_bytes, int(ta), int(tb), int(m), int(n), int(k), buf)
        out[i] = buf[:]
    return out


class Model:
    """A cost model on the GPU (hwyblas_mlp_*)."""

    def __init__(self, outputs, seed=1, params=None):
        self.ptr = lib().hwyblas_mlp_create(outputs, seed)
        if not self.ptr:
            raise ValueError("bad output count")
        self.outputs = outputs
        if params is not None:
            self.set_params(params)

    @property
    def num_params(self):
        return lib().hwyblas_mlp_num_params(self.ptr)

    def params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        lib().hwyblas_mlp_get_params(self.ptr, p.ctypes.data)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        lib().hwyblas_mlp_set_params(self.ptr, p.ctypes.data)

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros((len(x), self.outputs), dtype=np.float32)
        if lib().hwyblas_mlp_predict(self.ptr, len(x), x.ctypes.data, out.ctypes.data, None):
            raise ValueError("predict failed")
        return out

    def gradients(self, x, cfg, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        cfg = np.ascontiguousarray(cfg, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = ctypes.c_float(0)
        if lib().hwyblas_mlp_gradients(self.ptr, len(x), x.ctypes.data, cfg.ctypes.data,
                                       y.ctypes.data, g.ctypes.data, ctypes.byref(loss), None):
            raise ValueError("gradients failed")
        return float(loss.value), g

    def train(self, x, cfg, y, epochs, batch=256, lr=2e-3, seed=7):
        x = np.ascontiguousarray(x, dtype=np.float32)
        cfg = np.ascontiguousarray(cfg, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        if lib().hwyblas_mlp_train(self.ptr, len(x), x.ctypes.data, cfg.ctypes.data,
                                   y.ctypes.data, epochs, batch, lr, seed, losses.ctypes.data,
                                   None):
            raise ValueError("train failed")
        return losses[:epochs]

    def install(self):
        lib().hwyblas_grouped_set_model(self.ptr)

    def destroy(self):
        if self.ptr:
            lib().hwyblas_mlp_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def load_model(path=DEFAULT_MODEL, install=True):
    with np.load(path, allow_pickle=False) as z:
        params, outputs = z["params"], int(z["outputs"])
    if outputs != lib().hwyblas_grouped_config_count():
        raise ValueError(f"{path}: model for {outputs} configurations, library has "
                         f"{lib().hwyblas_grouped_config_count()}")
    m = Model(outputs, params=params)
    if install:
        m.install()
    return m


def training_rows(dtype_bytes, shapes, us):
    """One (features, configuration, log2 us) row per sample."""
    f = features(dtype_bytes, shapes)
    ncfg = us.shape[1]
    x = np.repeat(f, ncfg, axis=0)
    cfg = np.tile(np.arange(ncfg, dtype=np.int32), len(shapes))
    y = np.log2(us.reshape(-1)).astype(np.float32)
    return x, cfg, y


def default_choice(dtype_bytes, shapes):
    return np.array([lib().hwyblas_grouped_default_config(dtype_bytes, int(m), int(n), int(k))
                     for _, _, m, n, k in shapes])


def regret(us, choice):
    """Per shape: time of the chosen configuration / fastest time measured."""
    return us[np.arange(len(us)), choice] / us.min(axis=1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shapes", type=int, default=1000, help="shapes per precision")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=TUNING_DIR)
    ap.add_argument("--samples", default=None, help="reuse a samples file instead of tuning")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    ncfg = lib().hwyblas_grouped_config_count()
    rng = np.random.default_rng(args.seed)
    if args.samples:
        with np.load(args.samples, allow_pickle=False) as z:
            data = {k: z[k] for k in z.files}
    else:
        parts = []
        for dtype_bytes in (8, 4):
            shapes = sample_shapes(rng, args.shapes)
            us = tune(dtype_bytes, shapes, args.reps)
            parts.append((np.full(len(shapes), dtype_bytes, np.int32), shapes, us))
        db = np.concatenate([p[0] for p in parts])
        shapes = np.concatenate([p[1] for p in parts])
        us = np.concatenate([p[2] for p in parts])
        split = (rng.uniform(size=len(db)) < 0.2).astype(np.int8)  # 1: held out
        data = {"dtype_bytes": db, "shapes": shapes, "us": us, "split": split,
                "configs": np.array([list(c) for c in grouped_configs()], dtype=np.int32)}
        np.savez_compressed(os.path.join(args.out, "grouped_samples.npz"), **data)
        log(f"{us.size} samples ({len(db)} shapes x {ncfg} configurations) written")
    db, shapes, us, split = data["dtype_bytes"], data["shapes"], data["us"], data["split"]

    xs, cs, ys = [], [], []
    for b in (8, 4):
        sel = (db == b) & (split == 0)
        x, c, y = training_rows(b, shapes[sel], us[sel])
        xs.append(x), cs.append(c), ys.append(y)
    x, c, y = np.concatenate(xs), np.concatenate(cs), np.concatenate(ys)
    model = Model(ncfg, seed=args.seed)
    t0 = time.time()
    losses = model.train(x, c, y, args.epochs)
    log(f"trained on {len(x)} samples, {args.epochs} epochs in {time.time() - t0:.1f} s: "
        f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")

    report = {}
    for b in (8, 4):
        sel = (db == b) & (split == 1)
        sh, u = shapes[sel], us[sel]
        pred = model.predict(features(b, sh))
        r_mlp = regret(u, pred.argmin(axis=1))
        r_def = regret(u, default_choice(b, sh))
        best_single = int(np.argmin(u.sum(axis=0)))
        r_single = regret(u, np.full(len(u), best_single))
        batch_t = {"best": float(u.min(axis=1).sum()),
                   "mlp": float(u[np.arange(len(u)), pred.argmin(axis=1)].sum()),
                   "default": float(u[np.arange(len(u)), default_choice(b, sh)].sum()),
                   "single": float(u[:, best_single].sum())}
        report[b] = (r_mlp.mean(), r_def.mean(), r_single.mean(), batch_t)
        log(f"dtype {b} held-out ({len(u)} shapes): mean regret mlp {r_mlp.mean():.3f}, "
            f"default {r_def.mean():.3f}, best single config {best_single} "
            f"{r_single.mean():.3f}; batch time (us) best {batch_t['best']:.1f} mlp "
            f"{batch_t['mlp']:.1f} default {batch_t['default']:.1f} single "
            f"{batch_t['single']:.1f}")
    np.savez_compressed(os.path.join(args.out, "grouped_mlp.npz"), params=model.params(),
                        outputs=np.int32(ncfg), losses=losses, epochs=np.int32(args.epochs),
                        train_samples=np.int32(len(x)))
    log(f"model written to {os.path.join(args.out, 'grouped_mlp.npz')}")


if __name__ == "__main__":
    main()
