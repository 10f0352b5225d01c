"""CPU restatement (float64 numpy) of the grouped-GEMM cost model.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
training and inference kernels; the library never calls it.

The model has no hwyBLAS counterpart (hwyBLAS fixes its blocking per SIMD
target at compile time, src/gemm_kernel-inl.h:60-75). It is a 4-layer MLP of
width 64: 6 inputs -> 64 -> 64 -> 64 (ReLU) -> `outputs` linear units, FP32
parameters row-major W[out][in] packed as W1 b1 W2 b2 W3 b3 W4 b4. The loss of
a batch is 0.5 * mean_i (out[i, cfg_i] - y_i)^2.

Pinned by: a forward pass computed by hand on a sparse parameter set, and the
gradient against central finite differences in float64
(tests/test_oracle.py).
"""
import numpy as np

K_IN = 6
WIDTH = 64


def offsets(outputs):
    """(name, offset, shape) of every parameter block."""
    out, off = [], 0
    for name, shape in (("W1", (WIDTH, K_IN)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)),
                        ("b2", (WIDTH,)), ("W3", (WIDTH, WIDTH)), ("b3", (WIDTH,)),
                        ("W4", (outputs, WIDTH)), ("b4", (outputs,))):
        out.append((name, off, shape))
        off += int(np.prod(shape))
    return out


def num_params(outputs):
    name, off, shape = offsets(outputs)[-1]
    return off + int(np.prod(shape))


def unpack(params, outputs):
    p = np.asarray(params, dtype=np.float64)
    return {name: p[off:off + int(np.prod(shape))].reshape(shape)
            for name, off, shape in offsets(outputs)}


def features(dtype_bytes, ta, tb, m, n, k):
    """log2 m, n, k over 9, the transpositions and the precision."""
    return np.array([np.log2(max(m, 1)) / 9.0, np.log2(max(n, 1)) / 9.0,
                     np.log2(max(k, 1)) / 9.0, 1.0 if ta else 0.0, 1.0 if tb else 0.0,
                     1.0 if dtype_bytes == 8 else 0.0])


def forward(params, outputs, x):
    """Outputs (n x outputs) and the hidden activations."""
    w = unpack(params, outputs)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ w["W1"].T + w["b1"], 0.0)
    h2 = np.maximum(h1 @ w["W2"].T + w["b2"], 0.0)
    h3 = np.maximum(h2 @ w["W3"].T + w["b3"], 0.0)
    out = h3 @ w["W4"].T + w["b4"]
    return out, (x, h1, h2, h3)


def loss(params, outputs, x, cfg, y):
    out, _ = forward(params, outputs, x)
    r = out[np.arange(len(cfg)), cfg] - np.asarray(y, dtype=np.float64)
    return 0.5 * float(np.mean(r * r))


def loss_and_grad(params, outputs, x, cfg, y):
    """The loss and its gradient w.r.t. the packed parameters (backpropagation)."""
    w = unpack(params, outputs)
    out, (x, h1, h2, h3) = forward(params, outputs, x)
    n = len(cfg)
    rows = np.arange(n)
    r = out[rows, cfg] - np.asarray(y, dtype=np.float64)
    d4 = np.zeros_like(out)
    d4[rows, cfg] = r / n
    d3 = (d4 @ w["W4"]) * (h3 > 0)
    d2 = (d3 @ w["W3"]) * (h2 > 0)
    d1 = (d2 @ w["W2"]) * (h1 > 0)
    grads = {"W4": d4.T @ h3, "b4": d4.sum(0), "W3": d3.T @ h2, "b3": d3.sum(0),
             "W2": d2.T @ h1, "b2": d2.sum(0), "W1": d1.T @ x, "b1": d1.sum(0)}
    g = np.zeros(num_params(outputs))
    for name, off, shape in offsets(outputs):
        g[off:off + int(np.prod(shape))] = grads[name].ravel()
    return 0.5 * float(np.mean(r * r)), g
