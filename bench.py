#!/usr/bin/env python3
"""Benchmark of the Level-3 hot path on one MI355X.

Workload (one "step" = one DGEMM call over device-resident operands):
  C = A B, column-major, transa = transb = N, m = n = k = 16384,
  lda = ldb = ldc = 16384, entries uniform[-1, 1] from the seeded generator
  shared with the CPU oracle (seeds 1, 2, 3), alpha = 1, beta = 0, FP64 in
  and out. Flops per step: 2 m n k (FP64-equivalent).

The headline value is the emulated DGEMM (INT8 Ozaki scheme II, moduli
chosen from the operands) in TFLOP/s; the native FP64 MFMA DGEMM is timed on
the same operands beside it, and both are checked on sampled entries
against the double-double oracle (max componentwise relative error
|C - C_ref| / (|A||B|)). Prints one JSON line; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DGEMM TFLOP/s @16384^3 emulated vs native FP64 (err<=2x native); batched GEMM/s"
PEAK_FP64 = 78.6      # MI355X dense FP64 matrix, TFLOP/s
PEAK_FP32 = 157.3     # MI355X dense FP32 matrix, TFLOP/s
PEAK_INT8 = 5000.0    # MI355X dense INT8 matrix, TOP/s
PEAK_HBM = 8.0        # MI355X HBM3E, TB/s
SEEDS = (1, 2, 3)


def measured_traffic(n, moduli):
    """HBM bytes per INT8-GEMM launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (scaled by a 1 GiB copy), when they cover this shape."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            t = json.load(f)["emulated_dgemm_16384_nn_14_moduli"]
    except (OSError, KeyError, ValueError):
        return None
    if t["n"] != n or t["moduli"] != moduli:
        return None
    return t["fetch_bytes"] + t["write_bytes"]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def gather(H, dC, n, rows, cols):
    """Sampled entries of a device n x n column-major matrix."""
    got = np.empty(len(rows))
    col = np.empty(n)
    last = -1
    for t in np.argsort(cols, kind="stable"):
        j = int(cols[t])
        if j != last:
            H.lib().hwyblas_amd_memcpy(col.ctypes.data, dC.ptr + j * n * 8, n * 8, 2)
            last = j
        got[t] = col[rows[t]]
    return got


def error_entries(n, samples, seed, full=16):
    """The checked entries: `full` whole rows, `full` whole columns and
    `samples` random entries of an n x n result."""
    rng = np.random.default_rng(seed)
    R = rng.choice(n, full, replace=False)
    Cc = rng.choice(n, full, replace=False)
    sr = rng.integers(0, n, samples)
    sc = rng.integers(0, n, samples)
    return R, Cc, sr, sc


def reference_entries(oracle, ha, hb, n, R, Cc, sr, sc):
    """Double-double reference (value, |A||B| scale) of the checked entries of
    A B: the rows through a contiguous copy of A's rows, the columns and the
    sampled entries through a transposed copy of A (contiguous dot products)."""
    a_rows = np.asfortranarray(ha[R, :].T)  # op = T: op(A)(i, p) = a_rows[p + i n]
    ri = np.repeat(np.arange(len(R), dtype=np.int32), n)
    rj = np.tile(np.arange(n, dtype=np.int32), len(R))
    e_r, b_r = oracle.gemm_ref_entries("T", "N", n, 1.0, a_rows, n, hb, n, 0.0, None, n, ri, rj)
    del a_rows
    haT = np.asfortranarray(ha.T)
    ci = np.tile(np.arange(n, dtype=np.int32), len(Cc))
    cj = np.repeat(np.asarray(Cc, dtype=np.int32), n)
    e_c, b_c = oracle.gemm_ref_entries("T", "N", n, 1.0, haT, n, hb, n, 0.0, None, n, ci, cj)
    e_s, b_s = oracle.gemm_ref_entries("T", "N", n, 1.0, haT, n, hb, n, 0.0, None, n, sr, sc)
    del haT
    return np.concatenate([e_r, e_c, e_s]), np.concatenate([b_r, b_c, b_s])


def result_entries(H, dC, n, R, Cc, sr, sc):
    """The same entries of a device result (whole matrix copied once)."""
    c = dC.to_host()
    out = np.concatenate([c[R, :].ravel(), c[:, Cc].T.ravel(), c[sr, sc]])
    del c
    return out


def error_summary(got, exp, bnd):
    """Max componentwise error |C - C_ref| / (|A||B|) and max plain relative
    error |C - C_ref| / |C_ref| (entries with C_ref != 0)."""
    err = np.abs(got - exp)
    nz = exp != 0
    return float(np.max(err / bnd)), float(np.max(err[nz] / np.abs(exp[nz])))


def time_steps(H, stream, step, steps, warmup):
    for _ in range(warmup):
        step()
    H.lib().hwyblas_amd_stream_synchronize(stream)
    H.synchronize()
    e0, e1 = H.Event(), H.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    H.lib().hwyblas_amd_stream_synchronize(stream)
    H.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_ms(e1) / steps, wall


def time_loop_graph(H, stream, loop_step, steps):
    """The same loop of single-call launches captured once in a hipGraph and
    replayed: the loop's GPU time without host launch gaps."""
    L = H.lib()
    L.hwyblas_amd_capture_begin(stream)
    loop_step()
    exe = L.hwyblas_amd_capture_end(stream)
    ms, _ = time_steps(H, stream, lambda: L.hwyblas_amd_graph_launch(exe, stream), steps, 1)
    L.hwyblas_amd_graph_destroy(exe)
    return ms


def log_uniform_ints(rng, lo, hi, size):
    """Integers log-uniform in [lo, hi]: floor(exp(U(ln lo, ln(hi + 1))))."""
    return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1), size))).astype(int).clip(lo, hi)


class DeviceBatch:
    """Operands of many column-major problems packed in three device arenas
    (A, B, C), each filled in one launch with the seeded uniform[-1, 1]
    generator (element t of an arena is uniform(seed, t))."""

    def __init__(self, H, dtype, shapes_a, shapes_b, shapes_c, seed):
        self.H, self.dtype = H, np.dtype(dtype)
        self.arenas, self.offsets = [], []
        for s, shapes in enumerate((shapes_a, shapes_b, shapes_c)):
            sizes = np.array([max(1, r * c) for r, c in shapes], dtype=np.int64)
            off = np.concatenate([[0], np.cumsum((sizes + 63) // 64 * 64)])
            arena = H.DeviceArray(int(off[-1]), 1, dtype)
            arena.fill_uniform(int(off[-1]), 1, seed + s)
            self.arenas.append(arena)
            self.offsets.append(off[:-1])
        H.synchronize()

    def ptr(self, which, t):
        return self.arenas[which].ptr + int(self.offsets[which][t]) * self.dtype.itemsize

    def download(self, which, t, rows, cols):
        out = np.empty((rows, cols), dtype=self.dtype, order="F")
        if out.size:
            self.H.lib().hwyblas_amd_memcpy(out.ctypes.data, self.ptr(which, t), out.nbytes, 2)
        return out


def batched_gemm_leg(H, oracle, dtype, count, steps, warmup, stream, model=None, configs=None,
                     check=16, seed=7, with_loop=True):
    """`count` uneven GEMMs, m, n, k log-uniform in [8, 512], random trans,
    uniform[-1, 1], alpha = 1, beta = 0: one grouped launch replayed from a
    hipGraph, against a loop of single-call hwyblas_?gemm_dev launches."""
    rng = np.random.default_rng(seed)
    m, n, k = (log_uniform_ints(rng, 8, 512, count) for _ in range(3))
    ta = rng.integers(0, 2, count)
    tb = rng.integers(0, 2, count)
    sa = [(k[t], m[t]) if ta[t] else (m[t], k[t]) for t in range(count)]
    sb = [(n[t], k[t]) if tb[t] else (k[t], n[t]) for t in range(count)]
    sc = [(m[t], n[t]) for t in range(count)]
    batch = DeviceBatch(H, dtype, sa, sb, sc, seed=100)
    probs = [H.GemmProblem(b"T" if ta[t] else b"N", b"T" if tb[t] else b"N", int(m[t]),
                           int(n[t]), int(k[t]), 1.0, 0.0, batch.ptr(0, t), int(sa[t][0]),
                           batch.ptr(1, t), int(sb[t][0]), batch.ptr(2, t), int(m[t]))
             for t in range(count)]
    flops = float(np.sum(2.0 * m * n * k))
    nbytes = float(np.sum((m * k + k * n + m * n) * np.dtype(dtype).itemsize))
    arr = (H.GemmProblem * count)(*probs)
    dbytes = np.dtype(dtype).itemsize

    def choose(policy):
        cfg = np.zeros(count, dtype=np.int32)
        assert H.lib().hwyblas_grouped_choose_configs(dbytes, count, arr, policy,
                                                       cfg.ctypes.data, stream) == 0
        return cfg

    def time_plan(cfg):
        t0 = time.perf_counter()
        pl = H.Plan("gemm", dtype, probs, configs=cfg)
        ps = time.perf_counter() - t0
        ms, _ = time_steps(H, stream, lambda: pl.launch(stream, graph=True), steps, warmup)
        return pl, ms, ps

    # The built-in rule, then the configurations the cost model picks.
    plan0, d_ms, _ = time_plan(choose(0))
    plan0.destroy()
    policy = "default"
    if model is not None:
        t0 = time.perf_counter()
        cfg_used = choose(1)
        choose_s = time.perf_counter() - t0
        policy = "mlp"
    else:
        cfg_used, choose_s = choose(0), 0.0
    if configs is not None:
        cfg_used, policy = np.asarray(configs, dtype=np.int32), "given"
    plan, g_ms, plan_s = time_plan(cfg_used)

    pre = "d" if np.dtype(dtype) == np.float64 else "s"
    f = getattr(H.lib(), f"hwyblas_{pre}gemm_dev")
    import ctypes
    args = [(p.transa, p.transb, p.m, p.n, p.k, 1.0, ctypes.c_void_p(p.a), p.lda,
             ctypes.c_void_p(p.b), p.ldb, 0.0, ctypes.c_void_p(p.c), p.ldc, stream)
            for p in probs]

    def loop_step():
        for a in args:
            f(*a)
    l_ms = lg_ms = float("nan")
    if with_loop:
        l_ms, _ = time_steps(H, stream, loop_step, max(1, steps // 4), 1)
        lg_ms = time_loop_graph(H, stream, loop_step, max(1, steps // 2))

    # Grouped result of the last replay, then checked problem by problem.
    plan.launch(stream, graph=True)
    H.lib().hwyblas_amd_stream_synchronize(stream)
    worst = 0.0
    for t in rng.choice(count, size=min(check, count), replace=False):
        a = batch.download(0, t, *sa[t])
        b = batch.download(1, t, *sb[t])
        c = batch.download(2, t, *sc[t])
        exp, bnd = oracle.gemm_ref("T" if ta[t] else "N", "T" if tb[t] else "N", int(m[t]),
                                   int(n[t]), int(k[t]), 1.0, a, sa[t][0], b, sb[t][0], 0.0,
                                   None, int(m[t]))
        tol = 4.0 * (k[t] + 2) * np.finfo(dtype).eps
        worst = max(worst, float(np.max(np.abs(c.astype(np.float64) - exp) / (tol * bnd))))
    cfg_hist = np.bincount([plan.problem_config(t) for t in range(count)],
                           minlength=len(H.grouped_configs())).tolist()
    out = {"count": count, "dtype": "f64" if pre == "d" else "f32", "configs": policy,
           "grouped_ms": round(g_ms, 4), "grouped_ms_default_configs": round(d_ms, 4),
           "gain_vs_default_configs": round(d_ms / g_ms, 3),
           "choose_configs_s": round(choose_s, 3), "loop_ms": round(l_ms, 4),
           "gemm_per_s": round(count / (g_ms * 1e-3), 1),
           "loop_gemm_per_s": round(count / (l_ms * 1e-3), 1),
           "speedup_vs_loop": round(l_ms / g_ms, 3),
           "loop_graph_ms": round(lg_ms, 4),
           "speedup_vs_loop_graph": round(lg_ms / g_ms, 3),
           "tflops": round(flops / (g_ms * 1e-3) / 1e12, 3),
           "gb_per_s_min_traffic": round(nbytes / (g_ms * 1e-3) / 1e9, 1),
           "flops": flops, "work_items": plan.work_items, "plan_create_s": round(plan_s, 3),
           "worst_bound_ratio": worst, "checked": int(min(check, count)),
           "config_histogram": cfg_hist}
    plan.destroy()
    del batch
    return out


def batched_trsm_leg(H, oracle, count, steps, warmup, stream, check=16, seed=11, with_loop=True):
    """`count` DTRSMs L/L/N/N, m log-uniform in [8, 256], n in [1, 64]; A
    lower triangular with uniform[-1, 1] off-diagonal entries and diagonal
    m + uniform[0, 1]: one grouped launch replayed from a hipGraph, against a
    loop of single-call hwyblas_dtrsm_dev launches."""
    rng = np.random.default_rng(seed)
    m = log_uniform_ints(rng, 8, 256, count)
    n = log_uniform_ints(rng, 1, 64, count)
    batch = DeviceBatch(H, np.float64, [(x, x) for x in m], [(m[t], n[t]) for t in range(count)],
                        [(1, 1)] * count, seed=200)
    # Diagonal m + U[0, 1]: one host pass over the A arena's diagonal entries.
    total = int(batch.offsets[0][-1] + m[-1] * m[-1])
    host = np.empty(total)
    H.lib().hwyblas_amd_memcpy(host.ctypes.data, batch.arenas[0].ptr, host.nbytes, 2)
    for t in range(count):
        o, mm = int(batch.offsets[0][t]), int(m[t])
        idx = o + np.arange(mm) * (mm + 1)
        host[idx] = mm + (host[idx] + 1.0) * 0.5
    H.lib().hwyblas_amd_memcpy(batch.arenas[0].ptr, host.ctypes.data, host.nbytes, 1)
    del host
    B0 = H.DeviceArray(batch.arenas[1].ld, 1)
    nb = batch.arenas[1].ld * 8
    H.lib().hwyblas_amd_memcpy(B0.ptr, batch.arenas[1].ptr, nb, 3)
    probs = [H.TrsmProblem(b"L", b"L", b"N", b"N", int(m[t]), int(n[t]), 1.0, batch.ptr(0, t),
                           int(m[t]), batch.ptr(1, t), int(m[t])) for t in range(count)]
    plan = H.Plan("trsm", np.float64, probs)
    flops = float(np.sum(m.astype(float) ** 2 * n))

    def grouped_step():
        plan.launch(stream, graph=True)
    g_ms, _ = time_steps(H, stream, grouped_step, steps, warmup)
    import ctypes
    f = H.lib().hwyblas_dtrsm_dev
    args = [(b"L", b"L", b"N", b"N", p.m, p.n, 1.0, ctypes.c_void_p(p.a), p.lda,
             ctypes.c_void_p(p.b), p.ldb, stream) for p in probs]

    def loop_step():
        for a in args:
            f(*a)
    l_ms = lg_ms = float("nan")
    if with_loop:
        l_ms, _ = time_steps(H, stream, loop_step, max(1, steps // 4), 1)
        lg_ms = time_loop_graph(H, stream, loop_step, max(1, steps // 2))
    # One solve from the original right-hand sides, checked by residual.
    H.lib().hwyblas_amd_memcpy(batch.arenas[1].ptr, B0.ptr, nb, 3)
    H.synchronize()
    plan.launch(stream, graph=True)
    H.lib().hwyblas_amd_stream_synchronize(stream)
    worst = 0.0
    for t in rng.choice(count, size=min(check, count), replace=False):
        mm, nn = int(m[t]), int(n[t])
        a = batch.download(0, t, mm, mm)
        x = batch.download(1, t, mm, nn)
        b0 = np.empty((mm, nn), order="F")
        H.lib().hwyblas_amd_memcpy(b0.ctypes.data, B0.ptr + int(batch.offsets[1][t]) * 8,
                                   b0.nbytes, 2)
        r = oracle.trsm_residual("L", "L", "N", "N", mm, nn, 1.0, a, mm, b0, mm, x, mm,
                                 np.arange(nn))
        worst = max(worst, float(np.max(r) / (8.0 * (mm + 2) * np.finfo(np.float64).eps)))
    out = {"count": count, "dtype": "f64", "grouped_ms": round(g_ms, 4),
           "loop_ms": round(l_ms, 4), "solves_per_s": round(count / (g_ms * 1e-3), 1),
           "loop_solves_per_s": round(count / (l_ms * 1e-3), 1),
           "speedup_vs_loop": round(l_ms / g_ms, 3),
           "loop_graph_ms": round(lg_ms, 4),
           "speedup_vs_loop_graph": round(lg_ms / g_ms, 3),
           "tflops": round(flops / (g_ms * 1e-3) / 1e12, 4), "work_items": plan.work_items,
           "worst_residual_ratio": worst, "checked": int(min(check, count))}
    plan.destroy()
    return out


def plain_sweep(H, oracle, sizes=(4096, 8192, 16384), reps=3, samples=64):
    """The plain kernels at scale: S/DGEMM m = n = k in `sizes`, every
    transa/transb, column-major (hwyblas_?gemm_dev, event-timed) and row-major
    (cblas_?gemm on device pointers, wall-timed, synchronous), uniform[-1, 1],
    alpha = 1, beta = 0; then DTRSM m = n = 8192 for the 8 side/uplo/trans
    combinations (diag N), A triangular with uniform[-1, 1] off-diagonal
    entries and diagonal n + uniform[0, 1]. Sampled entries are checked
    against the double-double reference (GEMM: componentwise bound ratio;
    TRSM: residual ratio)."""
    L = H.lib()
    stream = L.hwyblas_amd_stream_create()
    out = {"gemm": [], "trsm": []}
    rng = np.random.default_rng(5)
    for dtype, pre, peak in ((np.float64, "d", PEAK_FP64), (np.float32, "s", 157.3)):
        for n in sizes:
            dA, dB, dC = (H.DeviceArray(n, n, dtype) for _ in range(3))
            dA.fill_uniform(n, n, 1, stream=stream)
            dB.fill_uniform(n, n, 2, stream=stream)
            L.hwyblas_amd_stream_synchronize(stream)
            ha = oracle.fill_uniform(n, n, 1, dtype=dtype)
            hb = oracle.fill_uniform(n, n, 2, dtype=dtype)
            rows = rng.integers(0, n, samples)
            cols = rng.integers(0, n, samples)
            flops = 2.0 * n ** 3
            for layout in ("col", "row"):
                for ta in "NT":
                    for tb in "NT":
                        if layout == "col":
                            def step():
                                assert H.gemm_dev(ta, tb, n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n,
                                                  stream=stream, dtype=dtype) == 0
                            ms, _ = time_steps(H, stream, step, reps, 1)
                        else:
                            t_a = H.Trans if ta == "T" else H.NoTrans
                            t_b = H.Trans if tb == "T" else H.NoTrans
                            H.cblas_gemm(H.RowMajor, t_a, t_b, n, n, n, 1.0, dA, n, dB, n, 0.0,
                                         dC, n, dtype=dtype)
                            t0 = time.perf_counter()
                            for _ in range(reps):
                                H.cblas_gemm(H.RowMajor, t_a, t_b, n, n, n, 1.0, dA, n, dB, n,
                                             0.0, dC, n, dtype=dtype)
                            ms = (time.perf_counter() - t0) * 1e3 / reps
                        got = np.empty(samples, dtype=dtype)
                        for t in range(samples):
                            L.hwyblas_amd_memcpy(got[t:].ctypes.data,
                                                 dC.ptr + int(rows[t] + cols[t] * n) *
                                                 np.dtype(dtype).itemsize,
                                                 np.dtype(dtype).itemsize, 2)
                        if layout == "col":
                            exp, bnd = oracle.gemm_ref_entries(ta, tb, n, 1.0, ha, n, hb, n, 0.0,
                                                               None, n, rows, cols)
                        else:  # row-major C = op(A) op(B)  <=>  column-major C = op(B) op(A)
                            exp, bnd = oracle.gemm_ref_entries(tb, ta, n, 1.0, hb, n, ha, n, 0.0,
                                                               None, n, rows, cols)
                        ratio = float(np.max(np.abs(got.astype(np.float64) - exp) /
                                             (4.0 * (n + 2) * np.finfo(dtype).eps * bnd)))
                        tf = flops / (ms * 1e-3) / 1e12
                        r = {"dtype": pre, "n": n, "layout": layout, "transa": ta,
                             "transb": tb, "ms": round(ms, 3), "tflops": round(tf, 2),
                             "frac_peak": round(tf / peak, 3), "bound_ratio": ratio}
                        log(f"{pre}gemm {layout} {ta}{tb} {n}: {ms:.2f} ms {tf:.1f} TFLOP/s "
                            f"({tf / peak:.0%}), bound ratio {ratio:.3g}")
                        out["gemm"].append(r)
            del dA, dB, dC, ha, hb
    n = 8192
    ha = oracle.fill_uniform(n, n, 5)
    ha[np.arange(n), np.arange(n)] = n + (ha[np.arange(n), np.arange(n)] + 1.0) * 0.5
    hb = oracle.fill_uniform(n, n, 6)
    dA = H.DeviceArray.from_host(ha)
    dB0 = H.DeviceArray.from_host(hb)
    dB = H.DeviceArray(n, n)
    for side in "LR":
        for uplo in "LU":
            for ta in "NT":
                def reset():
                    L.hwyblas_amd_memcpy(dB.ptr, dB0.ptr, n * n * 8, 3)

                def step():
                    assert H.trsm_dev(side, uplo, ta, "N", n, n, 1.0, dA, n, dB, n,
                                      stream=stream, dtype=np.float64) == 0
                times = []
                for _ in range(reps):
                    reset()
                    ms, _ = time_steps(H, stream, step, 1, 0)
                    times.append(ms)
                ms = min(times)
                reset()
                step()
                L.hwyblas_amd_stream_synchronize(stream)
                which = rng.integers(0, n, 3)
                x = dB.to_host()
                rr = oracle.trsm_residual(side, uplo, ta, "N", n, n, 1.0, ha, n, hb, n, x, n,
                                          which)
                ratio = float(np.max(rr) / (8.0 * (n + 2) * np.finfo(np.float64).eps))
                tf = float(n) ** 3 / (ms * 1e-3) / 1e12
                out["trsm"].append({"side": side, "uplo": uplo, "transa": ta, "n": n,
                                    "ms": round(ms, 3), "tflops": round(tf, 2),
                                    "frac_peak": round(tf / PEAK_FP64, 3),
                                    "residual_ratio": ratio})
                log(f"dtrsm {side}{uplo}{ta}N {n}: {ms:.2f} ms {tf:.1f} TFLOP/s, residual ratio "
                    f"{ratio:.3g}")
    L.hwyblas_amd_stream_destroy(stream)
    return out


METHODS = {"int8": "INT8 Ozaki scheme II (v_mfma_i32_32x32x32_i8)",
           "bf16": "BF16 slices, Ozaki scheme I (v_mfma_f32_32x32x16_bf16, exact FP32 "
                   "chunks of 1024)"}
DISTRIBUTIONS = {
    "b": "+-u 2^e, u uniform in [1,2), e uniform integer in [-30,30] (seeds 11, 12)",
    "c": "uniform[-1,1] (seeds 1, 2), rows of A and columns of B scaled by 10^s, "
         "s uniform in [-8,8] (numpy seeds 21, 22)",
}


_REFERENCE_CACHE = {}


def distribution_leg(H, oracle, dist, n, dA, dB, dC, stream, samples, nmod=0, steps=2,
                     method="int8"):
    """Emulated (moduli or BF16 slices from the operands, or `nmod` of them)
    and native DGEMM on one of the operand distributions; sampled
    componentwise errors against the double-double reference on the same
    operands regenerated on the host. method: "int8" (Ozaki scheme II) or
    "bf16" (BF16 slices, Ozaki scheme I)."""
    L = H.lib()
    if dist == "b":
        L.hwyblas_amd_fill_pow2(dA.ptr, n, n, n, 11, -30, 30, stream)
        L.hwyblas_amd_fill_pow2(dB.ptr, n, n, n, 12, -30, 30, stream)
    elif dist == "c":
        rs = np.power(10.0, np.random.default_rng(21).uniform(-8, 8, n))
        cs = np.power(10.0, np.random.default_rng(22).uniform(-8, 8, n))
        drs = H.DeviceArray.from_host(rs.reshape(-1, 1))
        dcs = H.DeviceArray.from_host(cs.reshape(-1, 1))
        dA.fill_uniform(n, n, 1, stream=stream)
        dB.fill_uniform(n, n, 2, stream=stream)
        L.hwyblas_amd_scale_rows_cols(dA.ptr, n, n, n, drs.ptr, None, stream)
        L.hwyblas_amd_scale_rows_cols(dB.ptr, n, n, n, None, dcs.ptr, stream)
    else:  # "a"
        dA.fill_uniform(n, n, SEEDS[0], stream=stream)
        dB.fill_uniform(n, n, SEEDS[1], stream=stream)
    L.hwyblas_amd_stream_synchronize(stream)
    flops = 2.0 * n ** 3

    def native_step():
        assert H.gemm_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n, stream=stream,
                          dtype=np.float64) == 0
    nat_ms, _ = time_steps(H, stream, native_step, steps, 1)
    R, Cc, sr, sc = error_entries(n, samples, 321)
    nat_vals = result_entries(H, dC, n, R, Cc, sr, sc)

    def emu_step(report=False):
        if method == "bf16":
            info, u = H.dgemm_emulated_bf16_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC,
                                                n, slices=nmod, stream=stream, report=report)
        else:
            info, u = H.dgemm_emulated_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n,
                                           nmoduli=nmod, stream=stream, report=report)
        assert info == 0
        return u
    emu_ms, _ = time_steps(H, stream, emu_step, steps, 1)
    used = [emu_step(report=True)]
    emu_vals = result_entries(H, dC, n, R, Cc, sr, sc)
    key = (dist, n, samples)
    if key not in _REFERENCE_CACHE:  # the sweeps reuse one distribution's reference
        if dist == "b":
            ha, hb = oracle.fill_pow2(n, n, 11), oracle.fill_pow2(n, n, 12)
        else:
            ha, hb = oracle.fill_uniform(n, n, 1), oracle.fill_uniform(n, n, 2)
            if dist == "c":
                ha *= rs[:, None]
                hb *= cs[None, :]
        _REFERENCE_CACHE[key] = reference_entries(oracle, ha, hb, n, R, Cc, sr, sc)
        del ha, hb
    exp, bnd = _REFERENCE_CACHE[key]
    e_nat, r_nat = error_summary(nat_vals, exp, bnd)
    e_emu, r_emu = error_summary(emu_vals, exp, bnd)
    count_key = "slices" if method == "bf16" else "moduli"
    out = {"distribution": DISTRIBUTIONS.get(dist, "uniform[-1,1] (seeds 1, 2)"),
           "method": METHODS[method],
           count_key: used[0], "emulated_ms": round(emu_ms, 3), "native_ms": round(nat_ms, 3),
           "emulated_tflops": round(flops / (emu_ms * 1e-3) / 1e12, 2),
           "native_tflops": round(flops / (nat_ms * 1e-3) / 1e12, 2),
           "emulated_max_componentwise_rel": e_emu, "native_max_componentwise_rel": e_nat,
           "emulated_max_rel": r_emu, "native_max_rel": r_nat,
           "ratio_emulated_to_native": e_emu / e_nat if e_nat > 0 else None,
           "ratio_rel_emulated_to_native": r_emu / r_nat if r_nat > 0 else None,
           "checked": f"16 full rows, 16 full columns and {samples} random entries "
                      f"({len(exp)} entries)"}
    log(f"dist {dist} ({method}): {used[0]} {count_key}, emulated {out['emulated_tflops']} TF vs native "
        f"{out['native_tflops']} TF; componentwise error emulated {e_emu:.3e} native {e_nat:.3e}; "
        f"relative error emulated {r_emu:.3e} native {r_nat:.3e}")
    return out


def householder_blocks(n, seed, b=128):
    """Stewart's (1980) random orthogonal Q = H_0 H_1 ... H_{n-2}: H_i reflects
    a Gaussian vector x (rows i..n-1) to -sign(x_0) |x| e_i; this is the Q of
    a Householder QR of a Gaussian matrix (equal in distribution). Yields
    (row0, V, T) per block of b reflectors in compact WY form, the product of
    the block's reflectors being I - V T V^T (V: rows row0..n-1)."""
    rng = np.random.default_rng(seed)
    for j0 in range(0, n - 1, b):
        nb = min(b, n - 1 - j0)
        rows = n - j0
        V = np.zeros((rows, nb), order="F")
        tau = np.zeros(nb)
        for t in range(nb):
            x = rng.standard_normal(rows - t)
            v = x.copy()
            v[0] += np.copysign(np.linalg.norm(x), x[0])
            V[t:, t] = v
            tau[t] = 2.0 / (v @ v)
        G = V.T @ V
        T = np.zeros((nb, nb), order="F")
        for t in range(nb):
            T[t, t] = tau[t]
            if t:
                T[:t, t] = -tau[t] * (T[:t, :t] @ G[:t, t])
        yield j0, V, T


def apply_orthogonal(H, dM, n, seed, side, stream):
    """M := Q M (side "L") or M := M Q^T (side "R") with Stewart's Q, through
    the library's DGEMM (blocks applied last to first)."""
    blocks = list(householder_blocks(n, seed))
    for j0, V, T in reversed(blocks):
        rows, nb = V.shape
        dV = H.DeviceArray.from_host(V)
        dT = H.DeviceArray.from_host(T)
        W = H.DeviceArray(nb, n) if side == "L" else H.DeviceArray(n, nb)
        W2 = H.DeviceArray(nb, n) if side == "L" else H.DeviceArray(n, nb)
        sub = dM.ptr + 8 * j0 if side == "L" else dM.ptr + 8 * j0 * n
        if side == "L":  # W = V^T M[j0:, :]; W = T W; M[j0:, :] -= V W
            H.gemm_dev("T", "N", nb, n, rows, 1.0, dV, rows, sub, n, 0.0, W, nb, stream=stream,
                       dtype=np.float64)
            H.gemm_dev("N", "N", nb, n, nb, 1.0, dT, nb, W, nb, 0.0, W2, nb, stream=stream,
                       dtype=np.float64)
            H.gemm_dev("N", "N", rows, n, nb, -1.0, dV, rows, W2, nb, 1.0, sub, n, stream=stream,
                       dtype=np.float64)
        else:  # Y = M[:, j0:] V; Y = Y T^T; M[:, j0:] -= Y V^T
            H.gemm_dev("N", "N", n, nb, rows, 1.0, sub, n, dV, rows, 0.0, W, n, stream=stream,
                       dtype=np.float64)
            H.gemm_dev("N", "T", n, nb, nb, 1.0, W, n, dT, nb, 0.0, W2, n, stream=stream,
                       dtype=np.float64)
            H.gemm_dev("N", "T", n, rows, nb, -1.0, W2, n, dV, rows, 1.0, sub, n, stream=stream,
                       dtype=np.float64)
        H.lib().hwyblas_amd_stream_synchronize(stream)
        del dV, dT, W, W2


def lu_leg(H, n, kappas, stream, nrhs=16, reps=2):
    """Mixed-precision LU solve: A = Q1 diag(sigma) Q2^T (sigma geometric from 1
    to 10^kappa, Q1/Q2 Stewart-random orthogonal, seeds 31/32), B uniform[-1,1]
    (seed 33), nrhs right-hand sides. For each trailing-update precision:
    getrf + getrs + FP64 refinement + the forward error bound, timed end to
    end (and without the bound); beside it the all-FP64 getrf + getrs
    without refinement."""
    L = H.lib()
    out = {}
    for kappa in kappas:
        t0 = time.perf_counter()
        dA = H.DeviceArray(n, n)
        sig = np.power(10.0, kappa * np.arange(n) / (n - 1))
        host = np.zeros(n)  # build diag(sigma) column by column on the device
        L.hwyblas_amd_memset(dA.ptr, 0, n * n * 8)
        for i in range(n):
            host[0] = sig[i]
            L.hwyblas_amd_memcpy(dA.ptr + 8 * (i + i * n), host.ctypes.data, 8, 1)
        apply_orthogonal(H, dA, n, 31, "L", stream)
        apply_orthogonal(H, dA, n, 32, "R", stream)
        dB = H.DeviceArray(n, nrhs)
        dB.fill_uniform(n, nrhs, 33)
        H.synchronize()
        gen_s = time.perf_counter() - t0
        res = {"generation_s": round(gen_s, 1)}
        dAF, dX = H.DeviceArray(n, n), H.DeviceArray(n, nrhs)
        dP = H.DeviceArray(n, 1, np.int32)
        ferr, berr, nbe = np.zeros(nrhs), np.zeros(nrhs), np.zeros(nrhs)
        iters = np.zeros(nrhs, dtype=np.int32)
        for update, name in ((0, "fp64"), (1, "emulated"), (2, "fp32")):
            times = {}
            for with_ferr in (False, True):  # gerfs computes the forward error bound too
                best = None
                for _ in range(reps):
                    t0 = time.perf_counter()
                    info = L.hwyblas_dgesv_ir_dev(
                        n, nrhs, dA.ptr, n, dAF.ptr, n, dP.ptr, dB.ptr, n, dX.ptr, n, update, 5,
                        ferr.ctypes.data if with_ferr else None, berr.ctypes.data,
                        nbe.ctypes.data, iters.ctypes.data, stream)
                    dt = time.perf_counter() - t0
                    best = dt if best is None else min(best, dt)
                times[with_ferr] = best
            best = times[True]
            res[name] = {"ms": round(best * 1e3, 1), "solves_per_s": round(1.0 / best, 3),
                         "ms_without_ferr": round(times[False] * 1e3, 1),
                         "berr_max": float(berr.max()), "nbe_max": float(nbe.max()),
                         "ferr_max": float(ferr.max()),
                         "iterations": [int(iters.min()), int(iters.max())], "info": int(info)}
            log(f"lu n={n} kappa=1e{kappa} {name}: {best * 1e3:.1f} ms with ferr "
                f"({times[False] * 1e3:.1f} without), berr {berr.max():.2e}, "
                f"nbe {nbe.max():.2e}, ferr {ferr.max():.2e}, "
                f"corrections {iters.min()}-{iters.max()}")
        # All-FP64 getrf + getrs, no refinement.
        dF = H.DeviceArray(n, n)
        X = H.DeviceArray(n, nrhs)
        best = None
        for _ in range(reps):
            L.hwyblas_amd_memcpy(dF.ptr, dA.ptr, n * n * 8, 3)
            L.hwyblas_amd_memcpy(X.ptr, dB.ptr, n * nrhs * 8, 3)
            t0 = time.perf_counter()
            ipiv, info = H.getrf_dev(dF, 0, stream=stream)
            L.hwyblas_dgetrs_dev(b"N", n, nrhs, dF.ptr, n, ipiv.ptr, X.ptr, n, stream)
            L.hwyblas_amd_stream_synchronize(stream)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        res["fp64_getrf_getrs_no_refinement"] = {"ms": round(best * 1e3, 1),
                                                 "solves_per_s": round(1.0 / best, 3)}
        log(f"lu n={n} kappa=1e{kappa} all-FP64 getrf+getrs: {best * 1e3:.1f} ms")
        out[f"kappa_1e{kappa}"] = res
        del dA, dB, dF, X, dAF, dX, dP
    return out


def zgemm_leg(H, n=8192, reps=3):
    """ZGEMM and CGEMM, then ZTRSM and CTRSM, at n on device operands (interleaved complex, each
    filled through its real 2n x n view, uniform [-1, 1]), alpha = 1 - 0.5i,
    beta = 0.5 + 0.25i; 8 n^3 real flops per product, against the FP64 (FP32)
    MFMA peak."""
    out = {}
    for name, rt, ct, peak in (("zgemm", np.float64, np.complex128, PEAK_FP64),
                               ("cgemm", np.float32, np.complex64, PEAK_FP32)):
        arrs = []
        for seed in (21, 22, 23):
            x = H.DeviceArray(2 * n, n, rt)
            x.fill_uniform(2 * n, n, seed)
            arrs.append(x)
        A, B, C = arrs
        H.synchronize()
        for ta, tb in (("N", "N"), ("C", "N"), ("N", "T")):
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                info = H.zgemm_dev(ta, tb, n, n, n, 1.0 - 0.5j, A.ptr, n, B.ptr, n, 0.5 + 0.25j,
                                   C.ptr, n, dtype=ct)
                H.synchronize()
                times.append(time.perf_counter() - t0)
                assert info == 0
            t = min(times)
            tf = 8.0 * n ** 3 / t / 1e12
            out[f"{name}_{ta}{tb}"] = {"n": n, "ms": round(t * 1e3, 2), "tflops": round(tf, 2),
                                       "frac_peak": round(tf / peak, 3)}
            log(f"{name} {ta}{tb} n={n}: {t * 1e3:.2f} ms, {tf:.1f} TFLOP/s")
        # TRSM on A's triangles with the diagonal set to 2n (well conditioned),
        # 4 n^3 real flops; B is solved in place rep after rep.
        tname = name.replace("gemm", "trsm")
        diag = np.array([2.0 * n, 0.0], dtype=rt)
        isz = np.dtype(rt).itemsize
        for i in range(n):
            H.lib().hwyblas_amd_memcpy(A.ptr + 2 * isz * (i + i * n), diag.ctypes.data,
                                       diag.nbytes, 1)
        H.synchronize()
        for side, uplo, tr in (("L", "L", "N"), ("R", "U", "C")):
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                info = H.ztrsm_dev(side, uplo, tr, "N", n, n, 1.0, A.ptr, n, C.ptr, n, dtype=ct)
                H.synchronize()
                times.append(time.perf_counter() - t0)
                assert info == 0
            t = min(times)
            tf = 4.0 * n ** 3 / t / 1e12
            out[f"{tname}_{side}{uplo}{tr}"] = {"n": n, "ms": round(t * 1e3, 2),
                                               "tflops": round(tf, 2),
                                               "frac_peak": round(tf / peak, 3)}
            log(f"{tname} {side}{uplo}{tr} n={n}: {t * 1e3:.2f} ms, {tf:.1f} TFLOP/s")
        del A, B, C, arrs
    return out


def host_path_leg(H, sizes=(8192, 16384), reps=3):
    """dgemm_ on host (pageable numpy) operands, the path Eigen and the LAPACK
    port use: NN, alpha = 1, beta = 0 (C is not copied in), synchronous,
    best of `reps` wall-clock calls. Beside it the same call on device
    operands and the pageable copy rates of one n x n matrix each way."""
    L = H.lib()
    out = {}
    for n in sizes:
        dA, dB, dC = (H.DeviceArray(n, n) for _ in range(3))
        dA.fill_uniform(n, n, 1)
        dB.fill_uniform(n, n, 2)
        H.synchronize()
        a, b = dA.to_host(), dB.to_host()
        c = np.zeros((n, n), order="F")

        def best(fn):
            fn()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t0)
            return min(ts)
        t_host = best(lambda: H.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n))
        t_dev = best(lambda: (H.gemm("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n),
                              H.synchronize()))
        nbytes = n * n * 8
        t_h2d = best(lambda: L.hwyblas_amd_memcpy(dC.ptr, a.ctypes.data, nbytes, 1))
        t_d2h = best(lambda: L.hwyblas_amd_memcpy(c.ctypes.data, dC.ptr, nbytes, 2))
        flops = 2.0 * n ** 3
        r = {"n": n, "host_ms": round(t_host * 1e3, 1),
             "host_tflops": round(flops / t_host / 1e12, 2),
             "device_ms": round(t_dev * 1e3, 1),
             "device_tflops": round(flops / t_dev / 1e12, 2),
             "h2d_gbs": round(nbytes / t_h2d / 1e9, 1), "d2h_gbs": round(nbytes / t_d2h / 1e9, 1),
             "pcie_bytes": 3 * nbytes}
        out[str(n)] = r
        log(f"dgemm_ host operands n={n}: {r['host_ms']} ms ({r['host_tflops']} TFLOP/s) "
            f"vs device {r['device_ms']} ms; pageable H2D {r['h2d_gbs']} GB/s, "
            f"D2H {r['d2h_gbs']} GB/s")
        del dA, dB, dC, a, b, c
    return out


def qr_leg(H, sizes=(8192, 16384), reps=2):
    """Householder QR (hwyblas_dgeqrf_dev) of a square uniform[-1, 1] matrix on
    the device, best of `reps` factorizations of fresh copies; flops
    (4/3) n^3 (LAPACK's count for m = n). Check: ||R||_F = ||A||_F (Q is
    orthogonal), to 1e-10 relative."""
    L = H.lib()
    out = {}
    for n in sizes:
        A = H.DeviceArray(n, n)
        A.fill_uniform(n, n, 41)
        W = H.DeviceArray(n, n)
        tau = H.DeviceArray(n, 1)
        H.synchronize()
        a_host = A.to_host()
        fro_a = float(np.linalg.norm(a_host))
        del a_host
        times = []
        for _ in range(reps):
            L.hwyblas_amd_memcpy(W.ptr, A.ptr, n * n * 8, 3)
            H.synchronize()
            t0 = time.perf_counter()
            info = H.geqrf_dev(W, tau)
            H.synchronize()
            times.append(time.perf_counter() - t0)
            assert info == 0
        r = np.triu(W.to_host())
        rel = abs(float(np.linalg.norm(r)) - fro_a) / fro_a
        t = min(times)
        tf = 4.0 / 3.0 * n ** 3 / t / 1e12
        out[str(n)] = {"ms": round(t * 1e3, 1), "tflops": round(tf, 2),
                       "frac_fp64_peak": round(tf / PEAK_FP64, 3), "norm_check": rel}
        log(f"dgeqrf n={n}: {t * 1e3:.1f} ms, {tf:.2f} TFLOP/s, | ||R||_F - ||A||_F | / ||A||_F "
            f"= {rel:.1e}")
        assert rel < 1e-10
        del A, W, tau, r
    return out


def vendor_leg(n=16384, reps=5):
    """The vendor libraries on the same products, for reference: torch.mm
    (hipBLASLt / rocBLAS) in FP64 and FP32 and torch._int_mm (hipBLASLt INT8,
    int32 result) at m = n = k = `n`, HIP-event timed, best of `reps` after a
    warm-up. Not this library's code path."""
    import torch
    dev = "cuda:0"
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    out = {}

    def best(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ts = []
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return min(ts)
    a8 = torch.randint(-127, 128, (n, n), dtype=torch.int8, device=dev, generator=g)
    b8 = torch.randint(-127, 128, (n, n), dtype=torch.int8, device=dev, generator=g)
    ms = best(lambda: torch._int_mm(a8, b8))
    out["int8"] = {"ms": round(ms, 3), "tops": round(2 * n ** 3 / ms / 1e9, 1)}
    del a8, b8
    for name, dt, peak in (("fp64", torch.float64, PEAK_FP64), ("fp32", torch.float32, 157.3)):
        a = torch.rand(n, n, dtype=dt, device=dev, generator=g)
        b = torch.rand(n, n, dtype=dt, device=dev, generator=g)
        ms = best(lambda: torch.mm(a, b))
        tf = 2 * n ** 3 / ms / 1e9
        out[name] = {"ms": round(ms, 3), "tflops": round(tf, 1), "frac_peak": round(tf / peak, 3)}
        del a, b
    torch.cuda.empty_cache()
    log(f"vendor n={n}: FP64 {out['fp64']['ms']} ms ({out['fp64']['tflops']} TFLOP/s), FP32 "
        f"{out['fp32']['ms']} ms ({out['fp32']['tflops']} TFLOP/s), INT8 {out['int8']['ms']} ms "
        f"({out['int8']['tops']} TOP/s)")
    return out


def potrf_leg(H, pn=16384, reps=2):
    """DPOTRF (lower) at pn on A = G G^T / pn + I (SPD), G uniform[-1, 1] seed
    11, formed on the device with the library's DGEMM; the best of `reps`
    factorizations of fresh copies, n^3 / 3 flops."""
    L = H.lib()
    d = np.float64
    G = H.DeviceArray(pn, pn, d)
    G.fill_uniform(pn, pn, 11)
    P = H.DeviceArray(pn, pn, d)
    Q = H.DeviceArray(pn, pn, d)
    L.hwyblas_amd_memset(P.ptr, 0, pn * pn * 8)
    one = np.ones(1)
    for i in range(pn):
        L.hwyblas_amd_memcpy(P.ptr + 8 * (i + i * pn), one.ctypes.data, 8, 1)
    H.gemm("N", "T", pn, pn, pn, 1.0 / pn, G, pn, G, pn, 1.0, P, pn, dtype=d)
    H.synchronize()
    times, info = [], None
    for _ in range(reps):
        L.hwyblas_amd_memcpy(Q.ptr, P.ptr, pn * pn * 8, 3)
        H.synchronize()
        t0 = time.perf_counter()
        info = H.potrf("L", pn, Q, pn, dtype=d)
        times.append(time.perf_counter() - t0)
    t = min(times)
    tf = pn ** 3 / 3.0 / t / 1e12
    log(f"dpotrf n={pn}: {t * 1e3:.1f} ms, {tf:.1f} TFLOP/s, info {info}")
    return {"n": pn, "ms": round(t * 1e3, 1), "tflops": round(tf, 2),
            "frac_fp64_peak": round(tf / PEAK_FP64, 3), "info": int(info)}


def level3_leg(H, n=8192, reps=3):
    """DSYMM (L, U), DSYRK (L, N), DSYR2K (L, N), DGEMMTR (L, N, T) and DTRMM
    (L, U, N, N) at m = n = k = `n` on device operands through the Fortran
    entries (synchronous; best of `reps` wall-clock calls), uniform[-1, 1],
    alpha = 1, beta = 0.5 (TRMM: alpha = 1). Algorithmic flops: SYMM 2 n^3,
    SYRK n^2 (n + 1), SYR2K 2 n^2 (n + 1), GEMMTR n^2 (n + 1), TRMM n^3; beside
    them DGEMM 2 n^3. Then DGEMV N and T at 16384 (bytes of A) and DPOTRF (L)
    at 16384 (n^3 / 3 flops). Check: SYRK's triangle equals the same tile
    configuration's DGEMM A A^T bit for bit."""
    L = H.lib()
    d = np.float64
    A, B, C, X = (H.DeviceArray(n, n, d) for _ in range(4))
    A.fill_uniform(n, n, 1)
    B.fill_uniform(n, n, 2)
    C.fill_uniform(n, n, 3)
    H.synchronize()

    def best(call):
        call()
        H.synchronize()
        t = None
        for _ in range(reps):
            t0 = time.perf_counter()
            call()
            H.synchronize()
            dt = time.perf_counter() - t0
            t = dt if t is None else min(t, dt)
        return t

    cases = {
        "dgemm": (lambda: H.gemm("N", "N", n, n, n, 1.0, A, n, B, n, 0.5, C, n, dtype=d),
                  2.0 * n ** 3),
        "dsymm": (lambda: H.symm("L", "U", n, n, 1.0, A, n, B, n, 0.5, C, n, dtype=d),
                  2.0 * n ** 3),
        "dsyrk": (lambda: H.syrk("L", "N", n, n, 1.0, A, n, 0.5, C, n, dtype=d),
                  float(n) * n * (n + 1)),
        "dsyr2k": (lambda: H.syr2k("L", "N", n, n, 1.0, A, n, B, n, 0.5, C, n, dtype=d),
                   2.0 * n * n * (n + 1)),
        "dgemmtr": (lambda: H.gemmtr("L", "N", "T", n, n, 1.0, A, n, B, n, 0.5, C, n, dtype=d),
                    float(n) * n * (n + 1)),
        "dtrmm": (lambda: H.trmm("L", "U", "N", "N", n, n, 1.0, A, n, X, n, dtype=d),
                  float(n) ** 3),
    }
    out = {"n": n}
    for name, (call, flops) in cases.items():
        if name == "dtrmm":
            X.fill_uniform(n, n, 4)
            H.synchronize()
        t = best(call)
        tf = flops / t / 1e12
        out[name] = {"ms": round(t * 1e3, 2), "tflops": round(tf, 2),
                     "frac_fp64_peak": round(tf / PEAK_FP64, 3)}
        log(f"level3 {name} n={n}: {t * 1e3:.2f} ms, {tf:.1f} TFLOP/s")
    # DGEMV at 16384 x 16384 (HBM-bound: 8 m n bytes of A per call).
    gm = 16384
    GA = H.DeviceArray(gm, gm, d)
    GA.fill_uniform(gm, gm, 5)
    gx, gy = H.DeviceArray(gm, 1, d), H.DeviceArray(gm, 1, d)
    gx.fill_uniform(gm, 1, 6)
    gy.fill_uniform(gm, 1, 7)
    H.synchronize()
    for tr in "NT":
        t = best(lambda: H.gemv(tr, gm, gm, 1.0, GA, gm, gx, 1, 0.5, gy, 1, dtype=d))
        gbs = 8.0 * gm * gm / t / 1e9
        out[f"dgemv_{tr}"] = {"n": gm, "ms": round(t * 1e3, 3), "gb_per_s": round(gbs, 1),
                              "frac_hbm_peak": round(gbs / (PEAK_HBM * 1e3), 3)}
        log(f"level3 dgemv {tr} n={gm}: {t * 1e3:.3f} ms, {gbs:.0f} GB/s")
    # ZGEMV on the same memory read as a complex 8192 x 16384 matrix (ld 8192
    # complex): 16 m n bytes of A per call; x and y hold 16384 complex
    # elements each (enough for either op).
    zm, zn = gm // 2, gm
    zx, zy = H.DeviceArray(2 * gm, 1, d), H.DeviceArray(2 * gm, 1, d)
    zx.fill_uniform(2 * gm, 1, 8)
    zy.fill_uniform(2 * gm, 1, 9)
    H.synchronize()
    for tr in "NC":
        t = best(lambda: H.zgemv(tr, zm, zn, 1.0 - 0.5j, GA.ptr, zm, zx.ptr, 1, 0.5, zy.ptr, 1,
                                 dtype=np.complex128))
        gbs = 16.0 * zm * zn / t / 1e9
        out[f"zgemv_{tr}"] = {"m": zm, "n": zn, "ms": round(t * 1e3, 3),
                              "gb_per_s": round(gbs, 1),
                              "frac_hbm_peak": round(gbs / (PEAK_HBM * 1e3), 3)}
        log(f"level3 zgemv {tr} {zm}x{zn}: {t * 1e3:.3f} ms, {gbs:.0f} GB/s")
    del GA, zx, zy
    out["dpotrf"] = potrf_leg(H)
    out.update(zgemm_leg(H))
    # SYRK's lower triangle against DGEMM A A^T (beta = 0 on both).
    C.fill_uniform(n, n, 3)
    H.syrk("L", "N", n, n, 1.0, A, n, 0.0, C, n, dtype=d)
    G = H.DeviceArray(n, n, d)
    H.gemm("N", "T", n, n, n, 1.0, A, n, A, n, 0.0, G, n, dtype=d)
    H.synchronize()
    c, g = C.to_host(), G.to_host()
    low = np.tril(np.ones((n, n), dtype=bool))
    out["syrk_equals_gemm_triangle"] = bool(np.array_equal(c[low], g[low]))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def worker_cores():
    """The host cores this process may use, at most 16 (one GPU's share)."""
    return sorted(os.sched_getaffinity(0))[:16]


def pin(core):
    try:
        os.sched_setaffinity(0, {core})
    except OSError:
        pass


def best_time(fn, seconds, min_calls=2):
    fn()  # warm-up
    best, calls = float("inf"), 0
    t_end = time.perf_counter() + seconds
    while calls < min_calls or time.perf_counter() < t_end:
        t = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t)
        calls += 1
    return best, calls


def cpu_gemm_leg(oracle, seconds):
    """The oracle's single-threaded restatement of hwyBLAS's blocked DGEMM,
    NN 2048^3 on one pinned core, best of the calls that fit in about
    `seconds` (bench/blas_bench.cc:109-133's method)."""
    n = 2048
    a, b = oracle.fill_uniform(n, n, 1), oracle.fill_uniform(n, n, 2)
    c = np.zeros((n, n), order="F")
    best, calls = best_time(lambda: oracle.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n),
                            seconds)
    return {"value": round(2.0 * n ** 3 / best / 1e12, 5), "unit": "TFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle restatement of hwyBLAS's blocked DGEMM (AVX2 8x6 FMA tile, "
                      f"KC=256, beta on the first K block), NN {n}^3, alpha=1 beta=0, best of "
                      f"{calls} calls, 1 pinned core"}


def cpu_trsm_leg(oracle, seconds):
    """The restatement of hwyBLAS's recursive DTRSM (halving at multiples of
    32, dot-form base with the reciprocal), L/L/N/N 4096 x 4096, the
    diagonal n + U[0, 1], one pinned core."""
    n = 4096
    t = oracle.fill_uniform(n, n, 3)
    t[np.arange(n), np.arange(n)] = n + (oracle.fill_uniform(n, 1, 4)[:, 0] + 1) / 2
    b = oracle.fill_uniform(n, n, 5)
    x = b.copy(order="F")

    def step():
        x[:] = b
        oracle.trsm("L", "L", "N", "N", n, n, 1.0, t, n, x, n)
    best, calls = best_time(step, seconds)
    return {"value": round(float(n) ** 3 / best / 1e12, 5), "unit": "TFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle restatement of hwyBLAS's recursive DTRSM L/L/N/N {n} x {n} "
                      f"(m^2 n flops), best of {calls} calls (the copy of B included), 1 pinned "
                      f"core"}


def _batched_part(job):
    """One process's share of a batch: generate its problems (untimed), then
    time them with the oracle's restatement on the pinned core."""
    kind, dtype, core, shapes, seed = job
    import oracle
    pin(core)
    ops = []
    for t, sh in enumerate(shapes):
        if kind == "gemm":
            m, n, k, ta, tb = sh
            ra, ca = (k, m) if ta else (m, k)
            rb, cb = (n, k) if tb else (k, n)
            ops.append((m, n, k, ta, tb, oracle.fill_uniform(ra, ca, seed + 3 * t, dtype=dtype),
                        oracle.fill_uniform(rb, cb, seed + 3 * t + 1, dtype=dtype),
                        np.zeros((m, n), dtype=dtype, order="F")))
        else:
            m, n = sh
            a = oracle.fill_uniform(m, m, seed + 2 * t)
            a[np.arange(m), np.arange(m)] = m + (a[np.arange(m), np.arange(m)] + 1) / 2
            ops.append((m, n, a, oracle.fill_uniform(m, n, seed + 2 * t + 1)))
    def once():
        t0 = time.perf_counter()
        for o in ops:
            if kind == "gemm":
                m, n, k, ta, tb, a, b, c = o
                oracle.gemm("T" if ta else "N", "T" if tb else "N", m, n, k, 1.0, a, a.shape[0],
                            b, b.shape[0], 0.0, c, m)
            else:
                m, n, a, b = o
                oracle.trsm("L", "L", "N", "N", m, n, 1.0, a, m, b, m)
        return time.perf_counter() - t0
    # Best of the repetitions that fit in about 2 s (at least 3).
    times = [once()]
    while len(times) < 3 or sum(times) < 2.0:
        times.append(once())
    return min(times), len(times)


def cpu_batched_leg(kind, dtype, cores, count=10000):
    """The batched workload on the host: the same seeded shapes as the GPU
    leg, split round-robin over one process per core (each pinned), all
    running at once; the batch time is the slowest process's."""
    import multiprocessing
    if kind == "gemm":
        rng = np.random.default_rng(7)
        m, n, k = (log_uniform_ints(rng, 8, 512, count) for _ in range(3))
        ta = rng.integers(0, 2, count)
        tb = rng.integers(0, 2, count)
        shapes = [(int(m[t]), int(n[t]), int(k[t]), int(ta[t]), int(tb[t])) for t in range(count)]
        flops = float(np.sum(2.0 * m * n * k))
    else:
        rng = np.random.default_rng(11)
        m = log_uniform_ints(rng, 8, 256, count)
        n = log_uniform_ints(rng, 1, 64, count)
        shapes = [(int(m[t]), int(n[t])) for t in range(count)]
        flops = float(np.sum(m.astype(float) ** 2 * n))
    jobs = [(kind, dtype, c, shapes[i::len(cores)], 5000 + 100000 * i)
            for i, c in enumerate(cores)]
    with multiprocessing.get_context("fork").Pool(len(cores)) as pool:
        parts = pool.map(_batched_part, jobs)
    secs = max(t for t, _ in parts)
    reps = min(r for _, r in parts)
    unit = "GEMM/s" if kind == "gemm" else "solves/s"
    name = {"gemm": f"{'D' if np.dtype(dtype) == np.float64 else 'S'}GEMM",
            "trsm": "DTRSM L/L/N/N"}[kind]
    return {"value": round(count / secs, 1), "unit": unit, "cores": len(cores), "kind": "port",
            "tflops": round(flops / secs / 1e12, 5),
            "sample": f"the {count} seeded uneven {name} problems of the GPU leg, split over "
                      f"{len(cores)} processes, one pinned core each, by the oracle's "
                      f"restatement, all at once; each process's best of {reps}+ repetitions, "
                      f"batch time = the slowest process's ({secs:.4f} s)"}


def cpu_baseline_worker():
    """Runs in a process started before the benchmark touches the GPU: waits
    for "go" on stdin (sent once the GPU legs are done, so that nothing else
    runs on the host meanwhile), then times every CPU leg and prints one
    JSON line."""
    import oracle
    oracle.build()
    if sys.stdin.readline().strip() != "go":
        return
    cores = worker_cores()
    out = {}
    pin(cores[0])
    out["dgemm"] = cpu_gemm_leg(oracle, 8.0)
    out["dtrsm"] = cpu_trsm_leg(oracle, 1.0)
    try:
        os.sched_setaffinity(0, set(cores))
    except OSError:
        pass
    out["batched_gemm_f64"] = cpu_batched_leg("gemm", np.float64, cores)
    out["batched_gemm_f32"] = cpu_batched_leg("gemm", np.float32, cores)
    out["batched_trsm_f64"] = cpu_batched_leg("trsm", np.float64, cores)
    out["host"] = {"cpu_model": cpu_model(), "cores_used": len(cores),
                   "cores_visible": os.cpu_count()}
    print(json.dumps(out), flush=True)


def run_batched(H, oracle, args, stream):
    out = {}
    model = None
    if not args.no_model:
        from hwyblas_amd import tune
        try:
            model = tune.load_model()
        except (OSError, ValueError) as e:
            log(f"no cost model ({e}); built-in configuration rule only")
    for dtype, name in ((np.float64, "gemm_f64"), (np.float32, "gemm_f32")):
        r = batched_gemm_leg(H, oracle, dtype, args.batch_count, max(3, args.steps),
                             args.warmup, stream, model=model, with_loop=not args.no_loop)
        log(f"batched {name}: {r['gemm_per_s']:.0f} GEMM/s ({r['grouped_ms']:.2f} ms, configs "
            f"{r['configs']}; built-in rule {r['grouped_ms_default_configs']:.2f} ms), loop "
            f"{r['loop_ms']:.2f} ms ({r['speedup_vs_loop']:.2f}x), loop graph "
            f"{r['loop_graph_ms']:.2f} ms ({r['speedup_vs_loop_graph']:.2f}x), "
            f"{r['tflops']:.2f} TFLOP/s, "
            f"worst bound ratio {r['worst_bound_ratio']:.3g}")
        out[name] = r
    r = batched_trsm_leg(H, oracle, args.batch_count, max(3, args.steps), args.warmup, stream,
                         with_loop=not args.no_loop)
    log(f"batched trsm_f64: {r['solves_per_s']:.0f} solves/s ({r['grouped_ms']:.2f} ms), loop "
        f"{r['loop_ms']:.2f} ms ({r['speedup_vs_loop']:.2f}x), loop graph "
        f"{r['loop_graph_ms']:.2f} ms ({r['speedup_vs_loop_graph']:.2f}x), worst residual ratio "
        f"{r['worst_residual_ratio']:.3g}")
    out["trsm_f64"] = r
    if model is not None:
        H.lib().hwyblas_grouped_set_model(None)
        model.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nmod", type=int, default=0, help="moduli for the emulation (0 = auto)")
    ap.add_argument("--samples", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-only", action="store_true")
    ap.add_argument("--plain-sweep", action="store_true",
                    help="S/DGEMM at 4096-16384 (every transposition, both layouts) and DTRSM "
                         "8192 (every side/uplo/trans) instead of the headline")
    ap.add_argument("--sweep-sizes", default="4096,8192,16384",
                    help="comma-separated m = n = k for --plain-sweep's S/DGEMM")
    ap.add_argument("--batch-count", type=int, default=10000)
    ap.add_argument("--lu", action="store_true",
                    help="the mixed-precision LU solve (n = --lu-n, kappa 1e2, 1e6, 1e10) "
                         "instead of the headline")
    ap.add_argument("--level3", action="store_true",
                    help="DSYMM/DSYRK/DSYR2K/DGEMMTR/DTRMM at n = 8192 beside DGEMM, "
                         "DGEMV and DPOTRF at 16384, ZGEMM/CGEMM and ZTRSM/CTRSM at 8192")
    ap.add_argument("--host-path", action="store_true",
                    help="dgemm_ on host (pageable) operands at 8192 and 16384 beside the "
                         "device-operand call, and the pageable copy rates")
    ap.add_argument("--qr", action="store_true",
                    help="Householder QR (dgeqrf) of square matrices at 8192 and 16384")
    ap.add_argument("--vendor", action="store_true",
                    help="the vendor libraries' FP64/FP32/INT8 GEMM at 16384 through torch, for "
                         "reference")
    ap.add_argument("--potrf", action="store_true",
                    help="only DPOTRF (lower) at n = 16384")
    ap.add_argument("--zgemm", action="store_true",
                    help="only ZGEMM/CGEMM and ZTRSM/CTRSM at n = 8192")
    ap.add_argument("--lu-n", type=int, default=16384)
    ap.add_argument("--lu-kappas", default="2,6,10",
                    help="comma-separated log10 condition numbers for --lu")
    ap.add_argument("--dists", default="bc",
                    help="wide-range operand distributions to run after the headline (b, c)")
    ap.add_argument("--moduli-sweep", default="",
                    help="comma-separated moduli counts: error and speed of the emulation at "
                         "each, uniform operands")
    ap.add_argument("--bf16", default="a",
                    help="distributions (a, b, c) to run the BF16-slice emulation on, slice "
                         "count from the operands")
    ap.add_argument("--bf16-sweep", default="",
                    help="comma-separated BF16 slice counts: error and speed at each, uniform "
                         "operands")
    ap.add_argument("--no-loop", action="store_true",
                    help="batched legs without the loop-of-single-calls comparison")
    ap.add_argument("--no-model", action="store_true",
                    help="batched GEMM with the built-in configuration rule only")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device to device (a kernel of known traffic, "
                         "to scale rocprofv3 FETCH_SIZE/WRITE_SIZE readings)")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_baseline_worker()
        return
    if args.vendor:
        # The vendor libraries run through torch's own HIP runtime, which
        # cannot share a process with this library's: measured first, alone.
        print(json.dumps({"vendor": vendor_leg()}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the hot path is "
              "single-GPU by design; run without --gpus or with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # The CPU baseline runs in its own process, started before anything here
    # touches the GPU; it waits until the GPU legs are done.
    cpu_proc = None
    if not args.no_cpu and not (args.lu or args.plain_sweep or args.level3 or args.potrf or
                                args.zgemm or args.batched_only or args.host_path or
                                args.vendor or args.qr):
        import subprocess
        cpu_proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker"],
                                    stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)

    import hwyblas_amd as H
    import oracle

    n, steps, warmup = args.n, args.steps, args.warmup
    L = H.lib()
    stream = L.hwyblas_amd_stream_create()
    dA, dB, dC = H.DeviceArray(n, n), H.DeviceArray(n, n), H.DeviceArray(n, n)
    for d, s in zip((dA, dB, dC), SEEDS):
        d.fill_uniform(n, n, s, stream=stream)
    L.hwyblas_amd_stream_synchronize(stream)
    flops = 2.0 * n ** 3
    if args.calibrate_copy:
        src, dst = H.DeviceArray(1 << 27, 1), H.DeviceArray(1 << 27, 1)  # 2 x 1 GiB
        L.hwyblas_amd_memcpy(dst.ptr, src.ptr, 1 << 30, 3)
        H.synchronize()
        del src, dst

    if args.lu:
        kappas = tuple(int(x) for x in args.lu_kappas.split(",") if x)
        print(json.dumps({"lu": lu_leg(H, args.lu_n, kappas, stream)}), flush=True)
        return
    if args.plain_sweep:
        sizes = tuple(int(x) for x in args.sweep_sizes.split(",") if x)
        print(json.dumps({"plain_sweep": plain_sweep(H, oracle, sizes)}), flush=True)
        return
    if args.level3:
        print(json.dumps({"level3": level3_leg(H)}), flush=True)
        return
    if args.potrf:
        print(json.dumps({"dpotrf": potrf_leg(H)}), flush=True)
        return
    if args.host_path:
        print(json.dumps({"host_path": host_path_leg(H)}), flush=True)
        return
    if args.qr:
        print(json.dumps({"dgeqrf": qr_leg(H)}), flush=True)
        return
    if args.zgemm:
        print(json.dumps({"zgemm": zgemm_leg(H)}), flush=True)
        return
    if args.batched_only:
        out = run_batched(H, oracle, args, stream)
        print(json.dumps({"batched": out}), flush=True)
        return

    # Native FP64 MFMA DGEMM.
    def native_step():
        assert H.gemm_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n, stream=stream,
                          dtype=np.float64) == 0
    nat_ms, _ = time_steps(H, stream, native_step, max(2, steps // 2), 1)
    nat_tf = flops / (nat_ms * 1e-3) / 1e12
    log(f"native DGEMM {nat_ms:.2f} ms/step = {nat_tf:.2f} TFLOP/s")
    R, Cc, sr, sc = error_entries(n, args.samples, 123)
    nat_vals = result_entries(H, dC, n, R, Cc, sr, sc) if not args.no_check else None

    # Emulated DGEMM (the headline): timed without the count pointer (the
    # call does not synchronize); one more call reports the count.
    def emu_step(report=False):
        info, u = H.dgemm_emulated_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n,
                                       nmoduli=args.nmod, stream=stream, report=report)
        assert info == 0
        return u
    emu_ms, wall = time_steps(H, stream, emu_step, steps, warmup)
    used = [emu_step(report=True)]
    emu_tf = flops / (emu_ms * 1e-3) / 1e12
    log(f"emulated DGEMM ({used[0]} moduli) {emu_ms:.2f} ms/step = {emu_tf:.2f} TFLOP/s, "
        f"{emu_tf / nat_tf:.2f}x native; wall {wall:.2f}s")
    # Phase breakdown of one more call (events on the same stream).
    L.hwyblas_amd_profile_enable(1)
    emu_step()
    L.hwyblas_amd_stream_synchronize(stream)
    phases = {name: round(L.hwyblas_amd_profile_ms(i), 3)
              for i, name in enumerate(("stats", "slice", "int8_gemm", "reconstruct"))}
    L.hwyblas_amd_profile_enable(0)
    log(f"phases (ms): {phases}")
    int8_ops = used[0] * 2.0 * n ** 3
    int8_tops = int8_ops / (phases["int8_gemm"] * 1e-3) / 1e12

    err = None
    if not args.no_check:
        emu_vals = result_entries(H, dC, n, R, Cc, sr, sc)
        ha = oracle.fill_uniform(n, n, SEEDS[0])
        hb = oracle.fill_uniform(n, n, SEEDS[1])
        exp, bnd = reference_entries(oracle, ha, hb, n, R, Cc, sr, sc)
        del ha, hb
        e_nat, r_nat = error_summary(nat_vals, exp, bnd)
        e_emu, r_emu = error_summary(emu_vals, exp, bnd)
        err = {"emulated_max_componentwise_rel": e_emu, "native_max_componentwise_rel": e_nat,
               "ratio_emulated_to_native": e_emu / e_nat if e_nat > 0 else None,
               "emulated_max_rel": r_emu, "native_max_rel": r_nat,
               "ratio_rel_emulated_to_native": r_emu / r_nat if r_nat > 0 else None,
               "checked": f"16 full rows, 16 full columns and {args.samples} random entries "
                          f"({len(exp)} entries)",
               "reference": "double-double (oracle)",
               "bound_4(k+2)u": 4 * (n + 2) * 2.0 ** -53}
        log(f"error: componentwise emulated {e_emu:.3e} native {e_nat:.3e}; relative emulated "
            f"{r_emu:.3e} native {r_nat:.3e}")

    dists = {}
    if not args.no_check:
        for dist in args.dists:
            dists[dist] = distribution_leg(H, oracle, dist, n, dA, dB, dC, stream, args.samples)
        for nm in [int(x) for x in args.moduli_sweep.split(",") if x]:
            dists[f"a_moduli_{nm}"] = distribution_leg(H, oracle, "a", n, dA, dB, dC, stream,
                                                       args.samples, nmod=nm)
        for dist in args.bf16:
            dists[f"{dist}_bf16"] = distribution_leg(H, oracle, dist, n, dA, dB, dC, stream,
                                                     args.samples, method="bf16")
        for sl in [int(x) for x in args.bf16_sweep.split(",") if x]:
            dists[f"a_bf16_slices_{sl}"] = distribution_leg(H, oracle, "a", n, dA, dB, dC, stream,
                                                            args.samples, nmod=sl, method="bf16")
    del dA, dB, dC

    batched = None if args.no_batched else run_batched(H, oracle, args, stream)

    cpu = None
    if cpu_proc is not None:
        cpu_proc.stdin.write("go\n")
        cpu_proc.stdin.flush()
        legs = json.loads(cpu_proc.stdout.readline())
        cpu_proc.wait()
        cpu = dict(legs["dgemm"])
        cpu["host"] = legs["host"]
        cpu["legs"] = {k: v for k, v in legs.items() if k not in ("dgemm", "host")}
        log(f"cpu baseline {cpu['value'] * 1e3:.2f} GFLOP/s (DGEMM 2048, 1 core); DTRSM "
            f"{legs['dtrsm']['value'] * 1e3:.2f} GFLOP/s; batched DGEMM "
            f"{legs['batched_gemm_f64']['value']:.0f} GEMM/s on {legs['host']['cores_used']} "
            f"cores ({legs['host']['cpu_model']})")

    out = {
        "metric": METRIC,
        "value": round(emu_tf, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(emu_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"emulated DGEMM NN column-major m=n=k={n}, lda=ldb=ldc={n}, "
                               f"uniform[-1,1] seeds 1/2/3, alpha=1 beta=0, FP64 in/out, "
                               f"device-resident",
                   "method": f"INT8 Ozaki scheme II, {used[0]} moduli (auto), "
                             f"v_mfma_i32_32x32x32_i8",
                   "native_tflops": round(nat_tf, 3),
                   "native_ms_per_step": round(nat_ms, 4),
                   "speedup_vs_native": round(emu_tf / nat_tf, 3),
                   "native_roofline_frac_fp64": round(nat_tf / PEAK_FP64, 4),
                   "phases_ms": phases},
        "roofline": {"bound": "mfma", "kernel": "ModGemmDmaKernel (INT8, all moduli)",
                     "achieved": round(int8_tops, 2), "peak": PEAK_INT8, "unit": "TFLOP/s",
                     "ops": "int8 multiply-adds counted as 2 ops (TOP/s); algorithmic "
                            "2 m n k per modulus",
                     "frac": round(int8_tops / PEAK_INT8, 4),
                     "traffic": measured_traffic(n, used[0]),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                     "copy-calibrated)",
                     "algorithmic_bytes": used[0] * 3 * n * n},
        "error": err,
        "distributions": dists,
        "batched": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    L.hwyblas_amd_stream_destroy(stream)


if __name__ == "__main__":
    main()
