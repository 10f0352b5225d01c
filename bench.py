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
PEAK_INT8 = 5000.0    # MI355X dense INT8 matrix, TOP/s
SEEDS = (1, 2, 3)


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


def cpu_baseline(oracle, seconds=12.0):
    """The oracle's single-threaded restatement of hwyBLAS's blocked DGEMM,
    DGEMM NN 2048^3 on one pinned core, best of the calls that fit in about
    `seconds` (at least 2), as bench/blas_bench.cc:109-133 times it."""
    n = 2048
    a = oracle.fill_uniform(n, n, 1)
    b = oracle.fill_uniform(n, n, 2)
    c = np.zeros((n, n), order="F")
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except Exception:
        pass
    oracle.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n)  # warm-up
    best, calls = float("inf"), 0
    t_end = time.perf_counter() + seconds
    while calls < 2 or time.perf_counter() < t_end:
        t = time.perf_counter()
        oracle.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 0.0, c, n)
        best = min(best, time.perf_counter() - t)
        calls += 1
    return {"value": round(2.0 * n ** 3 / best / 1e12, 5), "unit": "TFLOP/s", "cores": 1,
            "kind": "port",
            "sample": f"oracle restatement of hwyBLAS's blocked DGEMM (AVX2 8x6 FMA tile, "
                      f"KC=256, beta on the first K block), NN {n}^3, alpha=1 beta=0, best of "
                      f"{calls} calls, 1 pinned core"}


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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device to device (a kernel of known traffic, "
                         "to scale rocprofv3 FETCH_SIZE/WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the hot path is "
              "single-GPU by design; run without --gpus or with --gpus 1.", file=sys.stderr)
        sys.exit(2)

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

    # Native FP64 MFMA DGEMM.
    def native_step():
        assert H.gemm_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n, stream=stream,
                          dtype=np.float64) == 0
    nat_ms, _ = time_steps(H, stream, native_step, max(2, steps // 2), 1)
    nat_tf = flops / (nat_ms * 1e-3) / 1e12
    log(f"native DGEMM {nat_ms:.2f} ms/step = {nat_tf:.2f} TFLOP/s")
    rng = np.random.default_rng(123)
    rows = rng.integers(0, n, args.samples)
    cols = rng.integers(0, n, args.samples)
    nat_vals = gather(H, dC, n, rows, cols) if not args.no_check else None

    # Emulated DGEMM (the headline).
    used = [0]

    def emu_step():
        info, u = H.dgemm_emulated_dev("N", "N", n, n, n, 1.0, dA, n, dB, n, 0.0, dC, n,
                                       nmoduli=args.nmod, stream=stream)
        assert info == 0
        used[0] = u
    emu_ms, wall = time_steps(H, stream, emu_step, steps, warmup)
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
        emu_vals = gather(H, dC, n, rows, cols)
        ha = oracle.fill_uniform(n, n, SEEDS[0])
        hb = oracle.fill_uniform(n, n, SEEDS[1])
        exp, bnd = oracle.gemm_ref_entries("N", "N", n, 1.0, ha, n, hb, n, 0.0, None, n, rows,
                                           cols)
        del ha, hb
        e_nat = float(np.max(np.abs(nat_vals - exp) / bnd))
        e_emu = float(np.max(np.abs(emu_vals - exp) / bnd))
        err = {"emulated_max_componentwise_rel": e_emu, "native_max_componentwise_rel": e_nat,
               "ratio_emulated_to_native": e_emu / e_nat if e_nat > 0 else None,
               "samples": int(args.samples), "reference": "double-double (oracle)",
               "bound_4(k+2)u": 4 * (n + 2) * 2.0 ** -53}
        log(f"sampled error: emulated {e_emu:.3e}, native {e_nat:.3e}")

    cpu = None if args.no_cpu else cpu_baseline(oracle)
    if cpu:
        log(f"cpu baseline {cpu['value'] * 1e3:.2f} GFLOP/s")

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
        "roofline": {"bound": "mfma", "kernel": "ModGemmKernel (INT8, all moduli)",
                     "achieved": round(int8_tops, 2), "peak": PEAK_INT8, "unit": "TFLOP/s",
                     "ops": "int8 multiply-adds counted as 2 ops (TOP/s); algorithmic "
                            "2 m n k per modulus",
                     "frac": round(int8_tops / PEAK_INT8, 4), "traffic": None},
        "error": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    L.hwyblas_amd_stream_destroy(stream)


if __name__ == "__main__":
    main()
