#!/usr/bin/env python3
"""Headline benchmark: SGEMM 8192^3 in the BF16x3 split form vs plain FP32 MFMA,
at their measured FP64-oracle error, plus DGEMV N 16384^2 GB/s, on one MI355X.

One step = one coral_sgemm_bf16x3 call (NN, m = n = k = 8192, f32 N(0,1) operands
already resident in HBM; the call includes its split/pack pre-passes).  `value` is
the FP32-equivalent throughput 2n^3 / step time.  The plain FP32-MFMA coral_sgemm
runs on the same operands beside it, and the error of both against an FP64 oracle
is measured on a 256-row sample.  DGEMV N (m = n = 16384, alpha = 1, beta = 0) is
timed as well.

`roofline` describes the dominant kernel (the BF16x3 MFMA GEMM kernel): its
duration comes from HIP events the library records around that launch on its
launch stream (coral_timing_*), its work is 2n^3 FP32-equivalent flops per launch
and its peak is the BF16 dense MFMA peak over the three products (2500/3 TFLOP/s).

`cpu_baseline` times the oracle's restatement of CORAL's CPU SGEMM (the Goto
5-loop with CORAL's blocking, oracle/coral_port.c) on one host core on a bounded
sample (SGEMM NN 3072^3, median of 3 runs).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SGEMM 8192^3 TFLOP/s BF16x3 vs FP32 MFMA at equal FP64-oracle error; GEMV GB/s"
PEAK_BF16 = 2500.0  # dense BF16 MFMA TFLOP/s (MI355X)
PEAK_FP32 = 157.3
PEAK_HBM = 8000.0  # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(fn, steps, warmup):
    from coral_amd import device

    for _ in range(warmup):
        fn()
    device.synchronize()
    e0, e1 = device.Event(), device.Event()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    e1.synchronize()
    device.synchronize()
    return e0.elapsed_ms(e1) / steps


def kernel_avg_ms(kind):
    import ctypes

    from coral_amd._lib import call

    tot, cnt = ctypes.c_double(), ctypes.c_int64()
    call("coral_timing_read", kind, ctypes.byref(tot), ctypes.byref(cnt))
    return (tot.value / cnt.value if cnt.value else float("nan")), cnt.value


def sample_error(A, B, C, rows, k):
    """Normwise |E|_F / (k eps |A|_F |B|_F) estimated from a row sample, and
    max |E| / max |C_ref| on that sample, against float64."""
    ref = A[rows, :].astype(np.float64) @ B.astype(np.float64)
    E = C[rows, :].astype(np.float64) - ref
    scale = np.sqrt(A.shape[0] / len(rows))
    fro_a = np.linalg.norm(A.astype(np.float64))
    fro_b = np.linalg.norm(B.astype(np.float64))
    eps = 2.0 ** -24
    normwise = scale * np.linalg.norm(E) / (k * eps * fro_a * fro_b)
    rel_fro = scale * np.linalg.norm(E) / (fro_a * fro_b)
    return {"normwise_bound_ratio": float(normwise), "fro_rel_to_AB": float(rel_fro),
            "max_abs_over_max_ref": float(np.abs(E).max() / np.abs(ref).max())}


def cpu_baseline(n=3072, runs=3):
    """CORAL's CPU SGEMM algorithm (oracle restatement) on one core."""
    from oracle import port

    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    rng = np.random.default_rng(0)
    a = rng.standard_normal(n * n).astype(np.float32)
    b = rng.standard_normal(n * n).astype(np.float32)
    c = np.zeros(n * n, np.float32)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        port.sgemm(0, 0, n, n, n, 1.0, a, n, b, n, 0.0, c, n)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {"value": 2.0 * n ** 3 / t / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"CORAL SGEMM NN algorithm restated in C (oracle/coral_port.c: MC/NC/KC 384/576/256, "
                      f"8x12 FMA microkernel), m=n=k={n} f32 N(0,1), median of {runs} runs on 1 host core "
                      f"({os.cpu_count()} visible), {t:.2f} s per run"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--gemv-n", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {max(args.gpus, world)}).",
              file=sys.stderr)
        sys.exit(2)

    from coral_amd import device, synthetic
    from coral_amd._lib import call, load

    lib = load()
    st = lib.coral_init()
    if st != 0:
        raise SystemExit(f"coral_init failed: {lib.coral_last_error().decode()}")
    log("device:", device.device_name())

    n = args.n
    A = device.DeviceArray(n * n, np.float32)
    B = device.DeviceArray(n * n, np.float32)
    C = device.DeviceArray(n * n, np.float32)
    C2 = device.DeviceArray(n * n, np.float32)
    device.fill_random(A, synthetic.seed_for(3, "A"), "normal")
    device.fill_random(B, synthetic.seed_for(3, "B"), "normal")

    def step_bf16x3():
        call("coral_sgemm_bf16x3", 0, 0, n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.0, C.ptr, n, None)

    def step_plain():
        call("coral_sgemm", 0, 0, n, n, n, 1.0, A.ptr, n, B.ptr, n, 0.0, C2.ptr, n, None)

    for _ in range(args.warmup):
        step_bf16x3()
    device.synchronize()
    call("coral_timing_reset")
    call("coral_timing_enable", 1)
    ms = timed_steps(step_bf16x3, args.steps, 0)
    k_ms, k_cnt = kernel_avg_ms(0)
    call("coral_timing_enable", 0)
    call("coral_timing_reset")
    flops = 2.0 * n ** 3
    value = flops / ms / 1e9
    log(f"bf16x3 sgemm {n}^3: {ms:.3f} ms/step -> {value:.1f} TFLOP/s; MFMA kernel {k_ms:.3f} ms x {k_cnt}")

    plain_steps = max(3, args.steps // 2)
    ms_plain = timed_steps(step_plain, plain_steps, 1)
    plain = flops / ms_plain / 1e9
    log(f"plain fp32 sgemm {n}^3: {ms_plain:.3f} ms -> {plain:.1f} TFLOP/s")

    # errors against FP64 on a row sample
    Ah = A.to_numpy().reshape(n, n).T
    Bh = B.to_numpy().reshape(n, n).T
    Cb = C.to_numpy().reshape(n, n).T
    Cp = C2.to_numpy().reshape(n, n).T
    rows = np.random.default_rng(1).choice(n, 256, replace=False)
    err_b = sample_error(Ah, Bh, Cb, rows, n)
    err_p = sample_error(Ah, Bh, Cp, rows, n)
    log(f"error bf16x3 {err_b}  plain {err_p}")
    del Ah, Bh, Cb, Cp, A, B, C, C2

    # DGEMV N
    gm = args.gemv_n
    GA = device.DeviceArray(gm * gm, np.float64)
    gx = device.DeviceArray(gm, np.float64)
    gy = device.DeviceArray(gm, np.float64)
    device.fill_random(GA, synthetic.seed_for(2, "A"), "uniform")
    device.fill_random(gx, synthetic.seed_for(2, "x"), "uniform")
    device.fill_random(gy, synthetic.seed_for(2, "y"), "uniform")

    def step_gemv():
        call("coral_dgemv", 0, gm, gm, 1.0, GA.ptr, gm, gx.ptr, 1, 0.0, gy.ptr, 1, None)

    step_gemv()
    device.synchronize()
    call("coral_timing_enable", 1)
    gms = timed_steps(step_gemv, 20, 2)
    gk_ms, _ = kernel_avg_ms(1)
    call("coral_timing_enable", 0)
    call("coral_timing_reset")
    gbytes = 8.0 * (gm * gm + gm + gm)
    gemv_gbps = gbytes / gms / 1e6
    gemv_kernel_gbps = 8.0 * (gm * gm + gm) / gk_ms / 1e6
    log(f"dgemv N {gm}^2: {gms * 1e3:.1f} us/call -> {gemv_gbps:.0f} GB/s (stream kernel {gemv_kernel_gbps:.0f} GB/s)")
    del GA, gx, gy

    achieved = flops / k_ms / 1e9
    peak = PEAK_BF16 / 3.0
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 via bf16x3 (bf16 MFMA, f32 accumulate)",
        "data": "synthetic",
        "config": {"workload": f"SGEMM NN m=n=k={n}, f32 N(0,1) (seeded counter RNG), BF16x3 split form; plain "
                               f"FP32-MFMA SGEMM on the same operands and DGEMV N {gm}^2 (f64 U[-1,1], alpha=1, "
                               f"beta=0) timed beside it",
                   "m": n, "n": n, "k": n, "op": "NN", "gemv_m": gm, "gemv_n": gm},
        "plain_fp32_tflops": round(plain, 2),
        "plain_fp32_frac_of_fp32_peak": round(plain / PEAK_FP32, 4),
        "bf16x3_over_plain": round(value / plain, 3),
        "error_vs_fp64": {"bf16x3": err_b, "plain_fp32": err_p},
        "dgemv": {"gbps": round(gemv_gbps, 1), "kernel_gbps": round(gemv_kernel_gbps, 1),
                  "frac_of_hbm": round(gemv_kernel_gbps / PEAK_HBM, 4), "us_per_call": round(gms * 1e3, 2)},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": round(peak, 2), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": "gemm_kernel<TBF3> (BF16x3 MFMA); work = 2n^3 FP32-equivalent flops per launch; "
                               "peak = 2500 TFLOP/s BF16 / 3 products",
                     "kernel_ms": round(k_ms, 4)},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (CORAL SGEMM restatement, 1 core) ...")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
