#!/usr/bin/env python3
"""Benchmark of the SGEMM hot path on one MI355X.

Workload: square SGEMM, M = N = K = 8192, row-major NN, alpha = 1.5,
beta = 0.5, A, B, C ~ U[-1, 1] (seeded), inputs resident in HBM.  A "step" is
one GEMM call of the headline form (split-BF16 on MFMA); the plain FP32 form
and the other forms are timed the same way and reported beside it.

Prints ONE JSON line (rank 0).  Timing: W untimed warm-up calls, then K calls
bracketed by device synchronisation, measured with HIP events recorded on the
stream the kernels run on.  Errors: max / mean relative error against an FP64
evaluation of 64 sampled rows.  cpu_baseline: the reference's own AVX SGEMM
(oracle/_ref, compiled from its sources) on a bounded slice of the same
workload, one thread, as the reference runs.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import ugemm_amd as ug  # noqa: E402
from ugemm_amd import hip  # noqa: E402

PEAK = {"fp32_vector": 157.3, "bf16_mfma": 2516.6, "fp8_mfma": 5033.2, "fp64": 78.6}  # TFLOP/s, MI355X
MODE_PEAK = {"fp32": ("fp32_vector", 1), "bf16x3": ("bf16_mfma", 3), "bf16x6": ("bf16_mfma", 6),
             "bf16x1": ("bf16_mfma", 1), "fp8": ("fp8_mfma", 1)}


def gen_inputs(n, seed):
    g = np.random.Generator(np.random.PCG64(seed))

    def u(shape):
        return (g.integers(0, 1 << 24, size=shape, dtype=np.int64).astype(np.float32) * np.float32(2.0 ** -23)
                - np.float32(1.0))
    return u((n, n)), u((n, n)), u((n, n))


def time_mode(mode_id, n, dA, dB, dC, alpha, beta, steps, warmup, stream):
    for _ in range(warmup):
        ug.sgemm_dev("R", "N", "N", n, n, n, alpha, dA, n, dB, n, beta, dC, n, mode=mode_id, stream=stream)
    stream.synchronize()
    hip.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        ug.sgemm_dev("R", "N", "N", n, n, n, alpha, dA, n, dB, n, beta, dC, n, mode=mode_id, stream=stream)
    e1.record(stream)
    e1.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1) / steps
    e0.destroy()
    e1.destroy()
    return ms, wall


def errors(mode_id, n, A, B, C0, dA, dB, dC, alpha, beta, stream, rows):
    dC.upload(C0)
    ug.sgemm_dev("R", "N", "N", n, n, n, alpha, dA, n, dB, n, beta, dC, n, mode=mode_id, stream=stream)
    stream.synchronize()
    C = dC.download()
    ref = alpha * (A[rows].astype(np.float64) @ B.astype(np.float64)) + beta * C0[rows].astype(np.float64)
    got = C[rows].astype(np.float64)
    nz = ref != 0
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    absmax = float(np.abs(got - ref).max())
    return float(rel.max()), float(rel.mean()), absmax


def cpu_baseline(n, seconds_target=10.0):
    """The reference's sgemm_avx (sgemm_avx256.h:392) on a row band of the
    same workload: M = band rows of C, N = K = n, one thread."""
    try:
        from oracle import oracle as orc
    except Exception:
        return None
    L = orc.ref_lib()
    kind = "reference"
    fn = None
    if L is not None:
        fn = L.sgemm_avx
    else:
        kind = "port"
        fn = orc.lib().oracle_sgemm_packed
    band = 256
    g = np.random.Generator(np.random.PCG64(7))
    A = (g.random((band, n), dtype=np.float32) * 2 - 1).astype(np.float32)
    B = (g.random((n, n), dtype=np.float32) * 2 - 1).astype(np.float32)
    C = (g.random((band, n), dtype=np.float32) * 2 - 1).astype(np.float32)
    calls, t_total = 0, 0.0
    while t_total < seconds_target and calls < 64:
        t0 = time.perf_counter()
        fn(b"R", b"N", b"N", band, n, n, 1.5, A.ctypes.data, n, B.ctypes.data, n, 0.5, C.ctypes.data, n)
        t_total += time.perf_counter() - t0
        calls += 1
    tflops = 2.0 * band * n * n * calls / t_total / 1e12
    cpu = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    fname = "sgemm_avx (sgemm_avx256.h:392)" if kind == "reference" else "oracle_sgemm_packed (gemm_cpu.h blocking)"
    return {"value": round(tflops, 6), "unit": "TFLOP/s", "cores": 1, "kind": kind,
            "sample": f"{fname}, R/N/N, M={band} rows x N=K={n} (a row band of the {n}^3 workload), "
                      f"alpha=1.5 beta=0.5, {calls} calls in {t_total:.1f} s, 1 thread on {cpu}, "
                      f"nproc={os.cpu_count()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--modes", default="fp32,bf16x3,bf16x6,bf16x1,fp8")
    ap.add_argument("--headline", default="bf16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build is a single-GPU GEMM backend); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    n = args.size
    alpha, beta = 1.5, 0.5
    ug.init(0, 0)
    A, B, C0 = gen_inputs(n, 0x55474D4D02)
    dA, dB, dC = hip.DeviceArray.from_host(A), hip.DeviceArray.from_host(B), hip.DeviceArray.from_host(C0)
    stream = hip.Stream()
    rows = np.sort(np.random.Generator(np.random.PCG64(3)).choice(n, 64, replace=False))
    flop = 2.0 * n * n * n

    results = {}
    for name in args.modes.split(","):
        mid = ug.MODES[name]
        try:
            ms, wall = time_mode(mid, n, dA, dB, dC, alpha, beta, args.steps, args.warmup, stream)
        except ug.UgemmError as e:
            results[name] = {"error": str(e)}
            continue
        mx, mean, absmax = errors(mid, n, A, B, C0, dA, dB, dC, alpha, beta, stream, rows)
        tf = flop / (ms * 1e-3) / 1e12
        peak_key, mult = MODE_PEAK[name]
        results[name] = {"tflops": round(tf, 3), "ms": round(ms, 4), "wall_s": round(wall, 3),
                         "max_rel_err": mx, "mean_rel_err": mean, "max_abs_err": absmax,
                         "frac_fp32_vector": round(tf / PEAK["fp32_vector"], 4),
                         "frac_native_peak": round(tf * mult / PEAK[peak_key], 4), "native_peak": peak_key}

    head = args.headline if args.headline in results and "tflops" in results[args.headline] else "fp32"
    h = results[head]
    peak_key, mult = MODE_PEAK[head]
    roof = {"bound": "mfma" if "mfma" in peak_key else "fp32_vector" if False else "mfma" if head != "fp32" else "mfma",
            "achieved": h["tflops"], "peak": PEAK[peak_key] if head != "fp32" else PEAK["fp32_vector"],
            "unit": "TFLOP/s"}
    if head == "fp32":
        roof["bound"] = "mfma"  # compute-bound; peak is the FP32 vector rate
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    roof["traffic"] = None
    if head != "fp32":
        roof["mfma_utilisation"] = round(h["tflops"] * mult / PEAK[peak_key], 4)
    line = {
        "metric": "SGEMM TFLOP/s at 8192^3: plain FP32 vs BF16x3-MFMA at FP32-level error",
        "value": h["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": h["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (bf16 MFMA, split)" if head.startswith("bf16") else "f32",
        "data": "synthetic U[-1,1], PCG64 seed 0x55474D4D02",
        "config": {"workload": f"sgemm {n}x{n}x{n} row-major NN alpha=1.5 beta=0.5, headline form {head}",
                   "M": n, "N": n, "K": n, "layout": "NN", "form": head},
        "forms": results,
        "roofline": roof,
    }
    if "fp32" in results and "tflops" in results["fp32"] and head != "fp32":
        line["speedup_vs_plain_fp32"] = round(h["tflops"] / results["fp32"]["tflops"], 3)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
