"""Throughput of the ViSQOL scoring hot path on one MI355X.

One step = one batch of `--batch` 10 s, 48 kHz audio pairs (seeds 0..B-1,
resident in HBM before timing) scored by visqol_hip_measure_batch_device
(the C call, reading each result's MOS and freeing the results; the Python
dict conversion of the full results is not timed) in the throughput mode (precision "fast": f64 with fused multiply-adds,
|dMOS| <= 1e-6 against the f64 oracle, checked on a sample every run).
Prints one JSON line: whole-job pairs/s, the gammatone kernel's roofline
fraction from HIP events on the library's stream, the plain per-pair exact
f64 path's throughput, and the CPU f64 oracle timed on the host's cores on
a bounded sample of the same pairs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "10s 48kHz pairs/sec at |ΔMOS|≤1e-6 vs f64; p50/p99 latency at fixed QPS"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, FMA counted as 2 flop


def _gen(seed: int):
    from akvisqol_amd import synth

    return synth.audio_pair(seed, duration=10.0, snr_db=25.0, delay=37)


def make_pairs(n: int, workers: int):
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_gen, range(n), chunksize=4))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--exact-pairs", type=int, default=8, help="pairs timed on the per-pair exact f64 path")
    ap.add_argument("--cpu-sample", type=int, default=0, help="oracle sample size (0: 2 per host thread, <= batch)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (pairs would shard "
              "across GPUs with no exchange between them)", file=sys.stderr)
        return 2

    from akvisqol_amd import native

    workers = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    pairs = make_pairs(args.batch, workers)
    print(f"bench: generated {len(pairs)} pairs in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    refs = [native.DeviceBuffer(r) for r, _ in pairs]
    degs = [native.DeviceBuffer(d) for _, d in pairs]

    ctx = native.Context(native.MODE_AUDIO, precision="fast")
    for _ in range(args.warmup):
        rc, res, st = native.measure_batch_device(ctx, refs, degs, 48000)
        if rc != 0:
            raise SystemExit(f"warmup failed: {rc} {st[:8]}")
    ctx.profile(True, True)
    times = []
    for _ in range(args.steps):
        t = time.perf_counter()
        rc, fast_mos, st = native.measure_batch_device(ctx, refs, degs, 48000, convert=False)
        times.append(time.perf_counter() - t)
        if rc != 0:
            raise SystemExit(f"step failed: {rc}")
        print(f"bench: step {len(times)} {times[-1] * 1e3:.1f} ms", file=sys.stderr, flush=True)
    prof = ctx.profile_read()
    phases = {k: round(v / args.steps, 3) for k, v in ctx.profile_phases().items()}
    print(f"bench: host wall ms per step by stretch {phases}", file=sys.stderr, flush=True)
    simp = ctx.profile_sim()
    if simp["launches"]:
        print(f"bench: similarity table {simp['ms'] / simp['launches']:.3f} ms/launch, "
              f"{simp['cells'] / simp['launches']:.4g} cells/launch", file=sys.stderr, flush=True)
    ctx.profile(False, True)
    step_s = float(np.mean(times))
    value = args.batch / step_s

    # Plain per-pair f64 path (exact precision, one pair per call).
    ex = native.Context(native.MODE_AUDIO, precision="exact")
    ne = min(args.exact_pairs, args.batch)
    ex.measure(pairs[0][0], pairs[0][1], 48000)
    t = time.perf_counter()
    exact_mos = [ex.measure(pairs[i][0], pairs[i][1], 48000)["moslqo"] for i in range(ne)]
    exact_pps = ne / (time.perf_counter() - t)

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded 10 s 48 kHz music-like references (8 AM sinusoids + pink noise), degraded by "
                "a 37-sample delay, 25 dB white noise and 16-bit quantisation",
        "config": {"workload": f"{args.batch} x 10 s 48 kHz audio pairs per step, one batched call, throughput "
                               "mode (f64 with FMA), signals resident in HBM",
                   "batch": args.batch, "pair_seconds": 10.0, "sample_rate": 48000, "bands": 32,
                   "precision": "fast"},
        "audio_seconds_per_s": round(value * 10.0, 2),
        "exact_per_pair_pairs_per_s": round(exact_pps, 3),
        "speedup_vs_exact_per_pair": round(value / exact_pps, 2),
        "max_abs_dmos_fast_vs_exact_gpu": float(np.max(np.abs(fast_mos[:ne] - np.array(exact_mos)))),
    }
    if prof["launches"] and prof["ms"] > 0:
        ms_per_launch = prof["ms"] / prof["launches"]
        achieved = prof["flop"] / (prof["ms"] * 1e-3) / 1e12
        line["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
                            "kernel": "k_gammatone<fast>", "ms_per_launch": round(ms_per_launch, 3),
                            "flop_per_launch": prof["flop"] / prof["launches"],
                            "algorithmic_bytes_per_launch": prof["bytes"] / prof["launches"]}

    if not args.no_cpu:
        from oracle import vo

        threads = os.cpu_count() or 1
        ns = args.cpu_sample or min(args.batch, max(4, 2 * min(threads, 64)))
        orc = vo.Oracle(mode=0)
        t = time.perf_counter()
        rc, ores, _ = orc.measure_batch([p[0] for p in pairs[:ns]], [p[1] for p in pairs[:ns]], 48000,
                                        num_threads=threads)
        cpu_s = time.perf_counter() - t
        omos = np.array([r["moslqo"] for r in ores])
        dmos = float(np.max(np.abs(omos - fast_mos[:ns])))
        emos = float(np.max(np.abs(omos[:ne] - np.array(exact_mos[:ns]))))
        line["cpu_baseline"] = {"value": round(ns / cpu_s, 3), "unit": "pairs/s", "cores": threads, "kind": "port",
                                "sample": f"first {ns} pairs of the batch, C f64 oracle, one pair per thread"}
        line["max_abs_dmos_vs_oracle"] = dmos
        line["max_abs_dmos_exact_vs_oracle"] = emos
        line["dmos_ok"] = dmos <= 1e-6
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
