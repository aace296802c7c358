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
import threading
import time
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "10s 48kHz pairs/sec at |ΔMOS|≤1e-6 vs f64; p50/p99 latency at fixed QPS"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, FMA counted as 2 flop
# f64 VALU instructions per (sample, band) of the throughput-mode
# spectrogram at 32 bands: 8 bands in the DF2T cascade (17 each), 24 in the
# FIR + all-pole form (14 each)
FAST_OPS_PER_SAMPLE_BAND = (8 * 17 + 24 * 14) / 32


def _gen(seed: int):
    from akvisqol_amd import synth

    return synth.audio_pair(seed, duration=10.0, snr_db=25.0, delay=37)


def make_pairs(n: int, workers: int):
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_gen, range(n), chunksize=4))


def _traffic_ratio(path: Path, shape: str):
    """`traffic_per_algorithmic_byte <shape> = X` from a counter summary, or None."""
    try:
        for ln in path.read_text().splitlines():
            if ln.startswith(f"traffic_per_algorithmic_byte {shape} ="):
                return float(ln.split("=")[1])
    except OSError:
        return None
    return None


def _gen_mixed(seed: int):
    from akvisqol_amd import synth

    return synth.mixed_pair(seed, max_duration=120.0)


def _gen_mixed60(seed: int):
    from akvisqol_amd import synth

    return synth.mixed_pair(50_000 + seed, max_duration=60.0)


def serve_level(srv, pool, qps: float, seconds: float, seed: int, submitters: int = 4):
    """Poisson arrivals at `qps` for `seconds`, each request a pool pair drawn
    uniformly; latency is the server's submit-to-completion time."""
    rng = np.random.default_rng(seed)
    n = max(1, int(round(qps * seconds)))
    times = np.cumsum(rng.exponential(1.0 / qps, n))
    picks = rng.integers(0, len(pool), n)
    tickets = [None] * n
    submitted = [threading.Event() for _ in range(n)]
    start = time.perf_counter() + 0.05

    def submitter(k):
        for i in range(k, n, submitters):
            delay = start + times[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ref, deg, sr, mode = pool[picks[i]]
            tickets[i] = srv.submit(mode, ref, deg, sr)
            submitted[i].set()

    threads = [threading.Thread(target=submitter, args=(k,)) for k in range(submitters)]
    for t in threads:
        t.start()
    lat = np.zeros(n)
    mos = np.full(n, np.nan)
    for i in range(n):
        submitted[i].wait()
        rc, m, lat[i] = srv.wait(tickets[i], convert=False)
        if rc == 0:
            mos[i] = m
    end = time.perf_counter()
    for t in threads:
        t.join()
    return dict(offered_qps=qps, seconds=seconds, requests=n, pairs_per_s=round(n / (end - start), 2),
                p50_ms=round(float(np.percentile(lat, 50)), 2), p99_ms=round(float(np.percentile(lat, 99)), 2),
                failed=int(np.sum(np.isnan(mos)))), picks, mos


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--exact-pairs", type=int, default=8, help="pairs timed on the per-pair exact f64 path")
    ap.add_argument("--cpu-sample", type=int, default=0, help="oracle sample size (0: 2 per host thread, <= batch)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--qps", default="200,500,1000", help="serving loads (requests/s); empty: no serving run")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--uneven", type=int, default=512, help="pairs of the 1-60 s mix scored as batches (0: skip)")
    ap.add_argument("--pool", type=int, default=64, help="distinct pairs of the 1-120 s mix drawn by requests")
    ap.add_argument("--grad-pairs", type=int, default=64,
                    help="pairs of the batch scored with dMOS/ddeg (visqol_hip_measure_batch_grad); 0: skip")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer three times (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        src = torch.ones(args.copy_probe << 17, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench: copied {args.copy_probe} MiB three times", file=sys.stderr, flush=True)
        return 0
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
    # spectrogram block shapes timed by the engine on its first large launch
    autotune = {}
    for kind, name in ((0, "spectrogram"), (1, "realignment_spectrogram")):
        t = ctx.gammatone_tuning(48000, kind)
        if t["chosen"] >= 0:
            autotune[name] = {"chosen": t["shapes"][t["chosen"]],
                              "ms_per_shape": {f"{sh['waves']}w{sh['bands_per_lane']}b{sh['chunk']}c": m
                                               for sh, m in zip(t["shapes"], t["ms"])}}
    print(f"bench: autotune {autotune}", file=sys.stderr, flush=True)
    value = args.batch / step_s

    # Gradients (visqol_hip_measure_batch_grad) of the first pairs, host
    # arrays in and out: the batch scored alone, then with dMOS/ddeg, each
    # timed after a warm call of the same size.
    grad = None
    ng = min(args.grad_pairs, args.batch)
    if ng > 0:
        gr = [p[0] for p in pairs[:ng]]
        gd = [p[1] for p in pairs[:ng]]
        ctx.measure_batch(gr, gd, 48000, convert=False)
        ctx.measure_batch_grad(gr, gd, 48000)
        t = time.perf_counter()
        ctx.measure_batch(gr, gd, 48000, convert=False)
        t_score = time.perf_counter() - t
        t = time.perf_counter()
        rc_g, _, _, g = ctx.measure_batch_grad(gr, gd, 48000)
        t_grad = time.perf_counter() - t
        grad = {"pairs": ng, "score_ms": round(t_score * 1e3, 2), "score_and_grad_ms": round(t_grad * 1e3, 2),
                "grad_ms_per_pair": round((t_grad - t_score) * 1e3 / ng, 3), "rc": rc_g,
                "grad_l2_median": round(float(np.median([np.linalg.norm(x) for x in g])), 4)}
        print(f"bench: gradients {grad}", file=sys.stderr, flush=True)

    # The same batch handed over as host arrays (visqol_measure_batch): the
    # PCIe-inclusive rate of the plain C entry point, the median of three
    # steps after one untimed step (which also allocates the pinned staging).
    host_refs, host_degs = [p[0] for p in pairs], [p[1] for p in pairs]
    ctx.measure_batch(host_refs, host_degs, 48000, convert=False)
    host_ts, rc_h = [], 0
    for _ in range(3):
        t = time.perf_counter()
        rc_h, _, _ = ctx.measure_batch(host_refs, host_degs, 48000, convert=False)
        host_ts.append(time.perf_counter() - t)
        if rc_h != 0:
            break
    host_pps = args.batch / float(np.median(host_ts)) if rc_h == 0 else None

    # Plain per-pair f64 path (exact precision, one pair per call).
    ex = native.Context(native.MODE_AUDIO, precision="exact")
    ne = min(args.exact_pairs, args.batch)
    ex.measure(pairs[0][0], pairs[0][1], 48000)
    t = time.perf_counter()
    exact_mos = [ex.measure(pairs[i][0], pairs[i][1], 48000)["moslqo"] for i in range(ne)]
    exact_pps = ne / (time.perf_counter() - t) if ne > 0 else None

    # Per-stage error of the throughput mode against f64 (the exact path,
    # bit-identical to the oracle) on the first pairs of the batch.
    fctx = native.Context(native.MODE_AUDIO, precision="fast")
    err = {"spectrogram_max_rel": 0.0, "similarity_table_max_abs": 0.0, "fvnsim_max_abs": 0.0,
           "vnsim_max_abs": 0.0, "moslqo_max_abs": 0.0, "same_alignments": True}
    n_stage = min(4, args.batch)
    for i in range(n_stage):
        er, et = ex.measure_trace(pairs[i][0], pairs[i][1], 48000, want_stages=True)
        fr, ft = fctx.measure_trace(pairs[i][0], pairs[i][1], 48000, want_stages=True)
        for k in ("ref_spec", "deg_spec"):
            rel = np.abs(ft[k] - et[k]) / np.maximum(np.abs(et[k]), 1e-300)
            err["spectrogram_max_rel"] = max(err["spectrogram_max_rel"], float(np.max(rel)))
        if ft["sims"].shape == et["sims"].shape:
            err["similarity_table_max_abs"] = max(err["similarity_table_max_abs"],
                                                  float(np.max(np.abs(ft["sims"] - et["sims"]))))
        err["fvnsim_max_abs"] = max(err["fvnsim_max_abs"], float(np.max(np.abs(fr["fvnsim"] - er["fvnsim"]))))
        err["vnsim_max_abs"] = max(err["vnsim_max_abs"], abs(fr["vnsim"] - er["vnsim"]))
        err["moslqo_max_abs"] = max(err["moslqo_max_abs"], abs(fr["moslqo"] - er["moslqo"]))
        for k in ("lag", "dp_offset", "realign_lag", "realign_accepted"):
            err["same_alignments"] &= ft[k] == et[k]
    err["pairs"] = n_stage
    print(f"bench: fast-mode stage error vs f64 {err}", file=sys.stderr, flush=True)

    # Uneven batches: pairs of the 1-60 s mix (70% 48 kHz audio, 30% 16 kHz
    # speech, per-pair noise, delay, stretch, dropouts), device-resident,
    # scored as one batch per mode.
    uneven = None
    if args.uneven:
        t = time.perf_counter()
        with ProcessPoolExecutor(max_workers=workers) as ex:
            mix = list(ex.map(_gen_mixed60, range(args.uneven)))
        print(f"bench: generated {len(mix)} uneven pairs in {time.perf_counter() - t:.1f}s", file=sys.stderr,
              flush=True)
        groups = {}
        for ref, deg, sr, mode in mix:
            g = groups.setdefault((mode, sr), ([], []))
            g[0].append(native.DeviceBuffer(ref))
            g[1].append(native.DeviceBuffer(deg))
        uctx = {m: native.Context(m, precision="fast") for m in (0, 1)}
        for (mode, sr), (rb, db) in groups.items():  # warm the engines
            native.measure_batch_device(uctx[mode], rb, db, sr, convert=False)
        ts = []
        failed = 0
        for _ in range(2):
            t = time.perf_counter()
            failed = 0
            for (mode, sr), (rb, db) in groups.items():
                _, umos, _ = native.measure_batch_device(uctx[mode], rb, db, sr, convert=False)
                failed += int(np.sum(np.isnan(umos)))
            ts.append(time.perf_counter() - t)
        secs = float(sum(p[0].size / p[2] for p in mix))
        step = float(np.mean(ts))
        uneven = {"workload": f"{len(mix)} pairs of the uneven mix (durations log-uniform 1-60 s, "
                              f"{sum(1 for p in mix if p[3] == 1)} speech), one device-resident batch per mode",
                  "pairs_per_s": round(len(mix) / step, 2), "audio_seconds_per_s": round(secs / step, 1),
                  "ms_per_batch": round(step * 1e3, 2), "failed": failed}
        for g in groups.values():
            for b in g[0] + g[1]:
                b.free()
        print(f"bench: uneven {uneven}", file=sys.stderr, flush=True)

    serving = None
    qps_levels = [float(q) for q in args.qps.split(",") if q.strip()]
    if qps_levels:
        # Serving: Poisson arrivals of pairs from the uneven mix (70% 48 kHz
        # audio, 30% 16 kHz speech, 1-120 s log-uniform, noise, delay,
        # stretch, dropouts), continuous batching on one device.
        t = time.perf_counter()
        with ProcessPoolExecutor(max_workers=workers) as ex:
            pool = list(ex.map(_gen_mixed, range(args.pool)))
        secs = float(sum(p[0].size / p[2] for p in pool))
        print(f"bench: serving pool of {len(pool)} pairs ({secs / len(pool):.1f} s mean) in "
              f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
        levels = []
        served_mos = {}
        with native.Server(precision="fast") as srv:
            serve_level(srv, pool, 100.0, 1.0, seed=99)  # warm the workers' engines
            for k, q in enumerate(qps_levels):
                rep, picks, mos = serve_level(srv, pool, q, args.serve_seconds, seed=1000 + k)
                levels.append(rep)
                for pi, m in zip(picks, mos):
                    served_mos.setdefault(int(pi), m)
                print(f"bench: serving {rep}", file=sys.stderr, flush=True)
            stats = srv.stats()
        # the plain f64 per-pair path: service time per request, one at a time
        exc = {0: native.Context(0, precision="exact"), 1: native.Context(1, precision="exact")}
        svc = []
        for ref, deg, sr, mode in pool[:8]:
            t = time.perf_counter()
            exc[mode].measure(ref, deg, sr)
            svc.append((time.perf_counter() - t) * 1e3)
        serving = {"workload": f"Poisson arrivals, requests drawn from {len(pool)} pairs of the uneven mix "
                               "(70% 48 kHz audio / 30% 16 kHz speech, durations log-uniform 1-120 s, "
                               f"mean {secs / len(pool):.1f} s)",
                   "precision": "fast", "levels": levels,
                   "batches": stats["batches"], "pairs": stats["pairs"],
                   "exact_per_pair": {"p50_ms": round(float(np.percentile(svc, 50)), 2),
                                      "p99_ms": round(float(np.percentile(svc, 99)), 2),
                                      "capacity_pairs_per_s": round(1e3 * len(svc) / float(np.sum(svc)), 2),
                                      "note": "service time of one request on an idle device (first 8 pool pairs); "
                                              "offered loads above capacity queue without bound"}}

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
        "host_buffers_pairs_per_s": round(host_pps, 3) if host_pps else None,
        "exact_per_pair_pairs_per_s": round(exact_pps, 3) if exact_pps else None,
        "speedup_vs_exact_per_pair": round(value / exact_pps, 2) if exact_pps else None,
        "max_abs_dmos_fast_vs_exact_gpu": (float(np.max(np.abs(fast_mos[:ne] - np.array(exact_mos)))) if ne > 0
                                           else None),
        "stage_error_fast_vs_f64": err,
        "autotune": autotune,
        "gradients": grad,
    }
    if prof["launches"] and prof["ms"] > 0:
        ms_per_launch = prof["ms"] / prof["launches"]
        achieved = prof["flop"] / (prof["ms"] * 1e-3) / 1e12
        alg_bytes = prof["bytes"] / prof["launches"]
        # HBM bytes per algorithmic byte of this kernel on this workload for
        # the block shape the engine chose, from the committed FETCH_SIZE /
        # WRITE_SIZE counter runs (copy-calibrated)
        traffic = None
        ratio = None
        if "spectrogram" in autotune:
            sh = autotune["spectrogram"]["chosen"]
            ratio = _traffic_ratio(ROOT / "profiles" / "k_gammatone_fast_1024x10s48k_hbm_traffic.txt",
                                   f"{sh['waves']}w{sh['bands_per_lane']}b{sh['chunk']}c")
        if ratio is not None and args.batch == 1024:
            traffic = round(ratio * alg_bytes)
        line["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                            "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                            "kernel": "k_gammatone<fast>", "ms_per_launch": round(ms_per_launch, 3),
                            "flop_per_launch": prof["flop"] / prof["launches"],
                            "algorithmic_bytes_per_launch": alg_bytes,
                            # 30 flop per (sample, band) issue as 13 FMA + 4 mul: the FP64
                            # instruction-issue ceiling is peak / 2 lane-ops per second
                            "issue_frac": round(achieved * (FAST_OPS_PER_SAMPLE_BAND / 30.0) / (FP64_PEAK_TFLOPS / 2.0), 4)}

    if uneven is not None:
        line["uneven_batch"] = uneven
    if serving is not None:
        line["serving"] = serving
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
        emos = float(np.max(np.abs(omos[:ne] - np.array(exact_mos[:ns])))) if min(ne, ns) > 0 else None
        line["cpu_baseline"] = {"value": round(ns / cpu_s, 3), "unit": "pairs/s", "cores": threads, "kind": "port",
                                "sample": f"first {ns} pairs of the batch, C f64 oracle, one pair per thread"}
        line["max_abs_dmos_vs_oracle"] = dmos
        if serving is not None and served_mos:
            # served results are batch-independent: check 4 served pool pairs
            # at length quantiles 0, 25, 50 and 75 %
            by_len = sorted((k for k in served_mos if not np.isnan(served_mos[k])), key=lambda k: pool[k][0].size)
            keys = sorted({by_len[int(f * (len(by_len) - 1))] for f in (0.0, 0.25, 0.5, 0.75)}) if by_len else []
            sd = []
            for mode in (0, 1):
                sel = [k for k in keys if pool[k][3] == mode]
                if not sel:
                    continue
                rc2, ores2, _ = vo.Oracle(mode=mode).measure_batch([pool[k][0] for k in sel], [pool[k][1] for k in sel],
                                                                     pool[sel[0]][2], num_threads=threads)
                sd += [abs(o["moslqo"] - served_mos[k]) for k, o in zip(sel, ores2) if o is not None]
            serving["max_abs_dmos_vs_oracle"] = float(max(sd)) if sd else None
            serving["oracle_checked_pairs"] = len(sd)
        line["max_abs_dmos_exact_vs_oracle"] = emos
        line["dmos_ok"] = dmos <= 1e-6
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
