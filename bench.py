"""Throughput and serving latency of the ViSQOL scoring hot path on one MI355X.

Headline (the JSON line's `value`): one step = one batch of `--batch` 10 s,
48 kHz audio pairs (seeds 0..B-1, resident in HBM before timing) scored by
visqol_hip_measure_batch_device (the C call, reading each result's MOS and
freeing the results) in the throughput mode (precision "fast": f64 with
fused multiply-adds, lags nominated from an f32 correlation and rescored in
f64; |dMOS| <= 1e-6 against the f64 oracle, checked every run).  Beside it:
the gammatone kernel's roofline fraction from HIP events on the library's
stream; the same batch as host arrays and in the exact f64 mode (the path
the reference's default backend is served with); the plain per-pair exact
f64 path;
per-stage fast-vs-f64 error for audio and speech; uneven batches of the
1-60 s mix; the continuous-batching server under Poisson arrivals of the
1-120 s mix (p50/p99 latency and where the tail's time went), with |dMOS|
against the oracle on a 1 % sample of served requests; and the CPU f64
oracle timed on the host's cores (best of two rounds) plus its one-thread
latency on the 10 s pair.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor
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
HOST_THREADS = 16  # host cores a process on one GPU has


def _gen(seed: int):
    from akvisqol_amd import synth

    return synth.audio_pair(seed, duration=10.0, snr_db=25.0, delay=37)


def make_pairs(n: int, workers: int):
    with ProcessPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(_gen, range(n), chunksize=4))


def make_mix(n_req: int, n_src: int, max_s: float, seed: int, threads: int):
    """n_req pairs of the uneven mix (70 % 48 kHz audio / 30 % 16 kHz
    speech): slices of log-uniform duration 1..max_s s of n_src long source
    pairs (delay 0-4800 samples, speech low-pass, 0-3 % stretch), each with
    its own noise (5-40 dB SNR), dropouts and quantisation."""
    from akvisqol_amd import synth

    with ThreadPoolExecutor(max_workers=threads) as ex:
        srcs = list(ex.map(lambda k: synth.mix_source(seed * 1000 + k, max_s), range(n_src)))
        return list(ex.map(lambda k: synth.mix_request(srcs[k % n_src], seed * 100_000 + k, 1.0, max_s),
                           range(n_req)))


def _traffic_ratio(path: Path, shape: str):
    """`traffic_per_algorithmic_byte <shape> = X` from a counter summary, or None."""
    try:
        for ln in path.read_text().splitlines():
            if ln.startswith(f"traffic_per_algorithmic_byte {shape} ="):
                return float(ln.split("=")[1])
    except OSError:
        return None
    return None


def _plain(x):
    """numpy scalars in the JSON line as plain numbers and booleans."""
    if isinstance(x, np.generic):
        return x.item()
    raise TypeError(f"{type(x).__name__} is not JSON serializable")


def say(msg: str) -> None:
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def serve_level(srv, reqs, qps: float, seconds: float, seed: int, submitters: int = 4):
    """Poisson arrivals at `qps` for `seconds`, each request one of `reqs`
    (distinct pairs of the mix, cycled in a seeded random order); latency is
    the server's submit-to-completion time, split by phase."""
    rng = np.random.default_rng(seed)
    n = max(1, int(round(qps * seconds)))
    times = np.cumsum(rng.exponential(1.0 / qps, n))
    order = np.concatenate([rng.permutation(len(reqs)) for _ in range(n // len(reqs) + 1)])[:n]
    tickets = [None] * n
    submitted = [threading.Event() for _ in range(n)]
    start = time.perf_counter() + 0.05

    def submitter(k):
        for i in range(k, n, submitters):
            delay = start + times[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ref, deg, sr, mode = reqs[order[i]]
            tickets[i] = srv.submit(mode, ref, deg, sr)
            submitted[i].set()

    threads = [threading.Thread(target=submitter, args=(k,)) for k in range(submitters)]
    for t in threads:
        t.start()
    lat = np.zeros(n)
    phases = np.zeros((n, 4))
    batch = np.zeros(n)
    mos = np.full(n, np.nan)
    for i in range(n):
        submitted[i].wait()
        rc, m, tm = srv.wait_timed(tickets[i], convert=False)
        lat[i] = tm["total_ms"]
        phases[i] = (tm["admit_ms"], tm["upload_ms"], tm["queue_ms"], tm["run_ms"])
        batch[i] = tm["batch_size"]
        if rc == 0:
            mos[i] = m
    end = time.perf_counter()
    for t in threads:
        t.join()
    p99 = float(np.percentile(lat, 99))
    tail = lat >= p99
    rep = dict(offered_qps=qps, seconds=seconds, requests=n, pairs_per_s=round(n / (end - start), 2),
               p50_ms=round(float(np.percentile(lat, 50)), 2), p99_ms=round(p99, 2),
               max_ms=round(float(lat.max()), 2), failed=int(np.sum(np.isnan(mos))),
               mean_batch=round(float(batch.mean()), 2),
               tail_phase_ms={k: round(float(phases[tail, j].mean()), 2)
                              for j, k in enumerate(("admit", "upload", "queue", "run"))})
    worst = int(np.argmax(lat))
    wref, _, wsr, wmode = reqs[order[worst]]
    rep["max_request"] = dict(seconds=round(wref.size / wsr, 2), mode=int(wmode), batch=int(batch[worst]),
                              phase_ms={k: round(float(phases[worst, j]), 2)
                                        for j, k in enumerate(("admit", "upload", "queue", "run"))})
    return rep, order, mos


def stage_errors(native, pairs, mode, sr, lattice=True):
    """Per-stage error of the throughput mode against f64 (the exact path,
    bit-identical to the oracle) on the given pairs."""
    kw = {} if mode == 0 else {"use_lattice_model": lattice}
    ex = native.Context(mode, precision="exact", **kw)
    fc = native.Context(mode, precision="fast", **kw)
    err = {"spectrogram_max_rel": 0.0, "similarity_table_max_abs": 0.0, "fvnsim_max_abs": 0.0,
           "vnsim_max_abs": 0.0, "moslqo_max_abs": 0.0, "same_alignments": True}
    for ref, deg in pairs:
        er, et = ex.measure_trace(ref, deg, sr, want_stages=True)
        fr, ft = fc.measure_trace(ref, deg, sr, want_stages=True)
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
    err["pairs"] = len(pairs)
    return err


def oracle_dmos(vo, reqs, idx, got_mos, threads):
    """max |dMOS| of got_mos[k] (for reqs[idx[k]]) against the CPU f64 oracle."""
    worst, checked = 0.0, 0
    for mode in (0, 1):
        sel = [k for k, i in enumerate(idx) if reqs[i][3] == mode and not np.isnan(got_mos[k])]
        if not sel:
            continue
        sr = reqs[idx[sel[0]]][2]
        _, ores, _ = vo.Oracle(mode=mode).measure_batch([reqs[idx[k]][0] for k in sel], [reqs[idx[k]][1] for k in sel],
                                                         sr, num_threads=threads)
        for k, o in zip(sel, ores):
            if o is not None:
                worst = max(worst, abs(o["moslqo"] - got_mos[k]))
                checked += 1
    return float(worst), checked


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--exact-pairs", type=int, default=8, help="pairs timed on the per-pair exact f64 path")
    ap.add_argument("--exact-batch", type=int, default=1,
                    help="also time the headline batch in the exact f64 mode (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="oracle pairs per round (0: 2 per host thread)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle checks")
    ap.add_argument("--qps", default="200,500,1000", help="serving loads (requests/s); empty: no serving run")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-requests", type=int, default=512, help="distinct requests of the 1-120 s mix")
    ap.add_argument("--serve-sources", type=int, default=64, help="long source pairs the requests are cut from")
    ap.add_argument("--uneven", type=int, default=4096, help="pairs of the 1-60 s mix scored as batches (0: skip)")
    ap.add_argument("--uneven-sources", type=int, default=128)
    ap.add_argument("--oracle-fraction", type=float, default=0.01, help="served / uneven pairs checked on the oracle")
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
        say(f"copied {args.copy_probe} MiB three times")
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (pairs would shard "
              "across GPUs with no exchange between them)", file=sys.stderr)
        return 2
    t_all = time.perf_counter()
    workers = max(1, min(HOST_THREADS, os.cpu_count() or 1))

    # ------------------------------------------------------------ inputs (host, before any device work)
    t0 = time.perf_counter()
    pairs = make_pairs(args.batch, workers)
    say(f"generated {len(pairs)} ten-second pairs in {time.perf_counter() - t0:.1f}s")
    mix = []
    if args.uneven:
        t0 = time.perf_counter()
        mix = make_mix(args.uneven, args.uneven_sources, 60.0, 3, workers)
        say(f"generated {len(mix)} pairs of the 1-60 s mix ({sum(p[0].size / p[2] for p in mix):.0f} s of audio, "
            f"{sum(p[3] for p in mix)} speech) in {time.perf_counter() - t0:.1f}s")
    qps_levels = [float(q) for q in args.qps.split(",") if q.strip()]
    reqs = []
    if qps_levels:
        t0 = time.perf_counter()
        reqs = make_mix(args.serve_requests, args.serve_sources, 120.0, 4, workers)
        say(f"generated {len(reqs)} serving requests of the 1-120 s mix "
            f"({sum(p[0].size / p[2] for p in reqs) / len(reqs):.1f} s mean) in {time.perf_counter() - t0:.1f}s")

    from akvisqol_amd import native

    # ------------------------------------------------------------ headline
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
        say(f"step {len(times)} {times[-1] * 1e3:.1f} ms")
    prof = ctx.profile_read()
    phases = {k: round(v / args.steps, 3) for k, v in ctx.profile_phases().items()}
    say(f"host wall ms per step by stretch {phases}")
    simp = ctx.profile_sim()
    if simp["launches"]:
        say(f"similarity table {simp['ms'] / simp['launches']:.3f} ms/launch, "
            f"{simp['cells'] / simp['launches']:.4g} cells/launch")
    ctx.profile(False, True)
    step_s = float(np.mean(times))
    autotune = {}
    for kind, name in ((0, "spectrogram"), (1, "realignment_spectrogram")):
        t = ctx.gammatone_tuning(48000, kind)
        if t["chosen"] >= 0:
            autotune[name] = {"chosen": t["shapes"][t["chosen"]],
                              "ms_per_shape": {f"{sh['waves']}w{sh['bands_per_lane']}b{sh['chunk']}c": m
                                               for sh, m in zip(t["shapes"], t["ms"])}}
    say(f"autotune {autotune}")
    value = args.batch / step_s

    # The exact f64 mode (what the reference's default backend, CPU, is
    # served with) on the same device-resident batch: one warm, one timed.
    exact_batch = None
    exact_batch_mos = None
    if args.exact_batch:
        ectx = native.Context(native.MODE_AUDIO, precision="exact")
        native.measure_batch_device(ectx, refs, degs, 48000, convert=False)
        t = time.perf_counter()
        rc_e, exact_batch_mos, _ = native.measure_batch_device(ectx, refs, degs, 48000, convert=False)
        te = time.perf_counter() - t
        exact_batch = {"pairs_per_s": round(args.batch / te, 2), "ms_per_batch": round(te * 1e3, 2), "rc": rc_e}
        say(f"exact batch {exact_batch}")
    for b in refs + degs:
        b.free()

    # Gradients of the first pairs, host arrays in and out.
    grad = None
    ng = min(args.grad_pairs, args.batch)
    if ng > 0:
        gr, gd = [p[0] for p in pairs[:ng]], [p[1] for p in pairs[:ng]]
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
        say(f"gradients {grad}")

    # The same batch as host arrays (visqol_measure_batch): the PCIe-inclusive
    # rate of the plain C entry point (median of three steps after one
    # untimed step); every score must equal the device-resident batch's.
    host_refs, host_degs = [p[0] for p in pairs], [p[1] for p in pairs]
    ctx.measure_batch(host_refs, host_degs, 48000, convert=False)
    host_ts, rc_h, host_match = [], 0, True
    for _ in range(3):
        t = time.perf_counter()
        rc_h, host_mos, _ = ctx.measure_batch(host_refs, host_degs, 48000, convert=False)
        host_ts.append(time.perf_counter() - t)
        host_match &= bool(np.array_equal(host_mos, fast_mos))
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

    # Per-stage error of the throughput mode against f64: the first audio
    # pairs of the batch, and 16 kHz speech pairs with the lattice mapping.
    from akvisqol_amd import synth

    err = {"audio": stage_errors(native, pairs[:4], 0, 48000),
           "speech_lattice": stage_errors(native, [synth.speech_pair(s, snr_db=10.0 + 5 * s, delay=17 * s)
                                                  for s in range(4)], 1, 16000)}
    say(f"fast-mode stage error vs f64 {err}")

    # Uneven batches of the 1-60 s mix, device-resident, one batch per mode.
    uneven = None
    uneven_mos = None
    if mix:
        groups = {}
        for i, (ref, deg, sr, mode) in enumerate(mix):
            g = groups.setdefault((mode, sr), ([], [], []))
            g[0].append(native.DeviceBuffer(ref))
            g[1].append(native.DeviceBuffer(deg))
            g[2].append(i)
        uctx = [native.Context(m, precision="fast") for m in (0, 1)]
        all_refs = [None] * len(mix)
        all_degs = [None] * len(mix)
        for (mode, sr), (rb, db, idx) in groups.items():  # warm the engines
            native.measure_batch_device(uctx[mode], rb, db, sr, convert=False)
            for j, i in enumerate(idx):
                all_refs[i], all_degs[i] = rb[j], db[j]
        modes = [p[3] for p in mix]
        rates = [p[2] for p in mix]
        ts, ts_seq = [], []
        uneven_mos = np.full(len(mix), np.nan)
        seq_mos = np.full(len(mix), np.nan)
        for _ in range(2):
            # both modes in one call, their batches at once
            t = time.perf_counter()
            _, umos, _ = native.measure_mixed_batch_device(uctx, modes, all_refs, all_degs, rates, convert=False)
            ts.append(time.perf_counter() - t)
            uneven_mos[:] = umos
            # the same pairs as one batch per mode, one after the other
            t = time.perf_counter()
            for (mode, sr), (rb, db, idx) in groups.items():
                _, smos, _ = native.measure_batch_device(uctx[mode], rb, db, sr, convert=False)
                seq_mos[idx] = smos
            ts_seq.append(time.perf_counter() - t)
        secs = float(sum(p[0].size / p[2] for p in mix))
        step = float(min(ts))
        uneven = {"workload": f"{len(mix)} pairs of the uneven mix (durations log-uniform 1-60 s, "
                              f"{sum(1 for p in mix if p[3] == 1)} of them 16 kHz speech), cut from "
                              f"{args.uneven_sources} long source pairs with per-pair noise, dropouts and "
                              "quantisation; one device-resident call for both modes "
                              "(visqol_hip_measure_mixed_batch_device), best of two",
                  "pairs_per_s": round(len(mix) / step, 2), "audio_seconds_per_s": round(secs / step, 1),
                  "ms_per_batch": round(step * 1e3, 2), "failed": int(np.sum(np.isnan(uneven_mos))),
                  "one_batch_per_mode_in_turn_pairs_per_s": round(len(mix) / float(min(ts_seq)), 2),
                  "mixed_equals_per_mode": bool(np.array_equal(uneven_mos, seq_mos, equal_nan=True))}
        for g in groups.values():
            for b in g[0] + g[1]:
                b.free()
        say(f"uneven {uneven}")

    # Serving: Poisson arrivals of the 1-120 s mix, continuous batching.
    serving = None
    served = {}
    if qps_levels:
        levels = []
        with native.Server(precision="fast") as srv:
            serve_level(srv, reqs, 100.0, 1.0, seed=99)  # warm the workers' engines
            for k, q in enumerate(qps_levels):
                rep, order, mos = serve_level(srv, reqs, q, args.serve_seconds, seed=1000 + k)
                levels.append(rep)
                served[q] = (order, mos)
                say(f"serving {rep}")
            stats = srv.stats()
        # the plain f64 per-pair path: service time per request, one at a time
        exc = {0: native.Context(0, precision="exact"), 1: native.Context(1, precision="exact")}
        svc = []
        for ref, deg, sr, mode in reqs[:16]:
            t = time.perf_counter()
            exc[mode].measure(ref, deg, sr)
            svc.append((time.perf_counter() - t) * 1e3)
        secs = float(sum(p[0].size / p[2] for p in reqs))
        serving = {"workload": f"Poisson arrivals for {args.serve_seconds:g} s per load; each request one of "
                               f"{len(reqs)} distinct pairs of the uneven mix (70% 48 kHz audio / 30% 16 kHz "
                               "speech, durations log-uniform 1-120 s, cut from "
                               f"{args.serve_sources} source pairs, per-request noise, dropouts, quantisation; "
                               f"mean {secs / len(reqs):.1f} s), cycled in a seeded random order",
                   "precision": "fast", "levels": levels, "batches": stats["batches"], "pairs": stats["pairs"],
                   "exact_per_pair": {"p50_ms": round(float(np.percentile(svc, 50)), 2),
                                      "p99_ms": round(float(np.percentile(svc, 99)), 2),
                                      "capacity_pairs_per_s": round(1e3 * len(svc) / float(np.sum(svc)), 2),
                                      "note": "service time of one request on an idle device (first 16 requests); "
                                              "offered loads above capacity queue without bound"}}
        # every repeat of a request scored the same (results do not depend on the batch)
        consistent = True
        first = {}
        for order, mos in served.values():
            for r, m in zip(order, mos):
                if np.isnan(m):
                    continue
                consistent &= bool(first.setdefault(int(r), m) == m)
        serving["repeats_identical"] = consistent

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
                               "mode (f64 with FMA, f32-nominated lags rescored in f64), signals resident in HBM",
                   "batch": args.batch, "pair_seconds": 10.0, "sample_rate": 48000, "bands": 32,
                   "precision": "fast"},
        "audio_seconds_per_s": round(value * 10.0, 2),
        "host_buffers_pairs_per_s": round(host_pps, 3) if host_pps else None,
        "host_buffers_scores_match": host_match,
        "exact_per_pair_pairs_per_s": round(exact_pps, 3) if exact_pps else None,
        "exact_batch": (dict(exact_batch, matches_per_pair=bool(np.array_equal(exact_batch_mos[:ne], exact_mos)))
                        if exact_batch else None),
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

        threads = workers
        ns = min(args.batch, args.cpu_sample or 2 * threads)
        orc = vo.Oracle(mode=0)
        rounds = []
        for _ in range(2):  # best of two rounds
            t = time.perf_counter()
            rc, ores, _ = orc.measure_batch([p[0] for p in pairs[:ns]], [p[1] for p in pairs[:ns]], 48000,
                                            num_threads=threads)
            rounds.append(time.perf_counter() - t)
        lat1 = []
        for _ in range(2):
            t = time.perf_counter()
            orc.measure(pairs[0][0], pairs[0][1], 48000)
            lat1.append(time.perf_counter() - t)
        omos = np.array([r["moslqo"] for r in ores])
        dmos = float(np.max(np.abs(omos - fast_mos[:ns])))
        emos = float(np.max(np.abs(omos[:ne] - np.array(exact_mos[:ns])))) if min(ne, ns) > 0 else None
        line["cpu_baseline"] = {"value": round(ns / min(rounds), 3), "unit": "pairs/s", "cores": threads,
                                "kind": "port",
                                "sample": f"first {ns} pairs of the batch, C f64 oracle, one pair per thread, "
                                          "best of two rounds",
                                "rounds_s": [round(x, 3) for x in rounds],
                                "one_thread_latency_s": round(min(lat1), 4)}
        line["max_abs_dmos_vs_oracle"] = dmos
        line["max_abs_dmos_exact_vs_oracle"] = emos
        rng = np.random.default_rng(77)
        if uneven is not None:
            k = max(2, int(round(args.oracle_fraction * len(mix))))
            idx = sorted(rng.choice(len(mix), k, replace=False).tolist())
            worst, nchk = oracle_dmos(vo, mix, idx, uneven_mos[idx], threads)
            uneven["max_abs_dmos_vs_oracle"] = worst
            uneven["oracle_checked_pairs"] = nchk
            say(f"uneven |dMOS| vs oracle on {nchk} pairs: {worst:.3g}")
        if serving is not None:
            # a 1 % sample of all served requests; repeats of one request are
            # identical (checked above), so each distinct one is scored once
            picks = {}
            for order, mos in served.values():
                k = max(1, int(round(args.oracle_fraction * len(order))))
                for j in rng.choice(len(order), k, replace=False):
                    picks.setdefault(int(order[j]), float(mos[j]))
            idx = sorted(picks)
            worst, nchk = oracle_dmos(vo, reqs, idx, np.array([picks[i] for i in idx]), threads)
            serving["oracle_sample"] = {"served_requests_sampled": int(sum(max(1, int(round(
                args.oracle_fraction * len(o)))) for o, _ in served.values())), "distinct_pairs_checked": nchk,
                "max_abs_dmos_vs_oracle": worst}
            say(f"served |dMOS| vs oracle on {nchk} distinct pairs: {worst:.3g}")
        line["dmos_ok"] = dmos <= 1e-6 and (uneven is None or uneven["max_abs_dmos_vs_oracle"] <= 1e-6) and \
            (serving is None or serving["oracle_sample"]["max_abs_dmos_vs_oracle"] <= 1e-6)
    say(f"total {time.perf_counter() - t_all:.1f}s")
    print(json.dumps(line, default=_plain), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
