"""Long pairs, prepared references and the serving mix at full length (GPU).

* One 120 s 48 kHz audio pair (a 2^24-point lag search, 3601 search offsets
  per patch) and one 60 s 16 kHz speech pair: exact mode bit-identical to
  the oracle on every decision and score, fast mode within |dMOS| <= 1e-6
  with the same alignments.
* A 90 s pair with search radius 80, whose DP rows exceed the DP kernel's
  LDS rings (the global-memory path): exact bit-identical, fast within 1e-6.
* Prepared references in exact mode: bit-identical to the oracle on the
  same (reference, degraded) pairs.
* The server under Poisson arrivals (above its capacity) of a few hundred
  requests drawn from the serving mix (70 % 48 kHz audio / 30 % 16 kHz speech, durations
  log-uniform 1-120 s, noise, delay, stretch, dropouts): every result equals
  the same pair scored alone, bit for bit, and sampled results are within
  1e-6 of the oracle.
"""
import sys
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST_MOS_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _trace_equal(tr, otr):
    for k in ("lag", "nf_ref", "nf_deg", "num_patches", "ref_indices", "dp_offset", "match_valid", "realign_lag",
              "realign_frames", "realign_accepted"):
        assert tr[k] == otr[k], k


def _bits_equal(r, o):
    assert r["moslqo"] == o["moslqo"] and r["vnsim"] == o["vnsim"]
    assert r["alignment_lag_s"] == o["alignment_lag_s"]
    for k in ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy"):
        assert np.array_equal(r[k], o[k]), k
    for p, q in zip(r["patch_sims"], o["patch_sims"]):
        for k in ("similarity", "ref_patch_start_time", "ref_patch_end_time", "deg_patch_start_time",
                  "deg_patch_end_time"):
            assert p[k] == q[k], k


@pytest.mark.parametrize("case", ["audio_120s", "speech_60s"])
def test_long_pair_both_modes(case):
    from akvisqol_amd import native, synth
    from oracle import vo

    if case == "audio_120s":
        ref, deg = synth.audio_pair(77, duration=120.0, delay=2345, snr_db=20.0)
        mode, sr = 0, 48000
    else:
        ref, deg = synth.speech_pair(78, duration=60.0, delay=123, snr_db=15.0)
        mode, sr = 1, 16000
    say(f"{case}: oracle (single thread) ...")
    t = time.perf_counter()
    ores, otr = vo.Oracle(mode=mode).measure(ref, deg, sr)
    say(f"  oracle {time.perf_counter() - t:.1f} s, MOS {ores['moslqo']!r}, {otr['num_patches']} patches")
    res, tr = native.Context(mode, precision="exact").measure_trace(ref, deg, sr)
    say(f"  exact MOS {res['moslqo']!r}")
    _trace_equal(tr, otr)
    _bits_equal(res, ores)
    fres, ftr = native.Context(mode, precision="fast").measure_trace(ref, deg, sr)
    say(f"  fast MOS {fres['moslqo']!r} |dMOS| {abs(fres['moslqo'] - ores['moslqo']):.3g}")
    _trace_equal(ftr, otr)
    assert abs(fres["moslqo"] - ores["moslqo"]) <= FAST_MOS_TOL


def test_wide_search_window_exact_and_fast():
    """A 90 s audio pair searched with radius 80 (windows of up to 4801
    frames): the DP's scan ranges exceed its 4096-frame LDS rings, so every
    row takes the global-memory path.  Exact mode bit-identical to the
    oracle, fast mode within 1e-6 with the same alignments."""
    from akvisqol_amd import native, synth
    from oracle import vo

    ref, deg = synth.audio_pair(79, duration=90.0, delay=4321, snr_db=18.0)
    say("wide window: oracle (single thread) ...")
    t = time.perf_counter()
    ores, otr = vo.Oracle(mode=0, search_window_radius=80).measure(ref, deg, 48000)
    say(f"  oracle {time.perf_counter() - t:.1f} s, MOS {ores['moslqo']!r}, {otr['num_patches']} patches")
    res, tr = native.Context(0, precision="exact", search_window_radius=80).measure_trace(ref, deg, 48000)
    _trace_equal(tr, otr)
    _bits_equal(res, ores)
    fres, ftr = native.Context(0, precision="fast", search_window_radius=80).measure_trace(ref, deg, 48000)
    say(f"  fast |dMOS| {abs(fres['moslqo'] - ores['moslqo']):.3g}")
    _trace_equal(ftr, otr)
    assert abs(fres["moslqo"] - ores["moslqo"]) <= FAST_MOS_TOL


@pytest.mark.parametrize("mode", [0, 1])
def test_prepared_exact_equals_oracle(mode):
    from akvisqol_amd import native, synth
    from oracle import vo

    sr = 48000 if mode == 0 else 16000
    refs, pairs = [], []
    for k in range(3):
        for j, (snr, delay) in enumerate(((14.0, 300 + 50 * k), (28.0, 0))):
            if mode == 0:
                ref, deg = synth.audio_pair(90 + k, duration=4.0 + k, snr_db=snr, delay=delay)
            else:
                ref, deg = synth.speech_pair(90 + k, duration=5.0 + k, snr_db=snr, delay=delay // 4)
            if j == 0:
                refs.append(ref)
            pairs.append((k, deg))
    say(f"prepared exact mode={mode}: {len(refs)} refs, {len(pairs)} degraded signals")
    ctx = native.Context(mode, precision="exact")
    prep = ctx.prepare(refs, sr)
    rc, got, st = ctx.measure_prepared(prep, [d for _, d in pairs], ref_index=[k for k, _ in pairs])
    assert rc == 0
    rc2, want, st2 = vo.Oracle(mode=mode).measure_batch([refs[k] for k, _ in pairs], [d for _, d in pairs], sr)
    assert rc2 == 0
    for g, w in zip(got, want):
        _bits_equal(g, w)
    prep.free()


def _requests(n_pool, n_req, seed):
    """n_req requests of the serving mix: each a slice, of a duration drawn
    log-uniform 1-120 s (capped by the pair), at a random offset, of one of
    n_pool distinct pairs of that mix (generated on host threads)."""
    from concurrent.futures import ThreadPoolExecutor

    from akvisqol_amd import synth

    with ThreadPoolExecutor(max_workers=16) as ex:
        pool = list(ex.map(lambda k: synth.mixed_pair(20_000 + k, max_duration=120.0), range(n_pool)))
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        ref, deg, sr, mode = pool[int(rng.integers(0, n_pool))]
        d = int(np.exp(rng.uniform(np.log(1.0), np.log(120.0))) * sr)
        d = min(d, ref.size, deg.size)
        o = int(rng.integers(0, min(ref.size, deg.size) - d + 1))
        reqs.append((ref[o:o + d], deg[o:o + d], sr, mode))
    return reqs


def test_server_poisson_config4_mix():
    """240 requests of the 1-120 s serving mix arriving as a Poisson stream
    at 3000 req/s, about three times what the device serves of this mix,
    so requests queue while others run and batches form and overlap: each
    served result equals the same pair scored alone, bit for bit (failures
    with the same code); four sampled pairs are within 1e-6 of the
    oracle."""
    from akvisqol_amd import native
    from oracle import vo

    n = 240
    t = time.perf_counter()
    pairs = _requests(48, n, 4)
    secs = sum(p[0].size / p[2] for p in pairs)
    say(f"generated {n} requests ({secs:.0f} s of audio, {sum(p[3] for p in pairs)} speech) in "
        f"{time.perf_counter() - t:.1f} s")
    rng = np.random.default_rng(4)
    arrivals = np.cumsum(rng.exponential(1.0 / 3000.0, n))
    tickets = [None] * n
    done = [threading.Event() for _ in range(n)]
    out = [None] * n
    with native.Server(precision="fast") as srv:
        start = time.perf_counter()

        def submitter():
            for i in range(n):
                d = start + arrivals[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ref, deg, sr, mode = pairs[i]
                tickets[i] = srv.submit(mode, ref, deg, sr)
                done[i].set()

        th = threading.Thread(target=submitter)
        th.start()
        for i in range(n):
            done[i].wait()
            out[i] = srv.wait_timed(tickets[i])
        th.join()
        stats = srv.stats()
    lat = np.array([o[2]["total_ms"] for o in out])
    say(f"served {n} requests in {time.perf_counter() - start:.1f} s: {stats['batches']} batches, "
        f"p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms")
    assert stats["pairs"] == n and stats["inflight_bytes"] == 0
    assert stats["batches"] < n  # requests shared batches
    ctxs = {0: native.Context(0, precision="fast"), 1: native.Context(1, precision="fast")}
    failed = 0
    for (ref, deg, sr, mode), (rc, res, _) in zip(pairs, out):
        try:
            alone, arc = ctxs[mode].measure(ref, deg, sr), 0
        except native.NativeError as e:
            alone, arc = None, e.code
        assert rc == arc
        if rc:
            failed += 1
            continue
        _bits_equal(res, alone)
    say(f"  every result equals the pair alone ({failed} failed the same way alone)")
    assert failed < n // 10
    # a sample against the oracle: the two shortest of each mode
    for mode in (0, 1):
        idx = sorted((i for i in range(n) if pairs[i][3] == mode and out[i][0] == 0),
                     key=lambda i: pairs[i][0].size)[:2]
        for i in idx:
            ref, deg, sr, _ = pairs[i]
            ores, _ = vo.Oracle(mode=mode).measure(ref, deg, sr)
            say(f"  pair {i} ({ref.size / sr:.1f} s, mode {mode}): |dMOS| {abs(out[i][1]['moslqo'] - ores['moslqo']):.3g}")
            assert abs(out[i][1]["moslqo"] - ores["moslqo"]) <= FAST_MOS_TOL
