"""Uneven mixed batches and the continuous-batching server on the device.

* Uneven batches (audio 48 kHz and speech 16 kHz pairs of different
  lengths, with delay, time-stretch, dropouts and noise): exact mode equals
  the f64 oracle bit for bit, with identical alignments, pair by pair.
* Serving: each request's result from visqol_hip_server equals the same
  pair scored alone through visqol_measure, bit for bit (fast mode), whether
  requests share batches, overlap across workers, or are admitted one at a
  time by a tight footprint budget.
"""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def native():
    from akvisqol_amd import native as n

    return n


@pytest.fixture(scope="module")
def synth():
    from akvisqol_amd import synth as s

    return s


@pytest.fixture(scope="module")
def vo():
    from oracle import vo as v

    return v


@pytest.fixture(scope="module")
def mix(synth):
    """Twelve pairs of the uneven mix, 1-12 s, both modes."""
    pairs = [synth.mixed_pair(seed, max_duration=12.0) for seed in range(12)]
    modes = {p[3] for p in pairs}
    assert modes == {0, 1}, modes
    return pairs


def _same(a, b):
    assert a["moslqo"] == b["moslqo"]
    assert a["vnsim"] == b["vnsim"]
    assert a["alignment_lag_s"] == b["alignment_lag_s"]
    for k in ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy", "center_freq_bands"):
        assert np.array_equal(a[k], b[k]), k
    assert len(a["patch_sims"]) == len(b["patch_sims"])
    for p, q in zip(a["patch_sims"], b["patch_sims"]):
        for k in ("similarity", "ref_patch_start_time", "ref_patch_end_time", "deg_patch_start_time",
                  "deg_patch_end_time"):
            assert p[k] == q[k], k
        for k in ("freq_band_means", "freq_band_stddevs", "freq_band_deg_energy"):
            assert np.array_equal(p[k], q[k]), k


def test_uneven_mixed_batch_exact_matches_oracle(native, vo, mix):
    """One batch per mode over pairs of uneven length: bit-identical to the
    oracle, pair by pair."""
    for mode, sr in ((0, 48000), (1, 16000)):
        sel = [p for p in mix if p[3] == mode]
        say(f"uneven exact batch mode={mode} n={len(sel)} lens={[p[0].size for p in sel]}")
        ctx = native.Context(mode, precision="exact")
        rc, res, st = ctx.measure_batch([p[0] for p in sel], [p[1] for p in sel], sr)
        orc = vo.Oracle(mode=mode)
        orc_rc, ores, ost = orc.measure_batch([p[0] for p in sel], [p[1] for p in sel], sr)
        assert list(st) == list(ost)
        for r, o in zip(res, ores):
            if o is None:
                assert r is None
                continue
            _same(r, o)


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_mixed_batch_equals_per_mode_batches(native, mix, precision):
    """visqol_hip_measure_mixed_batch_device over audio and speech pairs
    (two contexts, their batches at once, interleaved in pair order, plus
    one pair too short to score): every result equals the same context's
    own device batch bit for bit, and the short pair fails alone with
    VISQOL_E_SIGNAL."""
    pairs = list(mix) + [(mix[0][0][:2000], mix[0][1][:2000], mix[0][2], mix[0][3])]
    ctxs = [native.Context(0, precision=precision), native.Context(1, precision=precision)]
    refs = [native.DeviceBuffer(p[0]) for p in pairs]
    degs = [native.DeviceBuffer(p[1]) for p in pairs]
    say(f"mixed batch {precision}: n={len(pairs)} modes={[p[3] for p in pairs]}")
    rc, res, st = native.measure_mixed_batch_device(ctxs, [p[3] for p in pairs], refs, degs, [p[2] for p in pairs])
    assert st[-1] == native.E_SIGNAL and res[-1] is None
    assert rc == native.E_SIGNAL
    assert all(s == 0 for s in st[:-1]), st
    for mode in (0, 1):
        idx = [i for i, p in enumerate(pairs[:-1]) if p[3] == mode]
        sr = pairs[idx[0]][2]
        rc1, own, st1 = native.measure_batch_device(ctxs[mode], [refs[i] for i in idx], [degs[i] for i in idx], sr)
        assert rc1 == 0 and all(s == 0 for s in st1)
        for i, r in zip(idx, own):
            _same(res[i], r)
    for b in refs + degs:
        b.free()


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_uneven_batch_equals_pairs_alone(native, mix, precision):
    """An uneven batch's whole-signal spectrograms run the persistent
    unit-list form (most of the grid form's blocks would lie past a shorter
    job's end); each pair alone runs the grid form: every result is the
    same, bit for bit."""
    for mode, sr in ((0, 48000), (1, 16000)):
        sel = [p for p in mix if p[3] == mode]
        ctx = native.Context(mode, precision=precision)
        say(f"uneven batch vs alone: mode={mode} precision={precision} lens={[p[0].size for p in sel]}")
        rc, batch, st = ctx.measure_batch([p[0] for p in sel], [p[1] for p in sel], sr)
        assert rc == 0 and set(st) == {0}
        for p, b in zip(sel, batch):
            _same(b, ctx.measure(p[0], p[1], sr))


def _serve(native, pairs, **cfg):
    with native.Server(precision="fast", **cfg) as srv:
        tickets = [None] * len(pairs)

        def submitter(idx):
            for i in idx:
                ref, deg, sr, mode = pairs[i]
                tickets[i] = srv.submit(mode, ref, deg, sr)

        # two submitting threads interleave the requests
        ts = [threading.Thread(target=submitter, args=(range(k, len(pairs), 2),)) for k in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        out = [srv.wait(t) for t in tickets]
        stats = srv.stats()
    return out, stats


def test_server_results_equal_alone(native, mix):
    alone = []
    ctxs = {0: native.Context(0, precision="fast"), 1: native.Context(1, precision="fast")}
    for ref, deg, sr, mode in mix:
        alone.append(ctxs[mode].measure(ref, deg, sr))
    say("serving the mix with 3 workers, batches of up to 4")
    out, stats = _serve(native, mix, workers=3, max_batch=4)
    say("server stats", stats)
    assert stats["pairs"] == len(mix) and stats["inflight_bytes"] == 0
    for (rc, res, lat), a in zip(out, alone):
        assert rc == 0 and lat > 0
        _same(res, a)


def test_server_tight_admission(native, mix):
    """A one-byte footprint budget admits one request at a time; results are
    unchanged."""
    ctx = native.Context(0, precision="fast")
    sel = [p for p in mix if p[3] == 0][:4]
    alone = [ctx.measure(p[0], p[1], p[2]) for p in sel]
    say("serving with a one-byte admission budget")
    out, stats = _serve(native, sel, workers=2, max_batch=8, max_inflight_bytes=1)
    assert stats["batches"] == len(sel)  # never more than one request in flight
    for (rc, res, lat), a in zip(out, alone):
        assert rc == 0
        _same(res, a)


def test_server_errors(native, synth):
    ref, deg = synth.audio_pair(3, duration=2.0)
    with native.Server(precision="fast", workers=1) as srv:
        t_bad = srv.submit(0, ref[:100], deg[:100], 48000)  # shorter than one window
        t_ok = srv.submit(0, ref, deg, 48000)
        rc, res, _ = srv.wait(t_bad)
        assert rc == native.E_SIGNAL and res is None
        rc, res, _ = srv.wait(t_ok)
        assert rc == 0 and 1.0 <= res["moslqo"] <= 5.0
        rc, res, _ = srv.wait(987654)  # never issued
        assert rc == native.E_INVALID_ARGUMENT and res is None


@pytest.mark.parametrize("mode,precision", [(0, "fast"), (0, "exact"), (1, "fast"), (1, "exact")])
def test_prepared_references_equal_batch(native, synth, mode, precision):
    """Prepared references (spectrogram, sums, VAD kept on the device) give
    the scores of the plain batch on the same pairs, several degraded
    signals per reference."""
    if mode == 0:
        sr = 48000
        refs, pairs = [], []
        for k, seed in enumerate((11, 12, 13)):
            dur = 3.0 + k
            for j, (snr, delay) in enumerate(((15.0, 120), (30.0, 0))):
                ref, deg = synth.audio_pair(seed, duration=dur, snr_db=snr, delay=delay)
                if j == 0:
                    refs.append(ref)
                pairs.append((k, deg))
    else:
        sr = 16000
        refs, pairs = [], []
        for k, seed in enumerate((21, 22)):
            for j, (snr, delay) in enumerate(((12.0, 40), (25.0, 0))):
                ref, deg = synth.speech_pair(seed, duration=6.0, snr_db=snr, delay=delay)
                if j == 0:
                    refs.append(ref)
                pairs.append((k, deg))
    ctx = native.Context(mode, precision=precision)
    say(f"prepared refs mode={mode} precision={precision}: {len(refs)} refs, {len(pairs)} degs")
    prep = ctx.prepare(refs, sr)
    rc, got, st = ctx.measure_prepared(prep, [d for _, d in pairs], ref_index=[k for k, _ in pairs])
    rc2, want, st2 = ctx.measure_batch([refs[k] for k, _ in pairs], [d for _, d in pairs], sr)
    assert rc == rc2 == 0 and list(st) == list(st2)
    for g, w in zip(got, want):
        _same(g, w)
    other = native.Context(mode, precision="exact" if precision == "fast" else "fast")
    rc, res, _ = other.measure_prepared(prep, [pairs[0][1]])
    assert rc == native.E_INVALID_ARGUMENT
    rc, res, _ = ctx.measure_prepared(prep, [pairs[0][1]], ref_index=[len(refs)])
    assert rc == native.E_INVALID_ARGUMENT
    prep.free()


@pytest.mark.parametrize("precision", ["fast", "exact"])
def test_multi_chunk_batch_equals_single_chunk(native, synth, precision):
    """A batch of more than 512 pairs runs as several chunks, every other
    chunk on a helper engine with a stream of its own: each pair's result
    equals the same pair scored in a one-chunk batch, bit for bit."""
    base = [synth.audio_pair(50 + s, duration=1.2, snr_db=12.0 + 3 * s, delay=20 * s) for s in range(8)]
    pairs = base * 65  # 520 pairs: chunks of 512 and 8
    ctx = native.Context(0, precision=precision)
    say(f"multi-chunk batch precision={precision}: {len(pairs)} pairs")
    rc, got, st = ctx.measure_batch([p[0] for p in pairs], [p[1] for p in pairs], 48000)
    assert rc == 0 and set(st) == {0}
    rc, want, _ = ctx.measure_batch([p[0] for p in base], [p[1] for p in base], 48000)
    assert rc == 0
    for i, g in enumerate(got):
        _same(g, want[i % len(base)])


@pytest.mark.parametrize("precision", ["fast", "exact"])
def test_device_multi_chunk_batch_equals_single_chunk(native, synth, precision):
    """A device-resident batch of 1030 pairs: greedy chunks of 512, 512
    and 6 are re-cut into four equal chunks of about 258 for the two
    scorers (each on its own stream, the spectrogram launches fenced); each
    pair's result equals the same pair in a one-chunk batch, bit for bit."""
    base = [synth.audio_pair(90 + s, duration=1.1, snr_db=14.0 + 2 * s, delay=15 * s) for s in range(10)]
    refs = [native.DeviceBuffer(p[0]) for p in base]
    degs = [native.DeviceBuffer(p[1]) for p in base]
    ctx = native.Context(0, precision=precision)
    rc, want, _ = native.measure_batch_device(ctx, refs, degs, 48000)
    assert rc == 0
    n = 1030
    say(f"device multi-chunk batch precision={precision}: {n} pairs")
    rc, got, st = native.measure_batch_device(ctx, [refs[i % 10] for i in range(n)], [degs[i % 10] for i in range(n)],
                                              48000)
    assert rc == 0 and set(st) == {0}
    for i, g in enumerate(got):
        _same(g, want[i % len(base)])
    for b in refs + degs:
        b.free()


def test_prefetched_chunks_equal_single_chunk(native, synth):
    """Host arrays of a batch of several chunks go up one chunk ahead of
    scoring, into a ring of three device buffers: with five chunks the ring
    is reused, and each pair's result still equals the one-chunk batch's,
    bit for bit, for plain and for prepared references (whose degraded
    signals alone are uploaded)."""
    base = [synth.audio_pair(70 + s, duration=1.2, snr_db=15.0 + 2 * s, delay=10 * s) for s in range(8)]
    pairs = base * 257  # 2056 pairs: chunks of 512 x 4 and 8
    ctx = native.Context(0, precision="fast")
    say(f"prefetched chunks: {len(pairs)} pairs")
    rc, want, _ = ctx.measure_batch([p[0] for p in base], [p[1] for p in base], 48000)
    assert rc == 0
    rc, got, st = ctx.measure_batch([p[0] for p in pairs], [p[1] for p in pairs], 48000)
    assert rc == 0 and set(st) == {0}
    for i, g in enumerate(got):
        _same(g, want[i % len(base)])
    say("prefetched chunks: prepared references")
    prep = ctx.prepare([p[0] for p in base], 48000)
    rc, pgot, pst = ctx.measure_prepared(prep, [p[1] for p in pairs], ref_index=[i % len(base) for i in range(len(pairs))])
    assert rc == 0 and set(pst) == {0}
    for i, g in enumerate(pgot):
        _same(g, want[i % len(base)])
    prep.free()
