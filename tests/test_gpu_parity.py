"""Parity of the gfx950 path against the CPU f64 oracle, through the C ABI.

Exact mode (VISQOL_PRECISION_EXACT) is held to bit equality on every
stage and on the final scores: the device uses the oracle's operation
order, host-designed coefficients and the same log10 algorithm, so any
difference is a bug.  (The stated tolerance is |dMOS| <= 1e-10 with
identical alignments; the tests assert the stronger bit equality.)  Fast
mode is held to |dMOS| <= 1e-6.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT_MOS_TOL = 1e-10
FAST_MOS_TOL = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def native():
    from akvisqol_amd import native as n

    return n


@pytest.fixture(scope="module")
def vo():
    from oracle import vo as v

    return v


@pytest.fixture(scope="module")
def synth():
    from akvisqol_amd import synth as s

    return s


# ---------------------------------------------------------------- device arithmetic
def test_device_log10_sqrt_div_bit_exact(native, vo):
    say("probe: log10/sqrt/div on 200k values")
    rng = np.random.default_rng(1)
    a = np.concatenate([np.exp(rng.uniform(-700, 700, 100000)), rng.uniform(1e-18, 1e3, 100000),
                        [2.220446049250313e-16, 1.0, 10.0, 1e-310, 5e-324]])
    got = native.math_probe(0, a)
    want = np.array([vo.lib().vo_log10(float(x)) for x in a])
    assert np.array_equal(got, want)
    got = native.math_probe(1, a)
    assert np.array_equal(got, np.sqrt(a))
    b = rng.uniform(-1e3, 1e3, a.size)
    got = native.math_probe(2, a, b)
    assert np.array_equal(got, a / b)


# ---------------------------------------------------------------- stages
@pytest.mark.parametrize("mode,sr,n", [(0, 48000, 3840 * 3 + 960 * 70 + 17), (1, 16000, 1280 * 2 + 320 * 41 + 5)])
def test_spectrogram_exact_bit_identical(native, vo, mode, sr, n):
    say(f"spectrogram exact mode={mode} sr={sr} n={n}")
    rng = np.random.default_rng(sr + n)
    x = rng.uniform(-1, 1, n)
    got = native.spectrogram(x, sr, mode, "exact")
    want = vo.spectrogram(x, sr, 32 if mode == 0 else 21, None if mode == 0 else 8000.0)
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode,sr", [(0, 48000), (1, 16000)])
def test_spectrogram_fast_close(native, vo, mode, sr):
    """The throughput spectrogram (fused cascade below band 8, FIR +
    all-pole above) against the oracle's, audio at 48 kHz and speech at
    16 kHz (21 bands up to 8 kHz)."""
    say(f"spectrogram fast mode={mode} sr={sr}")
    rng = np.random.default_rng(5 + mode)
    window, hop = (3840, 960) if sr == 48000 else (1280, 320)
    x = rng.uniform(-1, 1, window + hop * 40)
    got = native.spectrogram(x, sr, mode, "fast")
    want = vo.spectrogram(x, sr, 32 if mode == 0 else 21, None if mode == 0 else 8000.0)
    assert got.shape == want.shape
    rel = np.max(np.abs(got - want) / np.maximum(want, 1e-300))
    say("fast spectrogram max rel err", rel)
    assert rel < 1e-9


@pytest.mark.parametrize("n", [2, 64, 4096, 1 << 13, 1 << 16, 1 << 18, 1 << 20])
def test_fft_bit_identical(native, vo, n):
    say(f"fft DIF/DIT n={n}")
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    got = native.fft(x, dit=False)
    ref = vo.f64(np.stack([x.real, x.imag], 1).ravel())
    vo.lib().vo_fft_dif_forward(vo.dptr(ref), n)
    assert np.array_equal(got, ref[0::2] + 1j * ref[1::2])
    back = native.fft(got, dit=True)
    vo.lib().vo_fft_dit_inverse(vo.dptr(ref), n)
    assert np.array_equal(back, ref[0::2] + 1j * ref[1::2])


REF14 = [2.0, 2.0, 1.0, 0.1, -3.0, 0.1, 1.0, 2.0, 2.0, 6.0, 8.0, 6.0, 2.0, 2.0]


def test_best_lag_matches_oracle(native, vo):
    say("best lag toys + random")
    cases = [
        (REF14, [1.2, 0.1, -3.3, 0.1, 1.1, 2.2, 2.1, 7.1, 8.3, 6.8, 2.4, 2.2, 2.2, 2.1]),
        (REF14, [2.0, 2.0, 2.0, 2.0, 1.0, 0.1, -3.0, 0.1, 1.0, 2.0, 2.0, 6.0, 8.0, 6.0]),
    ]
    rng = np.random.default_rng(3)
    for _ in range(8):
        l1, l2 = int(rng.integers(50, 9000)), int(rng.integers(50, 9000))
        a = rng.uniform(-1, 1, l1)
        b = np.roll(np.resize(a, l2), int(rng.integers(0, 40))) + 0.1 * rng.standard_normal(l2)
        cases.append((a, b))
    for a, b in cases:
        a, b = np.asarray(a, float), np.asarray(b, float)
        ea, eb = np.zeros(a.size), np.zeros(b.size)
        vo.lib().vo_upper_envelope(vo.dptr(vo.f64(a)), a.size, vo.dptr(ea))
        vo.lib().vo_upper_envelope(vo.dptr(vo.f64(b)), b.size, vo.dptr(eb))
        want = vo.lib().vo_best_lag(vo.dptr(ea), a.size, vo.dptr(eb), b.size)
        assert native.best_lag(a, b) == want


@pytest.mark.parametrize("lg", [2, 5, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 24])
def test_xcorr_bit_identical(native, vo, lg):
    """Every value of the device cross-correlation (envelopes built in the
    first forward pass, fused middle pass, inverse passes) equals the
    oracle's, for transform sizes 2^lg covering each pass plan."""
    rng = np.random.default_rng(100 + lg)
    la = max(2, (1 << (lg - 1)) - 3)
    lb = max(1, la - 17)
    say(f"xcorr n=2^{lg} la={la} lb={lb}")
    a = rng.uniform(-1, 1, la)
    b = np.roll(np.resize(a, lb), 11) + 0.05 * rng.standard_normal(lb)
    ea, eb = np.zeros(la), np.zeros(lb)
    vo.lib().vo_upper_envelope(vo.dptr(vo.f64(a)), la, vo.dptr(ea))
    vo.lib().vo_upper_envelope(vo.dptr(vo.f64(b)), lb, vo.dptr(eb))
    n = vo.lib().vo_fft_size_for(max(la, lb))
    want = np.zeros(n)
    assert vo.lib().vo_xcorr_re(vo.dptr(ea), la, vo.dptr(eb), lb, vo.dptr(want)) == 0
    got = native.xcorr(a, b)
    assert got.size == n
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


# ---------------------------------------------------------------- whole pipeline
def _compare(res, tr, ores, otr, exact=True, tol=EXACT_MOS_TOL):
    assert tr["lag"] == otr["lag"]
    assert tr["nf_ref"] == otr["nf_ref"] and tr["nf_deg"] == otr["nf_deg"]
    assert tr["num_patches"] == otr["num_patches"]
    assert tr["ref_indices"] == otr["ref_indices"]
    assert tr["dp_offset"] == otr["dp_offset"]
    assert tr["match_valid"] == otr["match_valid"]
    assert tr["realign_lag"] == otr["realign_lag"]
    assert tr["realign_frames"] == otr["realign_frames"]
    assert tr["realign_accepted"] == otr["realign_accepted"]
    assert len(res["patch_sims"]) == len(ores["patch_sims"])
    for p, q in zip(res["patch_sims"], ores["patch_sims"]):
        for k in ("ref_patch_start_time", "ref_patch_end_time", "deg_patch_start_time", "deg_patch_end_time"):
            assert p[k] == q[k], k
    assert abs(res["moslqo"] - ores["moslqo"]) <= tol
    if exact:
        assert tr["spl_factor"] == otr["spl_factor"]
        assert res["moslqo"] == ores["moslqo"]
        assert res["vnsim"] == ores["vnsim"]
        for k in ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy", "center_freq_bands"):
            assert np.array_equal(res[k], ores[k]), k
        for p, q in zip(res["patch_sims"], ores["patch_sims"]):
            assert p["similarity"] == q["similarity"]
            for k in ("freq_band_means", "freq_band_stddevs", "freq_band_deg_energy"):
                assert np.array_equal(p[k], q[k]), k
    assert res["alignment_lag_s"] == ores["alignment_lag_s"]


def test_audio_config0_exact_stages_and_score(native, vo, synth):
    say("config-0 audio pair, exact, stage by stage")
    ref, deg = synth.audio_pair(0)
    ctx = native.Context(native.MODE_AUDIO, precision="exact")
    res, tr = ctx.measure_trace(ref, deg, 48000, want_stages=True)
    ores, otr = vo.Oracle(mode=0).measure(ref, deg, 48000, want_spec=True)
    say("gpu mos", res["moslqo"], "oracle mos", ores["moslqo"], "lag", tr["lag"])
    assert np.array_equal(tr["ref_spec"], otr["ref_spec"])
    assert np.array_equal(tr["deg_spec"], otr["deg_spec"])
    assert np.array_equal(tr["sims"], otr["sims"])
    _compare(res, tr, ores, otr)
    assert tr["lag"] == -37


@pytest.mark.parametrize("lattice", [True, False])
def test_speech_exact(native, vo, synth, lattice):
    say(f"speech pair exact lattice={lattice}")
    ref, deg = synth.speech_pair(3)
    ctx = native.Context(native.MODE_SPEECH, use_lattice_model=lattice, precision="exact")
    res, tr = ctx.measure_trace(ref, deg, 16000, want_stages=True)
    ores, otr = vo.Oracle(mode=1, use_lattice_model=lattice).measure(ref, deg, 16000, want_spec=True)
    say("gpu mos", res["moslqo"], "oracle mos", ores["moslqo"])
    assert np.array_equal(tr["ref_spec"], otr["ref_spec"])
    assert np.array_equal(tr["deg_spec"], otr["deg_spec"])
    _compare(res, tr, ores, otr)


def test_options_exact(native, vo, synth):
    say("options: no global alignment / no realignment / radius")
    ref, deg = synth.audio_pair(4, duration=6.0, delay=500, snr_db=15.0)
    for kw in (dict(disable_global_alignment=True), dict(disable_realignment=True), dict(search_window_radius=5)):
        res, tr = native.Context(native.MODE_AUDIO, precision="exact", **kw).measure_trace(ref, deg, 48000)
        ores, otr = vo.Oracle(mode=0, **kw).measure(ref, deg, 48000)
        say(kw, res["moslqo"], ores["moslqo"])
        _compare(res, tr, ores, otr)


def test_batch_equals_single_and_isolates_failures(native, vo, synth):
    say("batch == single; a too-short pair fails alone")
    pairs = [synth.audio_pair(s, duration=d) for s, d in ((10, 5.0), (11, 7.5), (12, 6.0))]
    short = (np.zeros(1000), np.zeros(1000))
    refs = [p[0] for p in pairs] + [short[0]]
    degs = [p[1] for p in pairs] + [short[1]]
    ctx = native.Context(native.MODE_AUDIO, precision="exact")
    rc, results, status = ctx.measure_batch(refs, degs, 48000)
    assert rc == native.E_SIGNAL and status == [0, 0, 0, native.E_SIGNAL] and results[3] is None
    for (r, d), got in zip(pairs, results[:3]):
        alone = ctx.measure(r, d, 48000)
        assert got["moslqo"] == alone["moslqo"]
        assert np.array_equal(got["fvnsim"], alone["fvnsim"])
        ores, _ = vo.Oracle(mode=0).measure(r, d, 48000)
        assert got["moslqo"] == ores["moslqo"]


def test_errors(native):
    say("error paths")
    ctx = native.Context(native.MODE_AUDIO)
    with pytest.raises(native.NativeError) as e:
        ctx.measure(np.zeros(100), np.zeros(100), 48000)
    assert e.value.code == native.E_SIGNAL
    with pytest.raises(native.NativeError) as e:
        ctx.measure(np.zeros(100000), np.zeros(100000), 0)
    assert e.value.code == native.E_INVALID_ARGUMENT


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_degenerate_inputs_match_oracle(native, vo, synth, precision):
    """Silence, one silent side, DC, identical signals, unequal lengths and
    extreme levels (below the dB floor, far above full scale), in both
    modes: exact equals the oracle bit for bit, fast within 1e-6 MOS with
    the same global lag."""
    ref, deg = synth.audio_pair(3, duration=2.0, delay=50)
    sref, sdeg = synth.speech_pair(2, duration=3.0)
    cases = [
        ("silence", 0, 48000, np.zeros(96000), np.zeros(96000)),
        ("silent degraded", 0, 48000, ref, np.zeros(96000)),
        ("silent reference", 0, 48000, np.zeros(96000), deg),
        ("dc", 0, 48000, np.full(96000, 0.3), np.full(96000, 0.3)),
        ("identical", 0, 48000, ref, ref.copy()),
        ("degraded shorter", 0, 48000, ref, deg[:50000]),
        ("degraded longer", 0, 48000, ref, np.concatenate([deg, deg])),
        ("below the floor", 0, 48000, ref * 1e-9, deg * 1e-9),
        ("loud", 0, 48000, ref * 1e4, deg * 1e4),
        ("speech silence", 1, 16000, np.zeros(48000), np.zeros(48000)),
        ("speech identical", 1, 16000, sref, sref.copy()),
        ("speech silent degraded", 1, 16000, sref, np.zeros(len(sref))),
    ]
    ctxs = {m: native.Context(m, precision=precision) for m in (0, 1)}
    for name, mode, sr, r, d in cases:
        say(f"degenerate {name} mode={mode} precision={precision}")
        res, tr = ctxs[mode].measure_trace(r, d, sr)
        ores, otr = vo.Oracle(mode=mode).measure(r, d, sr)
        assert tr["lag"] == otr["lag"], name
        if precision == "exact":
            assert res["moslqo"] == ores["moslqo"], name
            assert np.array_equal(res["fvnsim"], ores["fvnsim"]), name
            assert tr["dp_offset"] == otr["dp_offset"], name
        else:
            assert abs(res["moslqo"] - ores["moslqo"]) <= 1e-6, (name, res["moslqo"], ores["moslqo"])


def test_fast_mode_close(native, vo, synth):
    say("fast mode |dMOS| <= 1e-6")
    ref, deg = synth.audio_pair(0)
    res, tr = native.Context(native.MODE_AUDIO, precision="fast").measure_trace(ref, deg, 48000)
    ores, otr = vo.Oracle(mode=0).measure(ref, deg, 48000)
    say("fast mos", res["moslqo"], "oracle", ores["moslqo"], "diff", res["moslqo"] - ores["moslqo"])
    _compare(res, tr, ores, otr, exact=False, tol=FAST_MOS_TOL)


# ---------------------------------------------------------------- per-pair f64 check set
def _bitsame(r, o):
    assert r["moslqo"] == o["moslqo"] and r["vnsim"] == o["vnsim"]
    assert r["alignment_lag_s"] == o["alignment_lag_s"]
    for k in ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy"):
        assert np.array_equal(r[k], o[k]), k
    assert len(r["patch_sims"]) == len(o["patch_sims"])
    for p, q in zip(r["patch_sims"], o["patch_sims"]):
        for k in ("similarity", "ref_patch_start_time", "ref_patch_end_time", "deg_patch_start_time",
                  "deg_patch_end_time"):
            assert p[k] == q[k], k


@pytest.mark.parametrize("mode", [0, 1])
def test_plain_f64_loop_64_pairs(native, vo, synth, mode):
    """The per-pair check set: 64 audio pairs (5-15 s, SNR 10-40 dB, delay
    0-2000 samples) or 64 speech pairs (8 s, lattice, VAD), scored one at a
    time on the exact path: each bit-identical to the oracle, so every
    alignment (global lag, patch times, realignment) is identical."""
    if mode == 0:
        pairs = [synth.plain_audio_pair(s) for s in range(64)]
        sr = 48000
    else:
        pairs = [synth.speech_pair(s) for s in range(64)]
        sr = 16000
    say(f"per-pair f64 loop mode={mode}: 64 pairs")
    ctx = native.Context(mode, precision="exact")
    got = [ctx.measure(r, d, sr) for r, d in pairs]
    rc, want, st = vo.Oracle(mode=mode).measure_batch([p[0] for p in pairs], [p[1] for p in pairs], sr)
    assert rc == 0
    for g, w in zip(got, want):
        _bitsame(g, w)


# ---------------------------------------------------------------- spectrogram block shapes
def _ctx_with_shape(native, mode, precision, shape):
    import os

    key = "VISQOL_HIP_GAMMATONE_SHAPE"
    old = os.environ.get(key)
    os.environ[key] = str(shape)
    try:
        return native.Context(mode, precision=precision)
    finally:
        if old is None:
            del os.environ[key]
        else:
            os.environ[key] = old


@pytest.mark.parametrize("precision", ["exact", "fast"])
def test_gammatone_shapes_bit_identical(native, synth, precision):
    """Every block shape of the spectrogram kernel gives the same bits: the
    spectrograms, the similarity tables and the scores (realignment
    spectrograms included) equal those of shape 0."""
    shapes = native.gammatone_shapes()
    assert len(shapes) >= 2
    cases = [(0, 48000, synth.audio_pair(2, duration=4.0, delay=300)),
             (1, 16000, synth.speech_pair(3, duration=4.0, delay=40))]
    for mode, sr, (ref, deg) in cases:
        base = None
        for v, sh in enumerate(shapes):
            say(f"shape {v} {sh} mode={mode} precision={precision}")
            res, tr = _ctx_with_shape(native, mode, precision, v).measure_trace(ref, deg, sr, want_stages=True)
            if base is None:
                base = (res, tr)
                continue
            for k in ("ref_spec", "deg_spec", "sims"):
                assert np.array_equal(tr[k], base[1][k]), (v, k)
            _bitsame(res, base[0])


def test_tuned_batch_equals_first_shape(native, synth):
    """A batch large enough to tune the spectrogram shape (whole signals and
    realignment slices) scores exactly as the same batch on shape 0."""
    # 384 pairs x 2 signals x 3 blocks of 64 frames x 2 groups of 16 bands
    # = 4608 units of work in one chunk (signals on the device: chunks of
    # 2^28 samples), above the 4096 the engine needs before it times the
    # shapes
    pairs = [synth.audio_pair(40 + s, duration=2.7, delay=50 * s) for s in range(8)] * 48
    bufs = [(native.DeviceBuffer(r), native.DeviceBuffer(d)) for r, d in pairs[:8]]
    refs, degs = [bufs[i % 8][0] for i in range(len(pairs))], [bufs[i % 8][1] for i in range(len(pairs))]
    say(f"tuned batch: {len(pairs)} pairs of 2.7 s")
    ctx = native.Context(native.MODE_AUDIO, precision="fast")
    rc, tuned, st = native.measure_batch_device(ctx, refs, degs, 48000)
    assert rc == 0
    t = ctx.gammatone_tuning(48000, 0)
    say("tuning", t)
    assert t["chosen"] >= 0 and len(t["ms"]) == len(t["shapes"])
    rc, first, _ = native.measure_batch_device(_ctx_with_shape(native, native.MODE_AUDIO, "fast", 0), refs, degs, 48000)
    assert rc == 0
    for a, b in zip(tuned, first):
        _bitsame(a, b)
    for r, d in bufs:
        r.free()
        d.free()
