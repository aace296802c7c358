"""The throughput mode ("fast") against the f64 oracle, through the C ABI.

The throughput mode nominates global and realignment lags from an f32
cross-correlation and rescores the nominees in f64 (xcorr.zig:269-340
refineLag / rescore; engine.zig:800-822 lagTolerance and the cap of 64).
Held here:
* the f32 correlation is within the bound the nomination assumes,
  8·eps_f32·log2(n)·n·(‖env_a‖² + ‖env_b‖²), of the oracle's f64 one;
* its lags equal the oracle's f64 search, nominated or (past the cap, or
  with the cap forced to 0 as pipeline.zig:1514-1517 forces lag_cap) from
  the device's f64 search;
* whole pairs of both modes and every mapping (SVR, speech exponential,
  speech lattice) are within |dMOS| <= 1e-6 of the oracle with identical
  alignments: the global lag, the DP's patch offsets, every realignment
  lag and acceptance.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST_MOS_TOL = 1e-6
F32_EPS = float(np.finfo(np.float32).eps)


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


def _env(vo, x):
    e = np.zeros(x.size)
    vo.lib().vo_upper_envelope(vo.dptr(vo.f64(x)), x.size, vo.dptr(e))
    return e


def _oracle_xcorr(vo, ea, eb):
    n = vo.lib().vo_fft_size_for(max(ea.size, eb.size))
    out = np.zeros(n)
    assert vo.lib().vo_xcorr_re(vo.dptr(ea), ea.size, vo.dptr(eb), eb.size, vo.dptr(out)) == 0
    return out


def _pair(rng, la, lb, shift):
    a = rng.uniform(-1, 1, la) * np.sin(np.linspace(0, 40, la)) ** 2
    b = np.roll(np.resize(a, lb), shift) + 0.05 * rng.standard_normal(lb)
    return a, b


@pytest.mark.parametrize("lg", [14, 15, 16, 18, 20, 21, 22, 23])
def test_xcorr32_within_nomination_bound(native, vo, lg):
    """The f32 correlation (one strided pass up to 2^21 points, two beyond)
    stays inside the error bound the nomination relies on, with margin."""
    rng = np.random.default_rng(500 + lg)
    la = (1 << (lg - 1)) - 5
    lb = la - 1234
    a, b = _pair(rng, la, lb, 321)
    say(f"xcorr32 n=2^{lg}")
    got = native.xcorr32(a, b).astype(np.float64)
    ea, eb = _env(vo, a), _env(vo, b)
    want = _oracle_xcorr(vo, ea, eb)
    n = want.size
    assert got.size == n
    bound = 8.0 * F32_EPS * lg * n * (np.dot(ea, ea) + np.dot(eb, eb))
    err = float(np.max(np.abs(got - want)))
    say(f"  max |f32 - f64| = {err:.4g}, bound {bound:.4g}, ratio {err / bound:.3g}")
    assert err <= 0.25 * bound


def test_xcorr32_out_of_range_sizes(native):
    assert native.xcorr32(np.ones(100), np.ones(90)).size == 0  # n = 256 < 2^14: the f64 search serves it


@pytest.mark.parametrize("case", range(12))
def test_best_lag_fast_equals_oracle(native, vo, synth, case):
    rng = np.random.default_rng(900 + case)
    if case < 4:
        ref, deg = synth.audio_pair(case, duration=float(rng.uniform(1.0, 12.0)), delay=int(rng.integers(0, 3000)),
                                    snr_db=float(rng.uniform(5, 40)))
    elif case < 8:
        ref, deg = synth.speech_pair(case, duration=float(rng.uniform(1.0, 9.0)), delay=int(rng.integers(0, 800)))
    else:
        la, lb = int(rng.integers(9000, 300000)), int(rng.integers(9000, 300000))
        ref, deg = _pair(rng, la, lb, int(rng.integers(-500, 500)))
    want = vo.lib().vo_best_lag(vo.dptr(_env(vo, ref)), ref.size, vo.dptr(_env(vo, deg)), deg.size)
    lag, nominated = native.best_lag_fast(ref, deg)
    say(f"case {case}: lag {lag} (oracle {want}) nominated {nominated}")
    assert lag == want
    assert nominated


def test_forced_cap_takes_f64_search(native, vo, synth):
    """cap 0 sends every row to the device's f64 search (pipeline.zig:
    1514-1517 sets lag_cap to 0 the same way); the lag is unchanged."""
    ref, deg = synth.audio_pair(5, duration=3.0, delay=777)
    want = vo.lib().vo_best_lag(vo.dptr(_env(vo, ref)), ref.size, vo.dptr(_env(vo, deg)), deg.size)
    assert native.best_lag_fast(ref, deg, cap=0) == (want, False)
    assert native.best_lag_fast(ref, deg, cap=64) == (want, True)


def test_flat_correlation_overflows_cap(native, vo):
    """Silence: the envelope is a constant, its correlation a triangle whose
    top ~77 lags lie within the f32 bound, more than the cap of 64, so the
    row falls back; the lag is still the oracle's."""
    a, b = np.zeros(20000), np.zeros(19000)
    want = vo.lib().vo_best_lag(vo.dptr(_env(vo, a)), a.size, vo.dptr(_env(vo, b)), b.size)
    lag, nominated = native.best_lag_fast(a, b)
    say(f"silence: lag {lag} oracle {want} nominated {nominated}")
    assert lag == want and not nominated


def _compare_fast(res, tr, ores, otr):
    assert tr["lag"] == otr["lag"]
    assert tr["ref_indices"] == otr["ref_indices"]
    assert tr["dp_offset"] == otr["dp_offset"]
    assert tr["match_valid"] == otr["match_valid"]
    assert tr["realign_lag"] == otr["realign_lag"]
    assert tr["realign_frames"] == otr["realign_frames"]
    assert tr["realign_accepted"] == otr["realign_accepted"]
    for p, q in zip(res["patch_sims"], ores["patch_sims"]):
        for k in ("ref_patch_start_time", "ref_patch_end_time", "deg_patch_start_time", "deg_patch_end_time"):
            assert p[k] == q[k], k
    assert res["alignment_lag_s"] == ores["alignment_lag_s"]
    assert abs(res["moslqo"] - ores["moslqo"]) <= FAST_MOS_TOL


@pytest.mark.parametrize("mapping,count", [("svr", 16), ("lattice", 16), ("exponential", 8)])
def test_fast_pipeline_matches_oracle(native, vo, synth, mapping, count):
    """Fast mode pair by pair: audio pairs of the per-pair check set (SVR),
    speech pairs with the lattice or the exponential mapping."""
    if mapping == "svr":
        pairs, mode, sr, kw = [synth.plain_audio_pair(s) for s in range(count)], 0, 48000, {}
    else:
        pairs, mode, sr = [synth.speech_pair(100 + s, snr_db=10.0 + s, delay=13 * s) for s in range(count)], 1, 16000
        kw = {"use_lattice_model": mapping == "lattice"}
    ctx = native.Context(mode, precision="fast", **kw)
    orc = vo.Oracle(mode=mode, **kw)
    worst = 0.0
    for i, (r, d) in enumerate(pairs):
        res, tr = ctx.measure_trace(r, d, sr)
        ores, otr = orc.measure(r, d, sr)
        worst = max(worst, abs(res["moslqo"] - ores["moslqo"]))
        _compare_fast(res, tr, ores, otr)
    say(f"fast {mapping}: {count} pairs, max |dMOS| {worst:.3g}")


def test_fast_uneven_mix_batch_matches_oracle(native, vo, synth):
    """A device batch per mode of the uneven 1-60 s mix (noise, delay,
    stretch, dropouts): every pair within 1e-6 of the oracle, identical
    global lags, patch offsets and realignment decisions."""
    mix = [synth.mixed_pair(7000 + s, max_duration=60.0) for s in range(24)]
    for mode in (0, 1):
        sel = [p for p in mix if p[3] == mode]
        if not sel:
            continue
        sr = sel[0][2]
        say(f"uneven mix mode {mode}: {len(sel)} pairs, {sum(p[0].size for p in sel) / sr:.0f} s")
        ctx = native.Context(mode, precision="fast")
        rc, got, st = ctx.measure_batch([p[0] for p in sel], [p[1] for p in sel], sr)
        rc2, want, st2 = vo.Oracle(mode=mode).measure_batch([p[0] for p in sel], [p[1] for p in sel], sr)
        assert st == st2
        for g, w, p in zip(got, want, sel):
            if w is None:
                continue
            assert g["alignment_lag_s"] == w["alignment_lag_s"]
            assert abs(g["moslqo"] - w["moslqo"]) <= FAST_MOS_TOL
            for a, b in zip(g["patch_sims"], w["patch_sims"]):
                for k in ("ref_patch_start_time", "deg_patch_start_time", "deg_patch_end_time"):
                    assert a[k] == b[k], k


def test_fast_cap_zero_pipeline(native, vo, synth):
    """VISQOL_HIP_LAG_CAP=0 at context creation: every global and
    realignment lag comes from the device's f64 search; scores stay within
    1e-6 of the oracle with the same alignments."""
    ref, deg = synth.audio_pair(21, duration=7.0, delay=410, snr_db=17.0)
    old = os.environ.get("VISQOL_HIP_LAG_CAP")
    os.environ["VISQOL_HIP_LAG_CAP"] = "0"
    try:
        ctx = native.Context(native.MODE_AUDIO, precision="fast")
    finally:
        if old is None:
            del os.environ["VISQOL_HIP_LAG_CAP"]
        else:
            os.environ["VISQOL_HIP_LAG_CAP"] = old
    res, tr = ctx.measure_trace(ref, deg, 48000)
    ores, otr = vo.Oracle(mode=0).measure(ref, deg, 48000)
    _compare_fast(res, tr, ores, otr)
