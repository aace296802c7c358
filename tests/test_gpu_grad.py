"""∂moslqo/∂deg on the device (measureBatchGrad, backend.zig:101-110).

As the reference's gradient tests (pipeline.zig:1553-1705):
  * the spectrogram adjoint on one frame against central differences of the
    oracle's filterbank (worst difference below 1e-6 of the largest slope);
  * the full gradient against central differences of the exact forward
    pass along random unit directions, the discrete choices (lags, matches,
    realignments) unchanged by the steps: |fd - g·d| <= 1e-3 |fd| +
    1e-5 |g| (SVR and exponential mappings); the lattice mapping, piecewise
    linear, through the chain rule of its defined slopes instead;
  * the fast precision's gradient against the exact one's;
  * prepared references give the same gradient as plain batches, failed
    pairs zeros.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import vo

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


@pytest.mark.parametrize("mode,sr,nb,high", [(0, 48000, 32, None), (1, 16000, 21, 8000.0)])
def test_frame_adjoint_matches_central_differences(native, mode, sr, nb, high):
    # pipeline.zig:1637-1705: loss = Σ_b g_b · rms_b of one frame
    window = int(np.floor(sr * 0.08 + 0.5))
    rng = np.random.default_rng(3)
    x = 0.1 * rng.standard_normal(window)
    g = rng.standard_normal(nb)

    def loss(y):
        return float(g @ vo.spectrogram(y, sr, nb, high)[:, 0])

    say(f"frame adjoint mode={mode} sr={sr} window={window}")
    got = native.frame_grad(mode, sr, x, g)
    picks = [0, 1, 31, 32, 63, 64, window // 4 + 11, window // 2, window // 2 + 31, window - 929, window - 1]
    worst = 0.0
    for j in picks:
        y = x.copy()
        h = 1e-6
        y[j] = x[j] + h
        up = loss(y)
        y[j] = x[j] - h
        down = loss(y)
        worst = max(worst, abs((up - down) / (2 * h) - got[j]))
    largest = np.abs(got).max()
    assert largest > 0
    assert worst < 1e-6 * largest, (worst, largest)

    # samples outside [lo, hi) read as zero and get zero slope
    lo, hi = 100, window - 300
    got_m = native.frame_grad(mode, sr, x, g, lo, hi)
    assert not got_m[:lo].any() and not got_m[hi:].any()
    xm = x.copy()
    xm[:lo] = 0.0
    xm[hi:] = 0.0
    assert np.array_equal(got_m, native.frame_grad(mode, sr, xm, g, lo, hi))
    full = native.frame_grad(mode, sr, xm, g)
    assert np.array_equal(got_m[lo:hi], full[lo:hi])


def _choices(trace):
    return (trace["lag"], trace["dp_offset"], trace["match_valid"], trace["realign_lag"], trace["realign_accepted"],
            trace["realign_frames"])


# (name, mode, lattice, step, relative bar, context options, degraded lead
# in samples: the degraded signal starts that many samples into its
# delayed copy, so it leads the reference and the global lag prepends
# zeros instead of dropping samples)
CASES = [
    ("audio_svr", 0, True, 1e-5, 1e-3, {}, 0),
    ("audio_svr_leading_deg", 0, True, 1e-5, 1e-3, {}, 400),
    ("audio_no_realignment", 0, True, 1e-5, 1e-3, {"disable_realignment": True}, 0),
    ("audio_no_global_alignment", 0, True, 1e-5, 1e-3, {"disable_global_alignment": True}, 0),
    ("speech_exponential", 1, False, 1e-5, 1e-3, {}, 0),
    ("speech_exponential_unscaled", 1, False, 1e-5, 1e-3, {"use_unscaled_speech_mos_mapping": True}, 0),
]


@pytest.mark.parametrize("name,mode,lattice,eps,rel,opts,lead", CASES, ids=[c[0] for c in CASES])
def test_gradient_matches_central_differences(native, synth, name, mode, lattice, eps, rel, opts, lead):
    if mode == 0:
        ref, deg = synth.audio_pair(21, duration=3.0)
        sr = 48000
    else:
        ref, deg = synth.speech_pair(22, duration=3.0)
        sr = 16000
    deg = deg[lead:].copy()
    ctx = native.Context(mode, use_lattice_model=lattice, precision="exact", **opts)
    say(f"gradient {name}: {deg.size} samples at {sr}")
    rc, res, st, grads = ctx.measure_batch_grad([ref], [deg], sr)
    assert rc == 0 and st == [0]
    g = grads[0]
    base_res, base_tr = ctx.measure_trace(ref, deg, sr)
    assert res[0]["moslqo"] == base_res["moslqo"]
    assert 1.0 < base_res["moslqo"] < 5.0
    if lead:
        assert base_tr["lag"] > 0  # zeros prepended to the aligned signal
    g_norm = np.linalg.norm(g)
    assert g_norm > 0 and np.all(np.isfinite(g))
    rng = np.random.default_rng(0x5EED)
    worst = 0.0
    for _ in range(3):
        d = rng.standard_normal(deg.size)
        d /= np.linalg.norm(d)
        want = float(d @ g)
        mos = []
        for e in (eps, -eps):
            r, tr = ctx.measure_trace(ref, deg + e * d, sr)
            assert _choices(tr) == _choices(base_tr)
            mos.append(r["moslqo"])
        fd = (mos[0] - mos[1]) / (2 * eps)
        worst = max(worst, abs(fd - want) / max(abs(fd), 1e-300))
        assert abs(fd - want) <= rel * abs(fd) + 1e-5 * g_norm, (fd, want, g_norm)
    say(f"gradient {name}: worst relative difference from central differences {worst:.1e}")


FEATURES = ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy")


def test_lattice_gradient_is_the_chain_rule_of_its_slopes(native, synth):
    """The lattice's slopes are defined by central differences of the model
    with steps of 1e-3·max(1, |v|) per feature (visqol.zig:132-146): it is
    piecewise linear, so a step of the whole pipeline can cross different
    kinks than those steps and the two slopes need not agree.  What pins the
    gradient is the chain rule: along a unit direction d it must equal
    Σ (the model's slopes, from the oracle's lattice at the pair's features)
    · (the features' own directional derivatives, central differences of the
    exact forward pass, which is smooth in the signal with the choices
    fixed).  The pair is a clean one: at 20 dB SNR its score sits on the
    model's floor of 1, where the slope is zero."""
    import ctypes

    from oracle.vo import dptr, f64

    ref, deg = synth.speech_pair(22, duration=3.0, snr_db=50.0)
    sr = 16000
    ctx = native.Context(1, use_lattice_model=True, precision="exact")
    say(f"lattice gradient: {deg.size} samples")
    rc, res, st, grads = ctx.measure_batch_grad([ref], [deg], sr)
    assert rc == 0 and st == [0]
    g = grads[0]
    base, base_tr = ctx.measure_trace(ref, deg, sr)
    assert 1.0 < base["moslqo"] < 5.0 and base["moslqo"] == res[0]["moslqo"]
    feats = [f64(base[k]).copy() for k in FEATURES]
    nb = feats[0].size
    model = (Path(native.__file__).parent / "models" / "lattice_speech.tflite").read_bytes()

    def lattice(x):
        y = ctypes.c_double()
        assert vo.lib().vo_lattice_predict(model, len(model), *[dptr(v) for v in x], nb, ctypes.byref(y)) == 0
        return y.value

    assert lattice(feats) == base["moslqo"]
    slopes = np.zeros((4, nb))
    for k in range(4):
        for b in range(nb):
            v = feats[k][b]
            h = 1e-3 * max(1.0, abs(v))
            x = [f.copy() for f in feats]
            x[k][b] = v + h
            up = lattice(x)
            x[k][b] = v - h
            slopes[k, b] = (up - lattice(x)) / (2 * h)
    g_norm = np.linalg.norm(g)
    assert g_norm > 0
    rng = np.random.default_rng(0x5EED)
    eps = 1e-5
    for _ in range(3):
        d = rng.standard_normal(deg.size)
        d /= np.linalg.norm(d)
        fs = []
        for e in (eps, -eps):
            r, tr = ctx.measure_trace(ref, deg + e * d, sr)
            assert _choices(tr) == _choices(base_tr)
            fs.append(np.array([r[k] for k in FEATURES]))
        dfeat = (fs[0] - fs[1]) / (2 * eps)
        want = float((slopes * dfeat).sum())
        got = float(d @ g)
        assert abs(got - want) <= 1e-3 * abs(want) + 1e-5 * g_norm, (got, want, g_norm)


@pytest.mark.parametrize("mode,lattice", [(0, True), (1, False)])
def test_fast_gradient_tracks_exact(native, synth, mode, lattice):
    if mode == 0:
        ref, deg = synth.audio_pair(31, duration=3.0)
        sr = 48000
    else:
        ref, deg = synth.speech_pair(32, duration=3.0)
        sr = 16000
    say(f"fast vs exact gradient mode={mode}")
    ge = native.Context(mode, use_lattice_model=lattice, precision="exact").measure_batch_grad([ref], [deg], sr)[3][0]
    gf = native.Context(mode, use_lattice_model=lattice, precision="fast").measure_batch_grad([ref], [deg], sr)[3][0]
    assert np.linalg.norm(ge) > 0
    assert np.linalg.norm(gf - ge) <= 1e-6 * np.linalg.norm(ge)


def test_batch_prepared_and_failures(native, synth):
    pairs = [synth.audio_pair(s, duration=2.0 + 0.5 * s) for s in (41, 42, 43)]
    refs = [p[0] for p in pairs]
    degs = [p[1] for p in pairs]
    ctx = native.Context(0, precision="fast")
    say("gradient batch of 3 audio pairs + one too short")
    short = np.zeros(100)
    rc, res, st, grads = ctx.measure_batch_grad(refs + [short], degs + [short], 48000)
    assert st[:3] == [0, 0, 0] and st[3] != 0 and res[3] is None
    assert rc == st[3]
    assert not grads[3].any()
    for k in range(3):
        one = ctx.measure_batch_grad([refs[k]], [degs[k]], 48000)[3][0]
        assert np.array_equal(grads[k], one)
        assert res[k]["moslqo"] == ctx.measure(refs[k], degs[k], 48000)["moslqo"]
    prep = ctx.prepare(refs, 48000)
    rc, pres, pst, pgrads = ctx.measure_prepared_grad(prep, degs)
    assert rc == 0 and pst == [0, 0, 0]
    for k in range(3):
        assert pres[k]["moslqo"] == res[k]["moslqo"]
        assert np.array_equal(pgrads[k], grads[k])


def test_speech_prepared_gradient_equals_batch(native, synth):
    # prepared speech references carry their VAD patch indices
    pairs = [synth.speech_pair(s, duration=3.0, snr_db=35.0) for s in (51, 52)]
    refs = [p[0] for p in pairs]
    degs = [p[1] for p in pairs]
    ctx = native.Context(1, use_lattice_model=False, precision="fast")
    say("speech gradients: 2 pairs, plain and prepared")
    rc, res, st, grads = ctx.measure_batch_grad(refs, degs, 16000)
    assert rc == 0 and st == [0, 0]
    prep = ctx.prepare(refs, 16000)
    rc2, pres, pst, pgrads = ctx.measure_prepared_grad(prep, degs)
    assert rc2 == 0 and pst == [0, 0]
    for k in range(2):
        assert pres[k]["moslqo"] == res[k]["moslqo"]
        assert np.linalg.norm(grads[k]) > 0
        assert np.array_equal(pgrads[k], grads[k])
