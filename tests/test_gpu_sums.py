"""Reference-order sums on the device, both forms, against their definition.

The exact path's global sums are left-to-right f64 running sums
(envelope.zig:23-53: the plain sum, then the alternating sum of the
mean-centred signal; realign.zig:21-51: the energy of the aligned degraded
signal, from the lag's sample when the lag is negative).  The checker is
that definition itself, a sequential numpy cumsum over the same rounded
terms, and the bar is bit equality for every form visqol_hip_sum offers:
one serial chain, the speculative segmented form (segments run from 2048
candidate starts around an estimate, resolved in order, a segment whose
true start misses its candidates run again serially), and whichever of the
two the scoring path picks.  Cases include the ones that drive the
speculative form off its fast path: wide dynamic range (boundary drift far
beyond the candidates), running sums that stay near zero, exact
cancellation, non-finite values, lengths around the chunk (64) and
segment (16384) sizes, and every start rule of the aligned energy.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = (0, 1, 2)  # serial, speculative, as the scoring path chooses


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def native():
    from akvisqol_amd import native as n

    return n


def definition(kind, x, plain_sum=0.0, lag=0, half=0):
    """The sequential f64 sum of the rounded terms (np.cumsum adds left to
    right)."""
    x = np.asarray(x, np.float64)
    n = x.size
    if kind == 0:
        t = x
    elif kind == 1:
        t = x * x
    elif kind == 2:
        mean = plain_sum / n
        d = x - mean
        t = np.where(np.arange(n) & 1, -d, d)
    else:
        a = -lag if lag < 0 else lag
        start = min(a, n) if (lag < 0 and a <= half) else 0
        t = x[start:] * x[start:]
    return float(np.cumsum(t)[-1]) if t.size else 0.0


def same(a, b):
    if np.isnan(a) or np.isnan(b):
        return bool(np.isnan(a) and np.isnan(b))
    return np.float64(a).view(np.int64) == np.float64(b).view(np.int64)


def check_all_kinds(native, name, x, lags=((0, 0),)):
    x = np.asarray(x, np.float64)
    plain = definition(0, x)
    cases = [(0, 0.0, 0, 0), (1, 0.0, 0, 0)]
    if x.size:
        cases.append((2, plain, 0, 0))
    cases += [(3, 0.0, lag, half) for lag, half in lags]
    for kind, ps, lag, half in cases:
        want = definition(kind, x, ps, lag, half)
        for form in FORMS:
            say(f"case {name} n={x.size} kind={kind} lag={lag} half={half} form={form}")
            got = native.reference_sum(kind, x, ps, lag, half, form)
            assert same(got, want), (name, kind, lag, form, got.hex(), want.hex())


def test_audio_signals(native):
    from akvisqol_amd import synth

    for dur, seed in ((10.0, 3), (60.0, 4)):
        ref, deg = synth.audio_pair(seed, duration=dur, snr_db=20.0, delay=200)
        n = ref.size
        lags = ((-200, n // 2), (0, n // 2), (150, n // 2), (-(n // 2) - 1, n // 2))
        check_all_kinds(native, f"audio ref {dur:g}s", ref, lags)
        check_all_kinds(native, f"audio deg {dur:g}s", deg, lags[:1])


def test_wide_dynamic_range(native):
    # magnitudes over 12 decades: the sum of squares' rounded chain drifts
    # past the candidates at 6 of its 24 boundaries (a host simulation of
    # the boundary choice), so those segments take the serial fallback
    rng = np.random.default_rng(11)
    n = 400_000
    x = rng.standard_normal(n) * 10.0 ** rng.uniform(-6, 6, n)
    check_all_kinds(native, "wide range", x, ((-12345, n // 2),))


def test_every_boundary_misses(native):
    # pairs +B, -B + noise with B over 1e7..1e9: every step rounds at the
    # scale of B while the prefix at chunk edges stays at the noise's, so
    # the plain sum's chain is millions of ulps from every estimate and all
    # 24 boundaries miss: the fallback runs every segment after the first
    rng = np.random.default_rng(16)
    n = 400_000
    b = 10.0 ** rng.uniform(7, 9, n // 2)
    x = np.empty(n)
    x[0::2] = b
    x[1::2] = -b + rng.standard_normal(n // 2)
    check_all_kinds(native, "pairs", x)


def test_near_zero_running_sums(native):
    rng = np.random.default_rng(12)
    n = 300_001
    alt = np.where(np.arange(n) & 1, -1.0, 1.0)
    check_all_kinds(native, "alternating", alt + 1e-9 * rng.standard_normal(n))
    r = rng.standard_normal(150_000)
    check_all_kinds(native, "cancelling", np.concatenate([r, -r[::-1]]))
    check_all_kinds(native, "tiny", 1e-300 * rng.standard_normal(200_000))


def test_quantised_and_constant(native):
    rng = np.random.default_rng(13)
    q = np.round(rng.standard_normal(500_000) * 3000.0) / 32768.0
    check_all_kinds(native, "16-bit", q, ((-777, 250_000),))
    check_all_kinds(native, "constant", np.full(200_000, 0.1))
    check_all_kinds(native, "zeros", np.zeros(100_000))


def test_lengths(native):
    rng = np.random.default_rng(14)
    for n in (0, 1, 2, 63, 64, 65, 16383, 16384, 16385, 65535, 65536, 65537, 100_003, 16384 * 4 + 64 * 3 + 5):
        x = rng.standard_normal(n)
        check_all_kinds(native, "len", x, ((-(n // 3), n // 2), (-n, n)) if n else ((0, 0),))


def test_non_finite(native):
    rng = np.random.default_rng(15)
    n = 200_000
    for name, pos, v in (("inf", 123_456, np.inf), ("-inf", 70_000, -np.inf), ("nan", 150_001, np.nan),
                         ("huge", 99_999, 1e200)):
        x = rng.standard_normal(n)
        x[pos] = v
        with np.errstate(over="ignore", invalid="ignore"):
            check_all_kinds(native, name, x)
