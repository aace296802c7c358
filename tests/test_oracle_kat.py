"""Pins the CPU oracle against the reference's own known-answer tests and
against what the mathematics fixes (closed forms, brute force, invariants).

Every constant below is transcribed from a reference test block; the file
and line are cited next to each test.
"""
import ctypes
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import vo
from oracle.vo import dptr, f64, sptr

ROOT = Path(__file__).resolve().parents[1]
MODELS = ROOT / "akvisqol_amd" / "models"
L = vo.lib()


# ---------------------------------------------------------------- log10
def test_log10_exact_powers_and_ulp():
    # fdlibm's log10 returns N exactly for 10**N (a documented property of
    # the published algorithm); glibc's log10 is within 1 ulp elsewhere.
    for n in range(0, 23):
        assert L.vo_log10(10.0 ** n) == float(n)
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(1e-300, 1e300, 2000), np.exp(rng.uniform(-700, 700, 20000)),
                         rng.uniform(0.5, 2.0, 5000), [2.2250738585072014e-308, 5e-324, 1e-310]])
    got = np.array([L.vo_log10(float(x)) for x in xs])
    want = np.log10(xs)
    ulp = np.spacing(np.abs(want))
    assert np.all(np.abs(got - want) <= ulp)
    assert L.vo_log10(0.0) == -math.inf
    assert math.isnan(L.vo_log10(-1.0))
    assert L.vo_log10(math.inf) == math.inf


def test_log10_constants_bit_patterns():
    # The hex-encoded fdlibm constants agree with their published decimal forms.
    import struct

    def h(b):
        return struct.unpack("<d", struct.pack("<Q", b))[0]

    assert h(0x3fe5555555555593) == 6.666666666666735130e-01
    assert h(0x3fd999999997fa04) == 3.999999999940941908e-01
    assert h(0x3fd2492494229359) == 2.857142874366239149e-01
    assert h(0x3fcc71c51d8e78af) == 2.222219843214978396e-01
    assert h(0x3fc7466496cb03de) == 1.818357216161805012e-01
    assert h(0x3fc39a09d078c69f) == 1.531383769920937332e-01
    assert h(0x3fc2f112df3e5244) == 1.479819860511658591e-01
    assert abs(h(0x3fdbcb7b15200000) + h(0x3dbb9438ca9aadd5) - 1 / math.log(10)) < 1.2e-16
    assert abs(h(0x3fd34413509f6000) + h(0x3d59fef311f12b36) - math.log10(2)) < 1.2e-16


# ---------------------------------------------------------------- ERB / filters
def test_center_freqs():  # erb.zig:64-72, gammatone.zig:469-478
    cf = np.zeros(32)
    L.vo_center_freqs_desc(32, 50.0, 24000.0, dptr(cf))
    assert abs(cf[0] - 20844.7852) < 1e-3
    assert abs(cf[31] - 50.0) < 1e-6
    assert np.all(np.diff(cf) < 0)
    _, asc = vo.filters(48000, 32, 50.0, 24000.0)
    assert np.all(np.diff(asc) > 0) and abs(asc[0] - 50.0) < 1e-6


def test_filters_gain_finite_positive():  # gammatone.zig:480-489
    bands, _ = vo.filters(48000, 32, 50.0, 24000.0)
    for b in bands:
        assert math.isfinite(b.gain) and b.gain > 0 and math.isfinite(b.b1) and math.isfinite(b.b2)


def test_lowest_band_reference_output():  # gammatone.zig:491-505
    bands, _ = vo.filters(48000, 32, 50.0, 24000.0)
    x = f64([0.2, 0.4, 0.6, 0.8, 0.9, 0.1, 0.3, 0.5, 0.7, 0.9])
    out = np.zeros(10)
    L.vo_filter_band(ctypes.byref(bands[0]), dptr(x), 10, dptr(out))
    assert abs(out[0] - 1.028e-10) < 1e-13
    assert abs(out[1] - 6.15143e-10) < 1e-13
    assert abs(out[2] - 2.14718e-09) < 1e-13


def test_gammatone_is_the_lti_cascade():
    # Closed form: the 4-stage cascade is LTI, so its response to any input
    # equals the convolution with its impulse response; check the
    # recurrence (float64) against an exact-rational-free high-precision
    # brute force (Python floats of the summed convolution at 1e-12 rel).
    bands, _ = vo.filters(16000, 21, 50.0, 8000.0)
    rng = np.random.default_rng(3)
    x = f64(rng.uniform(-1, 1, 300))
    for bi in (0, 7, 20):
        imp = np.zeros(300)
        imp[0] = 1.0
        h = np.zeros(300)
        L.vo_filter_band(ctypes.byref(bands[bi]), dptr(f64(imp)), 300, dptr(h))
        y = np.zeros(300)
        L.vo_filter_band(ctypes.byref(bands[bi]), dptr(x), 300, dptr(y))
        conv = np.convolve(x, h)[:300]
        assert np.max(np.abs(conv - y)) <= 1e-9 * np.max(np.abs(y))
        # Transfer function: 4 biquads share the denominator 1 + b1 z^-1 + b2 z^-2.
        b = bands[bi]
        from scipy.signal import lfilter

        gi = 1.0 / b.gain
        z = x.copy()
        for n0, n1 in ((b.a0 * gi, b.a11 * gi), (b.a0, b.a12), (b.a0, b.a13), (b.a0, b.a14)):
            z = lfilter([n0, n1], [1.0, b.b1, b.b2], z)
        assert np.max(np.abs(z - y)) <= 1e-12 * np.max(np.abs(y))


def test_band_rms_matches_filter_band():
    bands, _ = vo.filters(48000, 32, 50.0, 24000.0)
    rng = np.random.default_rng(9)
    x = f64(rng.uniform(-1, 1, 384))
    for bi in (0, 15, 31):
        y = np.zeros(384)
        L.vo_filter_band(ctypes.byref(bands[bi]), dptr(x), 384, dptr(y))
        rms = L.vo_band_rms_fresh(ctypes.byref(bands[bi]), dptr(x), 384)
        assert abs(rms - math.sqrt(np.sum(y * y) / 384)) <= 1e-14 * rms


def test_hann():  # gammatone.zig:507-516
    w = vo.hann(16)
    assert abs(w[0]) < 1e-12 and abs(w[15]) < 1e-12
    for i in range(8):
        assert abs(w[i] - w[15 - i]) < 1e-12


def test_spectrogram_shape_finite():  # gammatone.zig:546-577
    rng = np.random.default_rng(0xABC123)
    sig = rng.uniform(-1, 1, 3840 * 4 + 960 * 7)
    s = vo.spectrogram(sig, 48000, 32)
    assert s.shape == (32, 1 + (sig.size - 3840) // 960)
    assert np.all(np.isfinite(s)) and np.all(s >= 0)


def test_spectrogram_lanes_equal_scalar():  # gammatone.zig:579-610
    bands, _ = vo.filters(48000, 32, 50.0, 24000.0)
    W, hop, nf = 384, 96, 3 * 8 + 1
    rng = np.random.default_rng(0x9A77)
    sig = f64(rng.uniform(-1, 1, W + (nf - 1) * hop))
    h = vo.hann(W)
    out = np.zeros(32 * nf)
    assert L.vo_spectrogram(bands, 32, dptr(sig), sig.size, W, hop, dptr(h), dptr(out)) == 0
    out = out.reshape(32, nf)
    for f in range(nf):
        frame = f64(sig[f * hop: f * hop + W] * h)
        for b in range(32):
            assert L.vo_band_rms_fresh(ctypes.byref(bands[b]), dptr(frame), W) == out[b, f]


# ---------------------------------------------------------------- dB prep
def test_db_reference_values():  # spectrogram.zig:95-103
    d = f64([10.21, -4.63, 0.54, 87.98, 0.065, 0.0, -53.2, 8.7, 0.0, -2.76])
    exp = [10.0903, 6.6558, -2.6761, 19.4438, -11.8709, -156.5356, 17.2591, 9.3952, -156.5356, 4.4091]
    L.vo_db_floor(dptr(d), 10, -1000.0)
    assert np.max(np.abs(d - exp)) < 1e-3


def test_db_clamps():  # spectrogram.zig:105-111
    d = f64([1.0, 0.001, 0.000001])
    L.vo_db_floor(dptr(d), 3, -45.0)
    assert abs(d[0]) < 1e-9 and abs(d[1] + 30) < 1e-9 and abs(d[2] + 45) < 1e-9


def test_joint_min():  # spectrogram.zig:113-123
    a, b = f64([5.0, -3.0, 2.0]), f64([4.0, -7.0, 1.0])
    L.vo_subtract_joint_min(dptr(a), 3, dptr(b), 3)
    assert list(a) == [12.0, 4.0, 9.0] and list(b) == [11.0, 0.0, 8.0]


def test_raise_floor_per_frame():  # spectrogram.zig:125-138
    r, d = f64([10.0, -50.0, 5.0, 20.0]), f64([3.0, -200.0, -100.0, -1.0])
    L.vo_raise_floor_per_frame(dptr(r), dptr(d), 2, 2, 2, 45.0)
    assert list(r) == [10.0, -25.0, 5.0, 20.0] and list(d) == [3.0, -25.0, -35.0, -1.0]


# ---------------------------------------------------------------- envelope
def test_envelope_kats():  # envelope.zig:55-69
    z = np.zeros(8)
    out = np.zeros(8)
    L.vo_upper_envelope(dptr(z), 8, dptr(out))
    assert np.all(np.abs(out - 1e-6) < 1e-12)
    c = np.full(16, 0.7)
    out = np.zeros(16)
    L.vo_upper_envelope(dptr(c), 16, dptr(out))
    assert np.all(np.abs(out - (0.7 + 1e-6)) < 1e-12)


@pytest.mark.parametrize("n", [5, 31, 32, 100, 1000, 4097])
def test_envelope_closed_form_equals_fft_definition(n):
    # envelope.zig:6-10: Re(IFFT(X·h)) of the mean-centred, zero-padded
    # signal, h = [1, 2…2, 0 (Nyquist), 0…], reduces to the closed form.
    rng = np.random.default_rng(n)
    x = f64(rng.standard_normal(n) + 0.3)
    out = np.zeros(n)
    L.vo_upper_envelope(dptr(x), n, dptr(out))
    N = max(1 << (n - 1).bit_length(), 32)
    mean = x.mean()
    X = np.fft.fft(np.concatenate([x - mean, np.zeros(N - n)]))
    h = np.zeros(N)
    h[0] = 1.0
    h[1: N // 2] = 2.0
    re = np.fft.ifft(X * h).real[:n]
    want = np.abs(2.0 * re - 1e-6) + mean
    assert np.max(np.abs(out - want)) < 1e-9 * max(1.0, np.max(np.abs(x)))


# ---------------------------------------------------------------- FFT / lag
def _cplx(z):
    return f64(np.stack([z.real, z.imag], axis=1).ravel())


def test_fft_dc_impulse():  # fft.zig:411-426
    z = _cplx(np.ones(16, complex))
    L.vo_fft_natural(dptr(z), 16, 0)
    c = z[0::2] + 1j * z[1::2]
    assert abs(c[0] - 16) < 1e-10 and np.max(np.abs(c[1:])) < 1e-10


def test_fft_matches_numpy_and_round_trips():  # fft.zig:428-531
    rng = np.random.default_rng(7)
    for n in (2, 4, 8, 32, 1 << 13, 1 << 15, 1 << 16):
        x = rng.uniform(-0.5, 0.5, n) + 1j * rng.uniform(-0.5, 0.5, n)
        z = _cplx(x)
        L.vo_fft_dif_forward(dptr(z), n)
        y = z[0::2] + 1j * z[1::2]
        want = np.fft.fft(x)
        lg = n.bit_length() - 1
        br = np.array([int(format(i, f"0{lg}b")[::-1], 2) if lg else 0 for i in range(n)])
        assert np.max(np.abs(y - want[br])) <= 1e-9 * n
        L.vo_fft_dit_inverse(dptr(z), n)
        back = (z[0::2] + 1j * z[1::2]) / n
        assert np.max(np.abs(back - x)) < 1e-12


def test_fft_size_for():  # xcorr.zig:342-347
    assert [L.vo_fft_size_for(k) for k in (14, 15, 16, 17)] == [32, 32, 32, 64]


REF14 = [2.0, 2.0, 1.0, 0.1, -3.0, 0.1, 1.0, 2.0, 2.0, 6.0, 8.0, 6.0, 2.0, 2.0]


@pytest.mark.parametrize("deg,lag", [  # xcorr.zig:349-381
    ([1.2, 0.1, -3.3, 0.1, 1.1, 2.2, 2.1, 7.1, 8.3, 6.8, 2.4, 2.2, 2.2, 2.1], 2),
    ([1.2, 0.1, -3.3, 0.1, 1.1, 2.2, 2.1, 7.1, 8.3, 6.8, 2.4, 2.2, 2.2, 2.1, 2.0], 2),
    ([1.2, 0.1, -3.3, 0.1, 1.1, 2.2, 2.1, 7.1, 8.3, 6.8, 2.4, 2.2, 2.2], 2),
    ([2.0, 2.0, 2.0, 2.0, 1.0, 0.1, -3.0, 0.1, 1.0, 2.0, 2.0, 6.0, 8.0, 6.0], -2),
])
def test_best_lag_toys(deg, lag):
    a, b = f64(REF14), f64(deg)
    assert L.vo_best_lag(dptr(a), a.size, dptr(b), b.size) == lag


def test_best_lag_identical():
    s = f64([0.5, -0.3, 0.8, 0.1, -0.2, 0.9, 0.0, -0.5])
    assert L.vo_best_lag(dptr(s), 8, dptr(s), 8) == 0


def test_best_lag_cached_equals_three_fft_and_brute_force():  # xcorr.zig:383-401
    rng = np.random.default_rng(1)
    for it in range(20):
        l1, l2 = 50 + int(rng.integers(0, 3000)), 50 + int(rng.integers(0, 3000))
        a = rng.uniform(-1, 1, l1)
        shift = int(rng.integers(0, 40))
        b = np.where((np.arange(l2) >= shift) & (np.arange(l2) - shift < l1),
                     np.concatenate([np.zeros(shift), a])[:l2] if l2 <= l1 + shift else
                     np.concatenate([np.zeros(shift), a, np.zeros(l2)])[:l2], 0.0)
        noise = np.where((np.arange(l2) >= shift) & (np.arange(l2) - shift < l1),
                         0.1 * (rng.uniform(size=l2) - 0.5), 0.05 * rng.uniform(size=l2))
        b = f64(b + noise)
        a = f64(a)
        fast = L.vo_best_lag(dptr(a), l1, dptr(b), l2)
        assert fast == L.vo_best_lag_reference(dptr(a), l1, dptr(b), l2)
        if it < 6:
            assert fast == L.vo_best_lag_direct(dptr(a), l1, dptr(b), l2)


# ---------------------------------------------------------------- slicing / trim
def test_slice_layouts():  # realign.zig:462-482
    pad, base, n = (ctypes.c_size_t() for _ in range(3))
    L.vo_slice_layout(10, 10, 0.2, 0.6, pad, base, n)
    assert (pad.value, base.value, n.value) == (0, 2, 4)
    L.vo_slice_layout(48000, 16000, 0.5, 2.5, pad, base, n)
    assert pad.value + n.value == 32000 and base.value == 8000


def test_lag_samples_back():  # realign.zig:528
    assert L.vo_lag_samples_back(27, 48000) == 26


def test_trim_geometry_matches_truncate_aligned():  # realign.zig:496-529
    # truncateAligned restated from its definition: applyLag, then the
    # three length cases.
    rng = np.random.default_rng(5)
    ln, ro, do, ls = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int64(), ctypes.c_double()
    for it in range(3000):
        sr = (16000, 44100, 48000)[it % 3]
        rl, dl = int(rng.integers(0, 401)), int(rng.integers(0, 401))
        lag = 27 if it % 4 == 0 else -55 if it % 4 == 1 else int(rng.integers(-rl, rl + 1))
        ref = np.arange(1, rl + 1, dtype=float)
        deg = -np.arange(1, dl + 1, dtype=float)
        if lag == 0 or abs(lag) > rl // 2:
            al, lag_s = deg.copy(), 0.0
        elif lag < 0:
            al, lag_s = deg[min(-lag, dl):].copy(), lag / sr
        else:
            al, lag_s = np.concatenate([np.zeros(lag), deg]), lag / sr
        if rl > al.size:
            r_out, d_out = ref[: al.size], al
        elif rl < al.size:
            k = int(lag_s * sr)
            if k > 0:
                r_out = ref[k:]
                d_out = al[k: min(k + r_out.size, al.size)]
            else:
                r_out, d_out = ref, al[:rl]
        else:
            r_out, d_out = ref, al
        m = min(r_out.size, d_out.size)
        r_out, d_out = r_out[:m], d_out[:m]
        L.vo_trim_geometry(rl, dl, lag, sr, ln, ro, do, ls)
        assert ln.value == m and ls.value == lag_s
        assert np.array_equal(r_out, ref[ro.value: ro.value + m])
        j = do.value + np.arange(m)
        want = np.where((j >= 0) & (j < dl), deg[np.clip(j, 0, max(dl - 1, 0))] if dl else 0.0, 0.0)
        assert np.array_equal(d_out, want)


# ---------------------------------------------------------------- patches / DP
def test_ref_patch_indices():  # dp.zig:389-396
    out = np.zeros(16, np.uintp)
    n = L.vo_ref_patch_indices(200, 30, sptr(out), 16)
    assert list(out[:n]) == [14, 44, 74, 104, 134, 164]


def test_calc_max_num_patches():  # dp.zig:398-402
    idx = np.array([0, 15, 30, 45, 60], np.uintp)
    assert L.vo_calc_max_num_patches(sptr(idx), 5, 45, 30) == 5
    assert L.vo_calc_max_num_patches(sptr(idx), 5, 44, 30) == 4


def test_dp_toy():  # dp.zig:404-450
    ref_rows = [[1.0, 1.0, 1.0, 2.0, 2.0, 2.0, 2.0, 2.0, 3.0, 3.0],
                [0.0, 1.0, 0.0, 2.0, 1.0, 1.0, 2.0, 3.0, 1.0, 2.0],
                [0.0, 1.0, 0.0, 2.0, 1.0, 1.0, 2.0, 3.0, 1.0, 2.0]]
    deg = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 1.0, 0.0, 0.0, 0.0, 0.0, 2.0, 2.0, 3.0, 2.0,
           0.0, 0.0, 0.0, 0.0, 3.0, 3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
           0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 2.0, 1.0, 2.0, 3.0, 0.0, 0.0, 0.0, 0.0,
           1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0,
           0.0, 0.0, 0.0, 0.0, 0.0, 2.0, 1.0, 2.0, 3.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.0, 0.0, 0.0,
           0.0, 0.0, 0.0, 0.0, 0.0]
    ref = f64(np.array(ref_rows).ravel())
    d = f64(deg)
    idx = np.arange(10, dtype=np.uintp)
    starts = np.zeros(10)
    n = L.vo_find_optimal_patches(dptr(ref), sptr(idx), 10, dptr(d), 3, 10, 30, 1, 1.0, 8, dptr(starts))
    assert n == 10
    assert abs(starts[3] - 0.0) < 1e-9 and abs(starts[4] - 7.0) < 1e-9 and abs(starts[5] - 8.0) < 1e-9


def _dp_brute_value(sims, lo, hi):
    # The DP's objective restated as a max over explicit paths: row p takes
    # offset s_p in its window; either s_p > t (t = previous row's offset,
    # t in the previous window, gain sims[p][s_p]) or s_p == t (no move,
    # gain 0).  Cells a row never computed read 0 (dp.zig:224, memset).
    P, cols = sims.shape
    best = {}
    for s in range(lo[0], hi[0] + 1):
        best[(0, s)] = sims[0, s]
    for p in range(1, P):
        for s in range(lo[p], hi[p] + 1):
            cands = []
            for t in range(lo[p - 1], s):
                cands.append(best.get((p - 1, t), 0.0) + sims[p, s])
            cands.append(best.get((p - 1, s), 0.0))
            best[(p, s)] = max(cands)
    return max(best[(P - 1, s)] for s in range(lo[P - 1], hi[P - 1] + 1))


def test_dp_value_brute_force():
    rng = np.random.default_rng(11)
    for _ in range(200):
        P, cols = int(rng.integers(1, 5)), int(rng.integers(3, 13))
        sims = rng.uniform(0, 1, (P, cols)).round(1)  # ties on purpose
        lo = np.sort(rng.integers(0, cols // 2 + 1, P)).astype(np.uintp)
        hi = np.array([int(rng.integers(l, cols)) for l in lo], np.uintp)
        offs = np.zeros(P, np.uintp)
        valid = np.zeros(P, np.uint8)
        S = f64(sims.ravel())
        got = L.vo_dp_search(dptr(S), P, cols, sptr(lo), sptr(hi), sptr(offs),
                             valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        want = _dp_brute_value(sims, [int(x) for x in lo], [int(x) for x in hi])
        assert got == want
        assert lo[-1] <= offs[-1] <= hi[-1]


# ---------------------------------------------------------------- NSIM
def test_conv2d_reference():  # nsim.zig:513-548
    cols = [[40.0392, 43.3409, 39.5270, 41.1731, 41.3591], [42.6852, 45.2083, 45.7769, 39.9689, 43.6190],
            [41.0119, 40.4244, 41.5932, 43.6027, 42.6204], [43.0624, 42.2610, 42.4725, 43.4258, 42.9079]]
    exp_cols = [[40.6634, 42.8407, 40.6395, 41.0129, 41.5407], [42.4677, 44.2760, 44.2031, 41.2263, 42.9752],
                [41.3784, 41.2656, 42.1388, 43.0366, 42.8042], [42.7613, 42.1817, 42.4590, 43.2709, 42.9377]]
    x = f64(np.array(cols).T.ravel())
    out = np.zeros(20)
    L.vo_conv2d3x3(dptr(x), dptr(out), 5, 4)
    assert np.max(np.abs(out - np.array(exp_cols).T.ravel())) < 1e-3


def test_conv2d_brute_force():
    # True 3x3 convolution with replicate padding, by explicit padding.
    from scipy.signal import convolve2d

    w = np.array([[0.0113033910173052, 0.0838251475442633, 0.0113033910173052],
                  [0.0838251475442633, 0.619485845753726, 0.0838251475442633],
                  [0.0113033910173052, 0.0838251475442633, 0.0113033910173052]])
    rng = np.random.default_rng(2)
    for nr, nc in ((1, 1), (1, 5), (4, 1), (6, 7), (32, 30)):
        x = rng.uniform(0, 40, (nr, nc))
        out = np.zeros(nr * nc)
        L.vo_conv2d3x3(dptr(f64(x.ravel())), dptr(out), nr, nc)
        want = convolve2d(np.pad(x, 1, mode="edge"), w, mode="valid")
        assert np.max(np.abs(out.reshape(nr, nc) - want)) < 1e-12


def test_patch_similarity_reference():  # nsim.zig:550-568
    means, sds, en, mm = np.zeros(3), np.zeros(3), np.zeros(3), ctypes.c_double()
    L.vo_patch_similarity(dptr(f64([1.0, 0.0, 0.0])), dptr(f64([0.0, 0.0, 0.0])), 3, 1, dptr(means), dptr(sds),
                          dptr(en), ctypes.byref(mm))
    assert abs(means[0] - 0.000125225) < 1e-4 and abs(means[1] - 0.00875062) < 1e-4 and abs(means[2] - 1.0) < 1e-4


def test_identical_patches_give_one():  # nsim.zig:570-581
    sig = f64(np.sin(np.arange(20) * 0.3) + 1.5)
    assert abs(L.vo_similarity_scalar(dptr(sig), dptr(sig), 4, 5) - 1.0) < 1e-9


def test_similarity_table_bit_equal_scalar():  # nsim.zig:606-637
    nr, nc, nf = 6, 5, 17
    rng = np.random.default_rng(0x5EED)
    spec = f64(rng.uniform(0, 40, nr * nf))
    ref = f64(rng.uniform(0, 40, nr * nc))
    width = nf + nc
    lo = np.array([0], np.uintp)
    hi = np.array([width - nc], np.uintp)
    sims = np.zeros(nf)
    # the table writes sims[p*nf + off]; extend to width by a second buffer
    sims = np.zeros(width)
    L.vo_similarity_table(dptr(ref), nc, dptr(spec), nf, nr, nc, sptr(np.array([0], np.uintp)), sptr(lo), sptr(hi), 1,
                          dptr(sims))
    s2 = spec.reshape(nr, nf)
    for off in range(width - nc + 1):
        deg = np.zeros((nr, nc))
        for c in range(nc):
            if off + c < nf:
                deg[:, c] = s2[:, off + c]
        want = L.vo_similarity_scalar(dptr(ref), dptr(f64(deg.ravel())), nr, nc)
        assert sims[off] == want


# ---------------------------------------------------------------- VAD / SPL
def test_vad_silence_drops_patches():  # vad.zig:83-95
    sig = np.zeros(64000)
    n = np.arange(16000, 32000)
    sig[16000:32000] = np.sin(n * 0.1)
    nf = 1 + (sig.size - 1280) // 320
    out = np.zeros(64, np.uintp)
    k = L.vo_vad_patch_indices(dptr(f64(sig)), sig.size, nf, 20, 1280, 0.25, sptr(out), 64)
    assert k >= 2 and k < (nf - 9) // 20
    assert out[0] == 9


def test_spl():  # audio.zig:201-212
    assert abs(L.vo_spl(dptr(np.ones(16)), 16) - 93.9794) < 1e-3
    ref = f64([0.5, -0.5] * 4)
    deg = f64([0.25, -0.25] * 4)
    L.vo_scale_to_match_spl(dptr(ref), 8, dptr(deg), 8)
    assert abs(L.vo_spl(dptr(ref), 8) - L.vo_spl(dptr(deg), 8)) < 1e-9 and abs(deg[0] - 0.5) < 1e-9


# ---------------------------------------------------------------- mapping
def _svr(text, obs):
    y = ctypes.c_double()
    b = text.encode()
    rc = L.vo_svr_predict(b, len(b), dptr(f64(obs)), len(obs), ctypes.byref(y))
    return rc, y.value


def test_svr_manual_rbf():  # svr.zig:190-201
    txt = ("svm_type nu_svr\nkernel_type rbf\ngamma 1.0\nnr_class 2\ntotal_sv 2\nrho 0.0\nSV\n"
           "1.0 1:0.0 2:0.0\n1.0 1:1.0 2:1.0\n")
    rc, y = _svr(txt, [0.5, 0.5])
    assert rc == 0 and abs(y - 1.2130613194) < 5e-8


def test_svr_synthetic_model():  # svr.zig:167-188: 2 SVs x 3 features, evaluated by hand
    txt = ("svm_type nu_svr\nkernel_type rbf\ngamma 0.5\nnr_class 2\ntotal_sv 2\nrho 0.1\nSV\n"
           "0.5 1:1.0 2:2.0 3:3.0\n-0.25 1:0.5 2:0.0 3:1.0\n")
    x = [0.1, 0.2, 0.3]
    want = 0.5 * math.exp(-0.5 * ((1 - .1) ** 2 + (2 - .2) ** 2 + (3 - .3) ** 2)) - 0.25 * math.exp(
        -0.5 * ((.5 - .1) ** 2 + (0 - .2) ** 2 + (1 - .3) ** 2)) - 0.1
    rc, y = _svr(txt, x)
    assert rc == 0 and abs(y - want) < 1e-15


def test_svr_default_model_known_fvnsim():  # svr.zig:203-219
    obs = [0.853862, 0.680331, 0.535649, 0.639760, 0.029999, 0.058591, 0.077462, 0.012432, 0.192035, 0.389230,
           0.479403, 0.419914, 0.521414, 0.858340, 0.884218, 0.864682, 0.868514, 0.845271, 0.850559, 0.877882,
           0.903985, 0.887572, 0.920558, 0.920375, 0.954934, 0.945048, 0.952716, 0.986600, 0.987345, 0.936462,
           0.856010, 0.829761]
    rc, y = _svr((MODELS / "libsvm_nu_svr_model.txt").read_text(), obs)
    assert rc == 0 and abs(y - 4.30533) < 1e-3


def test_svr_rejects_bad_models():
    assert _svr("svm_type c_svc\nkernel_type rbf\nSV\n1 1:1\n", [1.0])[0] == -4
    assert _svr("svm_type nu_svr\nkernel_type linear\nSV\n1 1:1\n", [1.0])[0] == -4
    assert _svr("svm_type nu_svr\nkernel_type rbf\n", [1.0])[0] == -4


def test_lattice_all_ones_high():  # lattice.zig:74-83
    b = (MODELS / "lattice_speech.tflite").read_bytes()
    y = ctypes.c_double()
    ones, zeros, energy = f64(np.ones(21)), f64(np.zeros(21)), f64(np.full(21, 50.0))
    rc = L.vo_lattice_predict(b, len(b), dptr(ones), dptr(ones), dptr(zeros), dptr(energy), 21, ctypes.byref(y))
    assert rc == 0 and 4.0 < y.value <= 5.5


def test_lattice_monotone_in_similarity():
    # A perfect match scores above a poor one (sanity of the graph walk).
    b = (MODELS / "lattice_speech.tflite").read_bytes()
    y1, y2 = ctypes.c_double(), ctypes.c_double()
    ones, zeros, energy = f64(np.ones(21)), f64(np.zeros(21)), f64(np.full(21, 50.0))
    half = f64(np.full(21, 0.4))
    L.vo_lattice_predict(b, len(b), dptr(ones), dptr(ones), dptr(zeros), dptr(energy), 21, ctypes.byref(y1))
    L.vo_lattice_predict(b, len(b), dptr(half), dptr(half), dptr(f64(np.full(21, 0.2))), dptr(energy), 21,
                         ctypes.byref(y2))
    assert y1.value > y2.value


def test_lattice_rejects_truncated():  # tflite.zig:825-828
    y = ctypes.c_double()
    z = f64(np.zeros(21))
    for bad in (bytes([1, 2]), bytes([8, 0, 0, 0, 0, 0, 0, 0])):
        assert L.vo_lattice_predict(bad, len(bad), dptr(z), dptr(z), dptr(z), dptr(z), 21, ctypes.byref(y)) == -4
