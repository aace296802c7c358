"""Seeded synthetic signal pairs for the benchmark workloads and tests.

No audio files ship with this package, so every workload is generated here
from a recorded PRNG (numpy PCG64) and a seed:

* ``audio_pair``: 48 kHz music-like reference (8 amplitude-modulated
  sinusoids at log-uniform 100 Hz-12 kHz plus pink noise at -30 dBFS); the
  degraded copy is delayed, has white noise at a given SNR and is quantised
  to 16-bit.
* ``speech_pair``: 16 kHz harmonic bursts (f0 100-250 Hz, 12 harmonics,
  150-400 ms voiced, 50-300 ms silences); degraded by a 4th-order
  Butterworth low-pass at 4 kHz plus white noise.
* ``plain_audio_pair``: the per-pair f64 check set: seeds 0-63, durations
  uniform 5-15 s, SNR uniform 10-40 dB, delay uniform 0-2000 samples.
* ``mixed_pair``: one pair of the uneven serving mix: 70 % audio / 30 %
  speech, log-uniform duration, noise 5-40 dB SNR, delay 0-4800 samples,
  0-3 % linear-resampling time stretch, 0-5 zeroed 20 ms dropouts.
* ``mix_source`` / ``mix_request``: the same mix for workloads of thousands
  of pairs: long source pairs carry the delay, low-pass and stretch, and
  each request is a slice of one (log-uniform duration, random offset) with
  its own noise, dropouts and quantisation.
"""
from __future__ import annotations

import numpy as np


def _pink(rng: np.random.Generator, n: int) -> np.ndarray:
    white = rng.standard_normal(n)
    spec = np.fft.rfft(white)
    f = np.arange(spec.size, dtype=np.float64)
    f[0] = 1.0
    spec /= np.sqrt(f)
    x = np.fft.irfft(spec, n)
    return x / np.sqrt(np.mean(x * x))


def _add_noise(rng: np.random.Generator, x: np.ndarray, snr_db: float) -> np.ndarray:
    p = np.mean(x * x)
    sigma = np.sqrt(p / (10.0 ** (snr_db / 10.0)))
    return x + sigma * rng.standard_normal(x.size)


def _quantise16(x: np.ndarray) -> np.ndarray:
    return np.clip(np.round(x * 32768.0), -32768, 32767) / 32768.0


def _delay(x: np.ndarray, k: int) -> np.ndarray:
    if k <= 0:
        return x.copy()
    out = np.zeros_like(x)
    out[k:] = x[: x.size - k]
    return out


def audio_reference(seed: int, duration: float = 10.0, sample_rate: int = 48000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = int(round(duration * sample_rate))
    t = np.arange(n, dtype=np.float64) / sample_rate
    freqs = np.exp(rng.uniform(np.log(100.0), np.log(12000.0), 8))
    amps = rng.uniform(0.02, 0.1, 8)
    mod_f = rng.uniform(2.0, 8.0, 8)
    ph = rng.uniform(0, 2 * np.pi, (2, 8))
    x = np.zeros(n)
    for k in range(8):
        env = 1.0 + 0.5 * np.sin(2 * np.pi * mod_f[k] * t + ph[0, k])
        x += amps[k] * env * np.sin(2 * np.pi * freqs[k] * t + ph[1, k])
    x += 10.0 ** (-30.0 / 20.0) * _pink(rng, n)
    return x


def audio_pair(seed: int, duration: float = 10.0, snr_db: float = 25.0, delay: int = 37,
               sample_rate: int = 48000):
    ref = audio_reference(seed, duration, sample_rate)
    rng = np.random.default_rng(seed + 1_000_003)
    deg = _quantise16(_add_noise(rng, _delay(ref, delay), snr_db))
    return ref, deg


def _butter_lowpass(x: np.ndarray, cutoff: float, fs: float) -> np.ndarray:
    from scipy.signal import butter, lfilter

    b, a = butter(4, cutoff, btype="low", fs=fs)
    return lfilter(b, a, x)


def speech_reference(seed: int, duration: float = 8.0, sample_rate: int = 16000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = int(round(duration * sample_rate))
    x = np.zeros(n)
    pos = int(rng.uniform(0.05, 0.3) * sample_rate)
    while pos < n:
        seg = int(rng.uniform(0.15, 0.4) * sample_rate)
        f0 = rng.uniform(100.0, 250.0)
        amps = rng.uniform(0.2, 1.0, 12) / np.arange(1, 13)
        t = np.arange(min(seg, n - pos)) / sample_rate
        burst = np.zeros(t.size)
        for h in range(12):
            if f0 * (h + 1) < sample_rate / 2:
                burst += amps[h] * np.sin(2 * np.pi * f0 * (h + 1) * t + rng.uniform(0, 2 * np.pi))
        burst *= np.hanning(t.size) if t.size > 2 else 1.0
        x[pos: pos + t.size] = 0.3 * burst
        pos += seg + int(rng.uniform(0.05, 0.3) * sample_rate)
    return x


def speech_pair(seed: int, duration: float = 8.0, snr_db: float = 20.0, delay: int = 0,
                sample_rate: int = 16000):
    ref = speech_reference(seed, duration, sample_rate)
    rng = np.random.default_rng(seed + 2_000_003)
    deg = _butter_lowpass(_delay(ref, delay), 4000.0, sample_rate)
    deg = _add_noise(rng, deg, snr_db)
    return ref, deg


def plain_audio_pair(seed: int):
    """(ref, deg) of the per-pair f64 check set, parameters drawn from the seed."""
    rng = np.random.default_rng(20_000_000 + seed)
    dur = float(rng.uniform(5.0, 15.0))
    snr = float(rng.uniform(10.0, 40.0))
    delay = int(rng.integers(0, 2001))
    return audio_pair(seed, duration=dur, snr_db=snr, delay=delay)


def _stretch(x: np.ndarray, factor: float) -> np.ndarray:
    n_out = max(2, int(round(x.size * factor)))
    pos = np.linspace(0.0, x.size - 1, n_out)
    return np.interp(pos, np.arange(x.size), x)


def mixed_pair(seed: int, max_duration: float = 60.0):
    """(ref, deg, sample_rate, mode) drawn from the uneven serving mix."""
    rng = np.random.default_rng(10_000_000 + seed)
    speech = rng.uniform() < 0.3
    dur = float(np.exp(rng.uniform(np.log(1.0), np.log(max_duration))))
    snr = rng.uniform(5.0, 40.0)
    delay = int(rng.integers(0, 4801))
    stretch = 1.0 + rng.uniform(0.0, 0.03)
    drops = int(rng.integers(0, 6))
    if speech:
        sr = 16000
        ref = speech_reference(seed, dur, sr)
        deg = _butter_lowpass(_delay(ref, delay), 4000.0, sr)
    else:
        sr = 48000
        ref = audio_reference(seed, dur, sr)
        deg = _delay(ref, delay)
    deg = _stretch(deg, stretch)
    deg = _add_noise(rng, deg, snr)
    dl = int(0.02 * sr)
    for _ in range(drops):
        if deg.size > dl:
            s = int(rng.integers(0, deg.size - dl))
            deg[s: s + dl] = 0.0
    if not speech:
        deg = _quantise16(deg)
    return ref, deg, sr, 1 if speech else 0


def mix_source(seed: int, duration: float):
    """A long pair of the uneven mix before its per-request degradations:
    (ref, deg, sample_rate, mode), 70 % 48 kHz audio / 30 % 16 kHz speech,
    the degraded copy delayed 0-4800 samples, low-passed (speech) and
    time-stretched 0-3 % by linear resampling.  mix_request cuts requests
    from it."""
    rng = np.random.default_rng(30_000_000 + seed)
    speech = rng.uniform() < 0.3
    delay = int(rng.integers(0, 4801))
    stretch = 1.0 + rng.uniform(0.0, 0.03)
    if speech:
        sr = 16000
        ref = speech_reference(40_000 + seed, duration, sr)
        deg = _butter_lowpass(_delay(ref, delay), 4000.0, sr)
    else:
        sr = 48000
        ref = audio_reference(40_000 + seed, duration, sr)
        deg = _delay(ref, delay)
    return ref, _stretch(deg, stretch), sr, 1 if speech else 0


def mix_request(src, seed: int, lo: float = 1.0, hi: float = 60.0):
    """One pair of the uneven mix cut from a mix_source pair: a duration
    log-uniform in [lo, hi] s (at most the source's), a random offset, then
    the pair's own degradations: white noise at 5-40 dB SNR, 0-5 zeroed
    20 ms dropouts, 16-bit quantisation for audio.  The reference is a view
    of the source; the degraded signal is new."""
    ref, deg, sr, mode = src
    rng = np.random.default_rng(50_000_000 + seed)
    n = min(ref.size, deg.size)
    d = min(n, max(2, int(np.exp(rng.uniform(np.log(lo), np.log(hi))) * sr)))
    o = int(rng.integers(0, n - d + 1))
    out = _add_noise(rng, deg[o:o + d], rng.uniform(5.0, 40.0))
    dl = int(0.02 * sr)
    for _ in range(int(rng.integers(0, 6))):
        if out.size > dl:
            s0 = int(rng.integers(0, out.size - dl))
            out[s0: s0 + dl] = 0.0
    if mode == 0:
        out = _quantise16(out)
    return ref[o:o + d], out, sr, mode


def uniform_durations(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    return np.random.default_rng(seed).uniform(lo, hi, n)
