"""visqol_measure_files and visqol_measure_files_batch on the device.

A WAV pair scores exactly as its decoded samples do through visqol_measure
(bit for bit, exact mode), and the file batch (capi.zig:375-456) loads every
pair, scores one device batch per sample rate, leaves failed pairs NULL with
their own status and returns the first failing code.
"""
import struct
import sys

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


def write_wav(path, x: np.ndarray, rate: int, fmt: str) -> None:
    """Mono float64, or 16-bit PCM stereo with both channels equal."""
    if fmt == "f64":
        tag, nch, bits, payload = 3, 1, 64, x.astype("<f8").tobytes()
    else:
        q = np.clip(np.round(x * 32768.0), -32768, 32767).astype("<i2")
        tag, nch, bits, payload = 1, 2, 16, np.repeat(q, 2).tobytes()
    block = nch * bits // 8
    fmt_body = struct.pack("<HHIIHH", tag, nch, rate, rate * block, block, bits)
    body = b"WAVE" + b"fmt " + struct.pack("<I", 16) + fmt_body + b"data" + struct.pack("<I", len(payload)) + payload
    path.write_bytes(b"RIFF" + struct.pack("<I", len(body)) + body)


def _same(a, b):
    assert a["moslqo"] == b["moslqo"]
    assert a["alignment_lag_s"] == b["alignment_lag_s"]
    for k in ("fvnsim", "fvnsim10", "fstdnsim", "fvdegenergy"):
        assert np.array_equal(a[k], b[k]), k
    assert len(a["patch_sims"]) == len(b["patch_sims"])


@pytest.mark.parametrize("fmt", ["f64", "pcm16"])
def test_measure_files_equals_measure(native, synth, tmp_path, fmt):
    ref, deg = synth.audio_pair(5, duration=3.0)
    rp, dp = tmp_path / "ref.wav", tmp_path / "deg.wav"
    write_wav(rp, ref, 48000, fmt)
    write_wav(dp, deg, 48000, fmt)
    r_dec, sr = native.decode_wav(rp.read_bytes())
    d_dec, _ = native.decode_wav(dp.read_bytes())
    if fmt == "f64":
        assert np.array_equal(r_dec, ref) and np.array_equal(d_dec, deg)
    say(f"measure_files fmt={fmt} lens {r_dec.size} {d_dec.size}")
    ctx = native.Context(0, precision="exact")
    _same(ctx.measure_files(rp, dp), ctx.measure(r_dec, d_dec, sr))


def test_measure_files_batch_groups_rates_and_reports_failures(native, synth, tmp_path):
    files = []
    for i, (kind, seed) in enumerate((("a", 1), ("s", 2), ("a", 3), ("s", 4))):
        if kind == "a":
            ref, deg = synth.audio_pair(seed, duration=2.0)
            rate = 48000
        else:
            ref, deg = synth.speech_pair(seed, duration=3.0)
            rate = 16000
        rp, dp = tmp_path / f"r{i}.wav", tmp_path / f"d{i}.wav"
        write_wav(rp, ref, rate, "f64")
        write_wav(dp, deg, rate, "f64")
        files.append((rp, dp, ref, deg, rate))
    mixed = tmp_path / "d_mixed.wav"
    write_wav(mixed, files[1][3], 16000, "f64")
    refs = [f[0] for f in files] + [files[0][0], tmp_path / "missing.wav"]
    degs = [f[1] for f in files] + [mixed, files[0][1]]
    ctx = native.Context(0, precision="exact")
    say("measure_files_batch: 2 audio + 2 speech-rate pairs, a rate mismatch, a missing file")
    rc, res, st = ctx.measure_files_batch(refs, degs)
    assert st[:4] == [0, 0, 0, 0]
    assert st[4] == native.E_SIGNAL and res[4] is None
    assert st[5] == native.E_IO and res[5] is None
    assert rc == native.E_SIGNAL  # the first failing pair's code
    for (rp, dp, ref, deg, rate), got in zip(files, res[:4]):
        _same(got, ctx.measure(ref, deg, rate))
