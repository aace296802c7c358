"""WAV decoding for visqol_measure_files(_batch), on the host.

Pinned by the reference's two decoder tests (audio.zig:178-199: 16-bit
stereo channel mean and /32768; float32 mono) and, for the other formats,
by an independent decoder: Python's `wave` module for the RIFF framing of
PCM files, numpy for the sample conversions, the channel mean summed in
channel order and divided by the channel count as audio.zig:100-111 does.
Errors carry the reference's names and codes (capi.zig:94-112): format
errors are E_IO, empty audio and zero channels E_SIGNAL.
"""
import io
import struct
import wave
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "akvisqol_amd" / "libvisqol.so"

pytestmark = pytest.mark.skipif(not LIB.exists(), reason="libvisqol.so not built (run __graft_entry__.build())")

PCM, FLOAT, EXTENSIBLE = 1, 3, 0xFFFE


@pytest.fixture(scope="module")
def native():
    from akvisqol_amd import native as n

    return n


def riff(chunks: list[tuple[bytes, bytes]], riff_size: int | None = None) -> bytes:
    """RIFF/WAVE bytes from (id, body) chunks; odd bodies get a pad byte."""
    body = b"WAVE"
    for cid, data in chunks:
        body += cid + struct.pack("<I", len(data)) + data + (b"\0" if len(data) & 1 else b"")
    return b"RIFF" + struct.pack("<I", len(body) if riff_size is None else riff_size) + body


def fmt_chunk(tag: int, nch: int, rate: int, bits: int, ext_tag: int | None = None) -> bytes:
    block = nch * bits // 8
    b = struct.pack("<HHIIHH", tag, nch, rate, rate * block, block, bits)
    if ext_tag is not None:  # cbSize, valid bits, channel mask, SubFormat GUID
        b += struct.pack("<HHI", 22, bits, 0) + struct.pack("<H", ext_tag) + bytes(14)
    return b


def wav(tag: int, nch: int, rate: int, bits: int, payload: bytes, ext_tag: int | None = None) -> bytes:
    return riff([(b"fmt ", fmt_chunk(tag, nch, rate, bits, ext_tag)), (b"data", payload)])


def channel_mean(x: np.ndarray) -> np.ndarray:
    """[frames, nch] -> mono: sum in channel order, then / nch (mono as is)."""
    nch = x.shape[1]
    if nch == 1:
        return x[:, 0].copy()
    s = x[:, 0].copy()
    for c in range(1, nch):
        s = s + x[:, c]
    return s / nch


def test_reference_16bit_stereo(native):  # audio.zig:178-189
    payload = bytes([0x00, 0x40, 0x00, 0xC0, 0xFF, 0x7F, 0xFF, 0x7F])
    x, sr = native.decode_wav(wav(PCM, 2, 48000, 16, payload))
    assert sr == 48000
    assert x.tolist() == [0.0, 32767.0 / 32768.0]


def test_reference_float32_mono(native):  # audio.zig:191-199
    x, sr = native.decode_wav(wav(FLOAT, 1, 48000, 32, struct.pack("<ff", 0.5, -0.25)))
    assert x.tolist() == [0.5, -0.25]


@pytest.mark.parametrize("bits", [8, 16, 24, 32])
@pytest.mark.parametrize("nch", [1, 2, 3, 6])
def test_pcm_matches_wave_module(native, bits, nch):
    rng = np.random.default_rng(bits * 10 + nch)
    frames = 257
    width = bits // 8
    buf = io.BytesIO()
    with wave.open(buf, "wb") as w:
        w.setnchannels(nch)
        w.setsampwidth(width)
        w.setframerate(44100)
        raw = rng.integers(0, 256, size=frames * nch * width, dtype=np.uint8).tobytes()
        w.writeframes(raw)
    data = buf.getvalue()
    with wave.open(io.BytesIO(data), "rb") as r:
        assert r.getnframes() == frames
        got = r.readframes(frames)
    b = np.frombuffer(got, dtype=np.uint8).reshape(frames, nch, width).astype(np.int64)
    if bits == 8:
        v = (b[..., 0].astype(np.float64) - 128.0) / 128.0
    else:
        u = np.zeros(b.shape[:2], dtype=np.int64)
        for k in range(width):
            u |= b[..., k] << (8 * k)
        top = 1 << (bits - 1)
        signed = np.where(u >= top, u - (1 << bits), u)
        v = signed.astype(np.float64) / float(top)
    want = channel_mean(v)
    x, sr = native.decode_wav(data)
    assert sr == 44100
    assert np.array_equal(x, want)


@pytest.mark.parametrize("bits,dt", [(32, "<f4"), (64, "<f8")])
@pytest.mark.parametrize("nch", [1, 2])
def test_float_formats(native, bits, dt, nch):
    rng = np.random.default_rng(bits + nch)
    v = rng.standard_normal((100, nch)).astype(dt)
    x, sr = native.decode_wav(wav(FLOAT, nch, 16000, bits, v.tobytes()))
    assert sr == 16000
    assert np.array_equal(x, channel_mean(v.astype(np.float64)))


def test_extensible_takes_subformat_tag(native):
    v = np.array([[1000, -2000], [32767, -32768]], dtype="<i2")
    x, _ = native.decode_wav(wav(EXTENSIBLE, 2, 48000, 16, v.tobytes(), ext_tag=PCM))
    assert np.array_equal(x, channel_mean(v.astype(np.float64) / 32768.0))
    f = np.array([0.125, -0.5], dtype="<f4")
    x, _ = native.decode_wav(wav(EXTENSIBLE, 1, 48000, 32, f.tobytes(), ext_tag=FLOAT))
    assert x.tolist() == [0.125, -0.5]


def test_chunk_walk(native):
    """Odd-sized chunks before fmt are skipped with their pad byte, a data
    chunk whose size overruns the file (streamed WAVs) is clamped, and a
    trailing partial frame is dropped."""
    v = np.arange(-6, 6, dtype="<i2") * 1000
    data = riff([(b"LIST", b"abc"), (b"fmt ", fmt_chunk(PCM, 2, 8000, 16)), (b"junk", b"12345"),
                 (b"data", v.tobytes() + b"\x01")])
    x, sr = native.decode_wav(data)
    assert sr == 8000
    assert np.array_equal(x, channel_mean(v.reshape(-1, 2).astype(np.float64) / 32768.0))
    streamed = bytearray(wav(PCM, 1, 8000, 16, v.tobytes()))
    streamed[40:44] = struct.pack("<I", 0xFFFFFFFF)
    x, _ = native.decode_wav(bytes(streamed))
    assert np.array_equal(x, v.astype(np.float64) / 32768.0)


@pytest.mark.parametrize("case,code,name", [
    ("short", "E_IO", "NotRiffWave"),
    ("not_wave", "E_IO", "NotRiffWave"),
    ("no_fmt", "E_IO", "MissingFmtChunk"),
    ("fmt_after_data", "E_IO", "MissingFmtChunk"),
    ("no_data", "E_IO", "MissingDataChunk"),
    ("fmt_short", "E_IO", "UnsupportedFormat"),
    ("bits12", "E_IO", "UnsupportedFormat"),
    ("float16", "E_IO", "UnsupportedFormat"),
    ("adpcm", "E_IO", "UnsupportedFormat"),
    ("zero_channels", "E_SIGNAL", "ChannelCountZero"),
    ("empty", "E_SIGNAL", "EmptyAudio"),
    ("partial_frame", "E_SIGNAL", "EmptyAudio"),
])
def test_errors(native, case, code, name):
    pcm = np.zeros(4, dtype="<i2").tobytes()
    data = {
        "short": b"RIFF\0\0\0\0WAV",
        "not_wave": b"RIFF\0\0\0\0AVI " + bytes(40),
        "no_fmt": riff([(b"data", pcm)]),
        "fmt_after_data": riff([(b"data", pcm), (b"fmt ", fmt_chunk(PCM, 1, 8000, 16))]),
        "no_data": riff([(b"fmt ", fmt_chunk(PCM, 1, 8000, 16))]),
        "fmt_short": riff([(b"fmt ", fmt_chunk(PCM, 1, 8000, 16)[:14]), (b"data", pcm)]),
        "bits12": wav(PCM, 1, 8000, 12, pcm),
        "float16": wav(FLOAT, 1, 8000, 16, pcm),
        "adpcm": wav(2, 1, 8000, 16, pcm),
        "zero_channels": riff([(b"fmt ", struct.pack("<HHIIHH", PCM, 0, 8000, 0, 0, 16)), (b"data", pcm)]),
        "empty": wav(PCM, 1, 8000, 16, b""),
        "partial_frame": wav(PCM, 2, 8000, 16, b"\x01\x02"),
    }[case]
    with pytest.raises(native.NativeError) as e:
        native.decode_wav(data)
    assert e.value.code == getattr(native, code)
    assert name in str(e.value)


def test_small_buffer_is_invalid_argument(native):
    import ctypes

    data = wav(PCM, 1, 8000, 16, np.arange(10, dtype="<i2").tobytes())
    n = ctypes.c_size_t(0)
    out = np.empty(4)
    rc = native.lib.visqol_hip_decode_wav(data, len(data), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 4,
                                          ctypes.byref(n), None)
    assert rc == native.E_INVALID_ARGUMENT and n.value == 10
