"""The C-ABI boundary without a GPU: the library loads, exports every symbol
the headers declare, keeps the reference's struct layouts and defaults,
shares the oracle's filter design bit for bit, and refuses to score
loudly when no gfx950 device is present."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "akvisqol_amd" / "libvisqol.so"

pytestmark = pytest.mark.skipif(not LIB.exists(), reason="libvisqol.so not built (run __graft_entry__.build())")


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        names |= set(re.findall(r"\b(visqol_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    declared = _declared()
    assert len(declared) >= 21
    for name in sorted(declared):
        assert hasattr(lib, name), name


def test_struct_layouts_match_reference():
    from akvisqol_amd import native

    # visqol.h layouts measured for x86-64 (capi.zig:28-63).
    assert ctypes.sizeof(native.Options) == 32
    assert native.Options.model_path.offset == 16 and native.Options.backend.offset == 24
    assert native.Options.precision.offset == 28
    assert ctypes.sizeof(native.Patch) == 64
    assert ctypes.sizeof(native.Result) == 88
    assert native.Result.num_patches.offset == 72 and native.Result.patch_sims.offset == 80


def test_abi_version_and_defaults():
    from akvisqol_amd import native

    assert native.lib.visqol_abi_version() == 1
    assert native.__version__ == "0.1.0"
    o = native.default_options()
    assert (o.mode, o.search_window_radius, o.use_lattice_model, o.precision) == (0, 60, 1, 3)
    # the reference's default (capi.zig:136-149): backend CPU, served here
    # by the exact f64 device path
    assert o.backend == native.BACKENDS["cpu"] == 0


def test_header_constants():
    h = (ROOT / "include" / "visqol.h").read_text()
    for name, val in [("VISQOL_E_INVALID_ARGUMENT", -1), ("VISQOL_E_OUT_OF_MEMORY", -2), ("VISQOL_E_IO", -3),
                      ("VISQOL_E_MODEL", -4), ("VISQOL_E_SIGNAL", -5), ("VISQOL_E_BACKEND", -6),
                      ("VISQOL_E_INTERNAL", -7), ("VISQOL_BACKEND_HIP", 4), ("VISQOL_PRECISION_AUTO", 3)]:
        assert re.search(rf"{name}\s*=\s*{val}\b", h), name


@pytest.mark.parametrize("mode,sr", [(0, 48000), (1, 16000), (0, 44100), (1, 8000)])
def test_filter_design_bit_identical_to_oracle(mode, sr):
    from akvisqol_amd import native
    from oracle import vo

    nb = 32 if mode == 0 else 21
    high = sr / 2.0 if mode == 0 else 8000.0
    bands, cf = native.filters(sr, mode)
    ob, ocf = vo.filters(sr, nb, 50.0, high)
    want = np.array([[b.a0, b.a11, b.a12, b.a13, b.a14, b.b1, b.b2, b.gain] for b in ob])
    assert np.array_equal(bands, want)
    assert np.array_equal(cf, ocf)


def test_create_fails_loudly_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from akvisqol_amd import native

    with pytest.raises(native.NativeError) as e:
        native.Context(native.MODE_AUDIO)
    assert e.value.code == native.E_BACKEND
    assert native.lib.visqol_available_backends() == 0


def test_invalid_arguments_before_device():
    from akvisqol_amd import native

    L = native.lib
    h = ctypes.c_void_p()
    o = native.default_options()
    o.mode = 7
    assert L.visqol_create(ctypes.byref(o), ctypes.byref(h)) == native.E_INVALID_ARGUMENT
    assert b"mode" in L.visqol_last_error()
    o = native.default_options()
    o.precision = 9
    assert L.visqol_create(ctypes.byref(o), ctypes.byref(h)) == native.E_INVALID_ARGUMENT
    for backend in (1, 2, 3):  # Metal, CUDA, Vulkan: not provided here
        o = native.default_options()
        o.backend = backend
        assert L.visqol_create(ctypes.byref(o), ctypes.byref(h)) == native.E_BACKEND
        assert b"VISQOL_BACKEND_HIP" in L.visqol_last_error()
    o = native.default_options()
    o.backend = 5  # past the enum: an invalid argument, as capi.zig:171-174
    assert L.visqol_create(ctypes.byref(o), ctypes.byref(h)) == native.E_INVALID_ARGUMENT
    assert L.visqol_create(ctypes.byref(o), None) == native.E_INVALID_ARGUMENT
    assert L.visqol_measure(None, None, 0, None, 0, 48000, None) == native.E_INVALID_ARGUMENT
    assert L.visqol_measure_batch(None, 0, None, None, None, None, 48000, 0, None, None) == native.E_INVALID_ARGUMENT
    M = L.visqol_hip_measure_mixed_batch_device
    assert M(None, 0, 0, None, None, None, None, None, None, None, None) == native.OK  # n == 0
    assert M(None, 0, 1, None, None, None, None, None, None, None, None) == native.E_INVALID_ARGUMENT
    hs = (ctypes.c_void_p * 2)(None, None)
    one = (ctypes.c_uint32 * 1)(0)
    ptrs = (ctypes.c_void_p * 1)(None)
    lens = (ctypes.c_size_t * 1)(0)
    out = (ctypes.POINTER(native.Result) * 1)()
    assert M(hs, 2, 1, one, ptrs, lens, ptrs, lens, one, out, None) == native.E_INVALID_ARGUMENT
    assert b"ctx is NULL" in L.visqol_last_error()
    L.visqol_result_free(None)
    L.visqol_destroy(None)


def test_gradient_and_probe_arguments_before_device():
    # the gradient entry points check their context and arrays first, as
    # visqol_measure_batch does (capi.zig:340-350); the host probes check
    # theirs without touching a device
    from akvisqol_amd import native

    L = native.lib
    assert L.visqol_hip_measure_batch_grad(None, 1, None, None, None, None, 48000, None, None, None) == \
        native.E_INVALID_ARGUMENT
    assert L.visqol_hip_measure_prepared_grad(None, None, 1, None, None, None, None, None, None) == \
        native.E_INVALID_ARGUMENT
    assert L.visqol_hip_patch_similarity_grad(None, None, 2, 2, None, None, None, None, None) == \
        native.E_INVALID_ARGUMENT
    assert L.visqol_hip_prepare_grad(None, None, 2, 2, 2, None, None, None, None) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_lattice_predict(None, 0, 21, 1, None, None, None, None, None) == native.E_INVALID_ARGUMENT
    frame = np.zeros(3840)
    g = np.zeros(32)
    out = np.zeros(3840)
    p = native._dptr
    # wrong window length, lo > hi, hi beyond the window: rejected before any device use
    assert L.visqol_hip_frame_grad(0, 48000, p(frame), 100, p(g), 0, 100, p(out)) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_frame_grad(0, 48000, p(frame), 3840, p(g), 10, 5, p(out)) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_frame_grad(0, 48000, p(frame), 3840, p(g), 0, 4000, p(out)) == native.E_INVALID_ARGUMENT
    # the reference-order sum probe: kind, form, pointers and an empty
    # alternating sum (its mean would divide by zero) checked first
    s = ctypes.c_double()
    assert L.visqol_hip_sum(4, p(frame), 10, 0.0, 0, 0, 0, ctypes.byref(s)) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_sum(0, p(frame), 10, 0.0, 0, 0, 3, ctypes.byref(s)) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_sum(0, None, 10, 0.0, 0, 0, 0, ctypes.byref(s)) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_sum(0, p(frame), 10, 0.0, 0, 0, 0, None) == native.E_INVALID_ARGUMENT
    assert L.visqol_hip_sum(2, p(frame), 0, 0.0, 0, 0, 0, ctypes.byref(s)) == native.E_INVALID_ARGUMENT


def test_noalloc_limit_checked_before_device():
    # the size limit is checked before any device use (the pointer is never
    # dereferenced): a transform beyond 2^20 points is refused
    from akvisqol_amd import native

    L = native.lib
    lag = ctypes.c_int64()
    fake = ctypes.c_void_p(256)
    assert L.visqol_hip_best_lag_noalloc(fake, (1 << 19) + 1, fake, 10, ctypes.byref(lag)) == native.E_INVALID_ARGUMENT
    assert b"limit" in L.visqol_last_error()
    assert L.visqol_hip_best_lag_noalloc(None, 5, fake, 5, ctypes.byref(lag)) == native.E_INVALID_ARGUMENT
