"""ctypes binding to libvisqol.so (include/visqol.h, include/visqol_hip.h).

Mirrors the reference's Python binding (src/python/visqol/_native.py:16-234):
the same structs, error codes and Context.measure / measure_batch
behaviour, with "hip" added to the backend table.  The library must have
been built (``python -m akvisqol_amd.build``); there is no fallback: a
missing library raises ImportError, and scoring needs a gfx950 device.
"""
from __future__ import annotations

import ctypes
import os
import threading
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint32, c_void_p
from pathlib import Path

import numpy as np

MODE_AUDIO = 0
MODE_SPEECH = 1

# _native.py:19 plus this library's backend.
BACKENDS = {"cpu": 0, "metal": 1, "cuda": 2, "vulkan": 3, "hip": 4}
PRECISIONS = {"exact": 0, "fast": 1, "balanced": 2, "auto": 3}

OK = 0
E_INVALID_ARGUMENT = -1
E_OUT_OF_MEMORY = -2
E_IO = -3
E_MODEL = -4
E_SIGNAL = -5
E_BACKEND = -6
E_INTERNAL = -7


class Options(ctypes.Structure):
    _fields_ = [
        ("mode", c_uint32),
        ("search_window_radius", c_int32),
        ("use_unscaled_speech_mos_mapping", c_uint8),
        ("disable_global_alignment", c_uint8),
        ("disable_realignment", c_uint8),
        ("use_lattice_model", c_uint8),
        ("model_path", c_char_p),
        ("backend", c_uint32),
        ("precision", c_uint32),
    ]


class Patch(ctypes.Structure):
    _fields_ = [
        ("similarity", c_double),
        ("ref_patch_start_time", c_double),
        ("ref_patch_end_time", c_double),
        ("deg_patch_start_time", c_double),
        ("deg_patch_end_time", c_double),
        ("freq_band_means", POINTER(c_double)),
        ("freq_band_stddevs", POINTER(c_double)),
        ("freq_band_deg_energy", POINTER(c_double)),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("moslqo", c_double),
        ("vnsim", c_double),
        ("alignment_lag_s", c_double),
        ("num_bands", c_size_t),
        ("fvnsim", POINTER(c_double)),
        ("fvnsim10", POINTER(c_double)),
        ("fstdnsim", POINTER(c_double)),
        ("fvdegenergy", POINTER(c_double)),
        ("center_freq_bands", POINTER(c_double)),
        ("num_patches", c_size_t),
        ("patch_sims", POINTER(Patch)),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("workers", c_uint32),
        ("max_batch", c_uint32),
        ("slab_bytes", ctypes.c_uint64),
        ("max_inflight_bytes", ctypes.c_uint64),
    ]


class RequestTiming(ctypes.Structure):
    _fields_ = [
        ("admit_ms", c_double),
        ("upload_ms", c_double),
        ("queue_ms", c_double),
        ("run_ms", c_double),
        ("total_ms", c_double),
        ("batch_size", c_uint32),
    ]


class Trace(ctypes.Structure):
    _fields_ = [
        ("lag", c_int64),
        ("spl_factor", c_double),
        ("nf_ref", c_size_t),
        ("nf_deg", c_size_t),
        ("num_patches", c_size_t),
        ("ref_indices", c_size_t * 512),
        ("realign_lag", c_int64 * 512),
        ("realign_frames", c_size_t * 512),
        ("realign_accepted", c_uint8 * 512),
        ("match_valid", c_uint8 * 512),
        ("dp_offset", c_size_t * 512),
    ]


LIB_PATH = Path(os.environ.get("VISQOL_LIBRARY", Path(__file__).resolve().parent / "libvisqol.so"))


def _load() -> ctypes.CDLL:
    if not LIB_PATH.is_file():
        raise ImportError(f"libvisqol.so not found at {LIB_PATH}; build it with `python -m akvisqol_amd.build`")
    L = ctypes.CDLL(str(LIB_PATH))
    D = POINTER(c_double)
    S = POINTER(c_size_t)
    sig = {
        "visqol_abi_version": (c_uint32, []),
        "visqol_version": (c_char_p, []),
        "visqol_last_error": (c_char_p, []),
        "visqol_available_backends": (c_uint32, []),
        "visqol_options_default": (None, [POINTER(Options)]),
        "visqol_create": (c_int, [POINTER(Options), POINTER(c_void_p)]),
        "visqol_destroy": (None, [c_void_p]),
        "visqol_measure": (c_int, [c_void_p, D, c_size_t, D, c_size_t, c_uint32, POINTER(POINTER(Result))]),
        "visqol_measure_files": (c_int, [c_void_p, c_char_p, c_char_p, POINTER(POINTER(Result))]),
        "visqol_measure_batch": (c_int, [c_void_p, c_size_t, POINTER(D), S, POINTER(D), S, c_uint32, c_uint32,
                                         POINTER(POINTER(Result)), POINTER(c_int32)]),
        "visqol_measure_files_batch": (c_int, [c_void_p, c_size_t, POINTER(c_char_p), POINTER(c_char_p), c_uint32,
                                               POINTER(POINTER(Result)), POINTER(c_int32)]),
        "visqol_result_free": (None, [POINTER(Result)]),
        "visqol_hip_measure_trace": (c_int, [c_void_p, D, c_size_t, D, c_size_t, c_uint32, POINTER(POINTER(Result)),
                                             POINTER(Trace), D, D, D]),
        "visqol_hip_measure_batch_device": (c_int, [c_void_p, c_size_t, POINTER(c_void_p), S, POINTER(c_void_p), S,
                                                    c_uint32, POINTER(POINTER(Result)), POINTER(c_int32)]),
        "visqol_hip_measure_mixed_batch_device": (c_int, [POINTER(c_void_p), c_size_t, c_size_t, POINTER(c_uint32),
                                                          POINTER(c_void_p), S, POINTER(c_void_p), S,
                                                          POINTER(c_uint32), POINTER(POINTER(Result)),
                                                          POINTER(c_int32)]),
        "visqol_hip_profile": (c_int, [c_void_p, c_int, c_int]),
        "visqol_hip_profile_read": (c_int, [c_void_p, D, D, D, POINTER(c_int64)]),
        "visqol_hip_profile_phases": (c_int, [c_void_p, D, c_int]),
        "visqol_hip_profile_sim": (c_int, [c_void_p, D, D, POINTER(c_int64)]),
        "visqol_hip_lattice_predict": (c_int, [ctypes.c_char_p, c_size_t, c_size_t, c_size_t, D, D, D, D, D]),
        "visqol_hip_frame_grad": (c_int, [c_uint32, c_uint32, D, c_size_t, D, c_size_t, c_size_t, D]),
        "visqol_hip_patch_similarity_grad": (c_int, [D, D, c_size_t, c_size_t, D, D, D, D, D]),
        "visqol_hip_prepare_grad": (c_int, [D, D, c_size_t, c_size_t, c_size_t, D, D, D, D]),
        "visqol_hip_gammatone_shape": (c_int, [c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "visqol_hip_gammatone_tuning": (c_int, [c_void_p, c_uint32, c_int, POINTER(c_int), POINTER(ctypes.c_float),
                                                c_int]),
        "visqol_hip_decode_wav": (c_int, [ctypes.c_char_p, c_size_t, D, c_size_t, S, POINTER(c_uint32)]),
        "visqol_hip_alloc": (c_int, [c_size_t, POINTER(c_void_p)]),
        "visqol_hip_free": (c_int, [c_void_p]),
        "visqol_hip_upload": (c_int, [c_void_p, c_void_p, c_size_t]),
        "visqol_hip_spectrogram": (c_int, [c_uint32, c_uint32, c_uint32, D, c_size_t, D, S]),
        "visqol_hip_filters": (c_int, [c_uint32, c_uint32, D, D]),
        "visqol_hip_best_lag": (c_int, [D, c_size_t, D, c_size_t, POINTER(c_int64)]),
        "visqol_hip_sum": (c_int, [c_int, D, c_size_t, c_double, c_int64, c_size_t, c_int, POINTER(c_double)]),
        "visqol_hip_best_lag_noalloc": (c_int, [c_void_p, c_size_t, c_void_p, c_size_t, POINTER(c_int64)]),
        "visqol_hip_math_probe": (c_int, [c_int, D, D, D, c_size_t]),
        "visqol_hip_fft": (c_int, [D, c_size_t, c_int]),
        "visqol_hip_xcorr": (c_int, [D, c_size_t, D, c_size_t, D, c_size_t, POINTER(c_size_t)]),
        "visqol_hip_xcorr32": (c_int, [D, c_size_t, D, c_size_t, POINTER(ctypes.c_float), c_size_t,
                                       POINTER(c_size_t)]),
        "visqol_hip_best_lag_fast": (c_int, [D, c_size_t, D, c_size_t, c_int, POINTER(c_int64), POINTER(c_int)]),
        "visqol_hip_prepare": (c_int, [c_void_p, c_size_t, POINTER(D), S, c_uint32, POINTER(c_void_p)]),
        "visqol_hip_measure_prepared": (c_int, [c_void_p, c_void_p, c_size_t, S, POINTER(D), S,
                                                POINTER(POINTER(Result)), POINTER(c_int32)]),
        "visqol_hip_prepared_free": (None, [c_void_p]),
        "visqol_hip_measure_batch_grad": (c_int, [c_void_p, c_size_t, POINTER(D), S, POINTER(D), S, c_uint32,
                                                  POINTER(POINTER(Result)), POINTER(c_int32), POINTER(D)]),
        "visqol_hip_measure_prepared_grad": (c_int, [c_void_p, c_void_p, c_size_t, S, POINTER(D), S,
                                                     POINTER(POINTER(Result)), POINTER(c_int32), POINTER(D)]),
        "visqol_hip_server_config_default": (None, [POINTER(ServerConfig)]),
        "visqol_hip_server_create": (c_int, [POINTER(Options), POINTER(ServerConfig), POINTER(c_void_p)]),
        "visqol_hip_server_submit": (c_int, [c_void_p, c_uint32, D, c_size_t, D, c_size_t, c_uint32,
                                             POINTER(ctypes.c_uint64)]),
        "visqol_hip_server_wait": (c_int, [c_void_p, ctypes.c_uint64, POINTER(POINTER(Result)), D]),
        "visqol_hip_server_wait_timed": (c_int, [c_void_p, ctypes.c_uint64, POINTER(POINTER(Result)),
                                                 POINTER(RequestTiming)]),
        "visqol_hip_server_stats": (c_int, [c_void_p, POINTER(ctypes.c_uint64), POINTER(ctypes.c_uint64),
                                            POINTER(ctypes.c_uint64)]),
        "visqol_hip_server_destroy": (None, [c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()

ABI_VERSION = 1
if lib.visqol_abi_version() != ABI_VERSION:
    raise ImportError(f"libvisqol ABI {lib.visqol_abi_version()} does not match the bindings ({ABI_VERSION})")

__version__ = lib.visqol_version().decode()


class NativeError(RuntimeError):
    """A visqol.h error code with the thread's last error message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"visqol error {code}: {message}")
        self.code = code


def _check(rc: int) -> None:
    if rc != OK:
        raise NativeError(rc, lib.visqol_last_error().decode(errors="replace"))


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


def result_dict(r: Result) -> dict:
    """A visqol_result as plain Python / numpy values (copied)."""
    nb, npch = r.num_bands, r.num_patches

    def arr(p, n):
        return np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0)

    patches = []
    for i in range(npch):
        p = r.patch_sims[i]
        patches.append(dict(
            similarity=p.similarity, ref_patch_start_time=p.ref_patch_start_time,
            ref_patch_end_time=p.ref_patch_end_time, deg_patch_start_time=p.deg_patch_start_time,
            deg_patch_end_time=p.deg_patch_end_time, freq_band_means=arr(p.freq_band_means, nb),
            freq_band_stddevs=arr(p.freq_band_stddevs, nb), freq_band_deg_energy=arr(p.freq_band_deg_energy, nb)))
    return dict(moslqo=r.moslqo, vnsim=r.vnsim, alignment_lag_s=r.alignment_lag_s, num_bands=nb,
                fvnsim=arr(r.fvnsim, nb), fvnsim10=arr(r.fvnsim10, nb), fstdnsim=arr(r.fstdnsim, nb),
                fvdegenergy=arr(r.fvdegenergy, nb), center_freq_bands=arr(r.center_freq_bands, nb),
                patch_sims=patches)


def default_options() -> Options:
    o = Options()
    lib.visqol_options_default(ctypes.byref(o))
    return o


class Context:
    """One visqol_ctx (_native.py:165-234); calls are serialised per context."""

    def __init__(self, mode: int = MODE_AUDIO, *, use_lattice_model: bool = True, search_window_radius: int = 60,
                 disable_global_alignment: bool = False, disable_realignment: bool = False,
                 use_unscaled_speech_mos_mapping: bool = False, model_path: str | None = None,
                 backend: str | int = "hip", precision: str | int = "exact"):
        o = default_options()
        o.mode = mode
        o.use_lattice_model = int(use_lattice_model)
        o.search_window_radius = search_window_radius
        o.disable_global_alignment = int(disable_global_alignment)
        o.disable_realignment = int(disable_realignment)
        o.use_unscaled_speech_mos_mapping = int(use_unscaled_speech_mos_mapping)
        o.model_path = model_path.encode() if model_path else None
        o.backend = BACKENDS[backend] if isinstance(backend, str) else backend
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
        h = c_void_p()
        _check(lib.visqol_create(ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._lock = threading.Lock()
        self.mode = mode

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.visqol_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def measure(self, ref, deg, sample_rate: int) -> dict:
        ref, deg = _f64(ref), _f64(deg)
        out = POINTER(Result)()
        with self._lock:
            rc = lib.visqol_measure(self._h, _dptr(ref), ref.size, _dptr(deg), deg.size, sample_rate,
                                    ctypes.byref(out))
            _check(rc)
        try:
            return result_dict(out.contents)
        finally:
            lib.visqol_result_free(out)

    def measure_files(self, ref_path, deg_path) -> dict:
        """visqol_measure_files: both WAVs decoded to mono f64 and scored."""
        out = POINTER(Result)()
        with self._lock:
            rc = lib.visqol_measure_files(self._h, os.fsencode(ref_path), os.fsencode(deg_path), ctypes.byref(out))
            _check(rc)
        try:
            return result_dict(out.contents)
        finally:
            lib.visqol_result_free(out)

    def measure_files_batch(self, ref_paths, deg_paths, num_threads: int = 0):
        """(first error code, [result dict or None], [status]) over WAV pairs;
        pairs are grouped into one device batch per sample rate."""
        n = len(ref_paths)
        rp = (c_char_p * n)(*[os.fsencode(p) for p in ref_paths])
        dp = (c_char_p * n)(*[os.fsencode(p) for p in deg_paths])
        out = (POINTER(Result) * n)()
        st = (c_int32 * n)()
        with self._lock:
            rc = lib.visqol_measure_files_batch(self._h, n, rp, dp, num_threads, out, st)
        results = []
        for i in range(n):
            if out[i]:
                results.append(result_dict(out[i].contents))
                lib.visqol_result_free(out[i])
            else:
                results.append(None)
        return rc, results, list(st)

    def measure_trace(self, ref, deg, sample_rate: int, want_stages: bool = False):
        """visqol_hip_measure_trace: (result, trace dict with optional stages)."""
        ref, deg = _f64(ref), _f64(deg)
        out = POINTER(Result)()
        tr = Trace()
        rs = ds = sims = None
        if want_stages:
            nb = 32 if self.mode == 0 else 21
            hop = int(int(np.floor(sample_rate * 0.08 + 0.5)) * 0.25)
            patch = 30 if self.mode == 0 else 20
            nfr = ref.size // hop + 1
            nfd = (deg.size + ref.size // 2) // hop + 1
            rs, ds = np.zeros(nb * nfr), np.zeros(nb * nfd)
            sims = np.zeros((nfr // patch + 2) * nfd)
        with self._lock:
            rc = lib.visqol_hip_measure_trace(self._h, _dptr(ref), ref.size, _dptr(deg), deg.size, sample_rate,
                                              ctypes.byref(out), ctypes.byref(tr),
                                              _dptr(rs) if rs is not None else None,
                                              _dptr(ds) if ds is not None else None,
                                              _dptr(sims) if sims is not None else None)
            _check(rc)
        try:
            res = result_dict(out.contents)
        finally:
            lib.visqol_result_free(out)
        npch = tr.num_patches
        trace = dict(lag=tr.lag, spl_factor=tr.spl_factor, nf_ref=tr.nf_ref, nf_deg=tr.nf_deg, num_patches=npch,
                     ref_indices=list(tr.ref_indices[:npch]), realign_lag=list(tr.realign_lag[:npch]),
                     realign_frames=list(tr.realign_frames[:npch]),
                     realign_accepted=list(tr.realign_accepted[:npch]), match_valid=list(tr.match_valid[:npch]),
                     dp_offset=list(tr.dp_offset[:npch]))
        if want_stages:
            nb = 32 if self.mode == 0 else 21
            trace["ref_spec"] = rs[: nb * tr.nf_ref].reshape(nb, tr.nf_ref).copy()
            trace["deg_spec"] = ds[: nb * tr.nf_deg].reshape(nb, tr.nf_deg).copy()
            trace["sims"] = sims[: npch * tr.nf_deg].reshape(npch, tr.nf_deg).copy()
        return res, trace

    def profile(self, enable: bool = True, reset: bool = True) -> None:
        _check(lib.visqol_hip_profile(self._h, int(enable), int(reset)))

    def profile_read(self) -> dict:
        ms, fl, by = c_double(), c_double(), c_double()
        n = c_int64()
        _check(lib.visqol_hip_profile_read(self._h, ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by),
                                           ctypes.byref(n)))
        return dict(ms=ms.value, flop=fl.value, bytes=by.value, launches=n.value)

    def profile_sim(self) -> dict:
        """Similarity-table launches since the last reset: device ms, cells, launches."""
        ms, cells = c_double(), c_double()
        n = c_int64()
        _check(lib.visqol_hip_profile_sim(self._h, ctypes.byref(ms), ctypes.byref(cells), ctypes.byref(n)))
        return dict(ms=ms.value, cells=cells.value, launches=n.value)

    def prepare(self, refs, sample_rate: int) -> "Prepared":
        """visqol_hip_prepare: references kept on the device for measure_prepared."""
        refs = [_f64(r) for r in refs]
        n = len(refs)
        D = POINTER(c_double)
        rp = (D * n)(*[_dptr(r) for r in refs])
        rl = (c_size_t * n)(*[r.size for r in refs])
        h = c_void_p()
        with self._lock:
            _check(lib.visqol_hip_prepare(self._h, n, rp, rl, sample_rate, ctypes.byref(h)))
        return Prepared(h, n)

    def measure_prepared(self, prep: "Prepared", degs, ref_index=None, convert: bool = True):
        """(first error code, [result dict or None], [status]): degs[i] against
        reference ref_index[i] (default i)."""
        degs = [_f64(d) for d in degs]
        n = len(degs)
        D = POINTER(c_double)
        dp = (D * n)(*[_dptr(d) for d in degs])
        dl = (c_size_t * n)(*[d.size for d in degs])
        ri = None if ref_index is None else (c_size_t * n)(*[int(k) for k in ref_index])
        out = (POINTER(Result) * n)()
        st = (c_int32 * n)()
        with self._lock:
            rc = lib.visqol_hip_measure_prepared(self._h, prep._h, n, ri, dp, dl, out, st)
        res = []
        for i in range(n):
            if out[i]:
                res.append(result_dict(out[i].contents) if convert else out[i].contents.moslqo)
                lib.visqol_result_free(out[i])
            else:
                res.append(None)
        return rc, res, list(st)

    def measure_batch_grad(self, refs, degs, sample_rate: int):
        """(first error code, [result dict or None], [status], [∂moslqo/∂deg])
        as the reference's measureBatchGrad (backend.zig:101-110): the
        forward pass's discrete choices held fixed; zeros for a failed pair."""
        refs = [_f64(r) for r in refs]
        degs = [_f64(d) for d in degs]
        n = len(refs)
        D = POINTER(c_double)
        grads = [np.zeros(d.size) for d in degs]
        rp = (D * n)(*[_dptr(r) for r in refs])
        dp = (D * n)(*[_dptr(d) for d in degs])
        gp = (D * n)(*[_dptr(g) for g in grads])
        rl = (c_size_t * n)(*[r.size for r in refs])
        dl = (c_size_t * n)(*[d.size for d in degs])
        out = (POINTER(Result) * n)()
        st = (c_int32 * n)()
        with self._lock:
            rc = lib.visqol_hip_measure_batch_grad(self._h, n, rp, rl, dp, dl, sample_rate, out, st, gp)
        return rc, self._take(out, n), list(st), grads

    def measure_prepared_grad(self, prep: "Prepared", degs, ref_index=None):
        """measure_prepared with ∂moslqo/∂deg of every pair."""
        degs = [_f64(d) for d in degs]
        n = len(degs)
        D = POINTER(c_double)
        grads = [np.zeros(d.size) for d in degs]
        dp = (D * n)(*[_dptr(d) for d in degs])
        gp = (D * n)(*[_dptr(g) for g in grads])
        dl = (c_size_t * n)(*[d.size for d in degs])
        ri = None if ref_index is None else (c_size_t * n)(*[int(k) for k in ref_index])
        out = (POINTER(Result) * n)()
        st = (c_int32 * n)()
        with self._lock:
            rc = lib.visqol_hip_measure_prepared_grad(self._h, prep._h, n, ri, dp, dl, out, st, gp)
        return rc, self._take(out, n), list(st), grads

    @staticmethod
    def _take(out, n):
        res = []
        for i in range(n):
            if out[i]:
                res.append(result_dict(out[i].contents))
                lib.visqol_result_free(out[i])
            else:
                res.append(None)
        return res

    PHASES = ("phase_a", "plan", "phase_b", "match_stats", "match_decode", "realign_lags", "realign_stats", "aggregate")

    def gammatone_tuning(self, sample_rate: int, kind: int = 0) -> dict:
        """Spectrogram block shape chosen for this context's precision and
        mode at `sample_rate` (kind 0: whole signals, 1: realignment
        slices): {"chosen": index or -1, "ms": per-shape times, "shapes": [...]}."""
        count = gammatone_shapes()
        chosen = c_int(-1)
        ms = (ctypes.c_float * len(count))()
        _check_count(lib.visqol_hip_gammatone_tuning(self._h, sample_rate, kind, ctypes.byref(chosen), ms, len(count)))
        return {"chosen": chosen.value, "ms": [round(float(x), 4) for x in ms] if chosen.value >= 0 else [],
                "shapes": count}

    def profile_phases(self) -> dict:
        """Host wall time (ms) of each stretch of the batch calls since the last reset."""
        buf = (c_double * len(self.PHASES))()
        rc = lib.visqol_hip_profile_phases(self._h, buf, len(self.PHASES))
        if rc < 0:
            _check(rc)
        return {k: buf[i] for i, k in enumerate(self.PHASES)}

    def measure_batch(self, refs, degs, sample_rate: int, num_threads: int = 0, convert: bool = True):
        """(first error code, [result dict or None], [status]) as _native.py:215-225.

        convert=False returns only the MOS values (NaN for a failed pair)."""
        refs = [_f64(r) for r in refs]
        degs = [_f64(d) for d in degs]
        n = len(refs)
        D = POINTER(c_double)
        rp = (D * n)(*[_dptr(r) for r in refs])
        dp = (D * n)(*[_dptr(d) for d in degs])
        rl = (c_size_t * n)(*[r.size for r in refs])
        dl = (c_size_t * n)(*[d.size for d in degs])
        out = (POINTER(Result) * n)()
        st = (c_int32 * n)()
        with self._lock:
            rc = lib.visqol_measure_batch(self._h, n, rp, rl, dp, dl, sample_rate, 0, out, st)
        if not convert:
            mos = np.full(n, np.nan)
            for i in range(n):
                if out[i]:
                    mos[i] = out[i].contents.moslqo
                    lib.visqol_result_free(out[i])
            return rc, mos, list(st)
        results = []
        for i in range(n):
            if out[i]:
                results.append(result_dict(out[i].contents))
                lib.visqol_result_free(out[i])
            else:
                results.append(None)
        return rc, results, list(st)


class DeviceBuffer:
    """A device allocation holding one f64 signal (visqol_hip_alloc)."""

    def __init__(self, x):
        x = _f64(x)
        self.size = x.size
        p = c_void_p()
        _check(lib.visqol_hip_alloc(max(1, x.nbytes), ctypes.byref(p)))
        self.ptr = p
        _check(lib.visqol_hip_upload(p, x.ctypes.data_as(c_void_p), x.nbytes))

    def free(self):
        if self.ptr:
            lib.visqol_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def measure_batch_device(ctx: Context, refs: list[DeviceBuffer], degs: list[DeviceBuffer], sample_rate: int,
                         convert: bool = True):
    """visqol_hip_measure_batch_device over device-resident signals.

    convert=False reads only each result's MOS (NaN for a failed pair)
    before freeing it, as a C caller that keeps the scores would; the
    results list is then a float64 array."""
    n = len(refs)
    rp = (c_void_p * n)(*[r.ptr for r in refs])
    dp = (c_void_p * n)(*[d.ptr for d in degs])
    rl = (c_size_t * n)(*[r.size for r in refs])
    dl = (c_size_t * n)(*[d.size for d in degs])
    out = (POINTER(Result) * n)()
    st = (c_int32 * n)()
    with ctx._lock:
        rc = lib.visqol_hip_measure_batch_device(ctx._h, n, rp, rl, dp, dl, sample_rate, out, st)
    if not convert:
        mos = np.full(n, np.nan)
        for i in range(n):
            if out[i]:
                mos[i] = out[i].contents.moslqo
                lib.visqol_result_free(out[i])
        return rc, mos, list(st)
    results = []
    for i in range(n):
        if out[i]:
            results.append(result_dict(out[i].contents))
            lib.visqol_result_free(out[i])
        else:
            results.append(None)
    return rc, results, list(st)


def measure_mixed_batch_device(ctxs: list[Context], ctx_index, refs: list[DeviceBuffer], degs: list[DeviceBuffer],
                               sample_rates, convert: bool = True):
    """visqol_hip_measure_mixed_batch_device: pair i scored by
    ctxs[ctx_index[i]] at sample_rates[i], the contexts' batches at once.
    Returns (rc, results, status) as measure_batch_device."""
    n = len(refs)
    hs = (c_void_p * len(ctxs))(*[c._h for c in ctxs])
    ci = (c_uint32 * n)(*[int(k) for k in ctx_index])
    rp = (c_void_p * n)(*[r.ptr for r in refs])
    dp = (c_void_p * n)(*[d.ptr for d in degs])
    rl = (c_size_t * n)(*[r.size for r in refs])
    dl = (c_size_t * n)(*[d.size for d in degs])
    sr = (c_uint32 * n)(*[int(x) for x in sample_rates])
    out = (POINTER(Result) * n)()
    st = (c_int32 * n)()
    locks = sorted({id(c): c._lock for c in ctxs}.items())
    for _, lk in locks:
        lk.acquire()
    try:
        rc = lib.visqol_hip_measure_mixed_batch_device(hs, len(ctxs), n, ci, rp, rl, dp, dl, sr, out, st)
    finally:
        for _, lk in reversed(locks):
            lk.release()
    if not convert:
        mos = np.full(n, np.nan)
        for i in range(n):
            if out[i]:
                mos[i] = out[i].contents.moslqo
                lib.visqol_result_free(out[i])
        return rc, mos, list(st)
    results = []
    for i in range(n):
        if out[i]:
            results.append(result_dict(out[i].contents))
            lib.visqol_result_free(out[i])
        else:
            results.append(None)
    return rc, results, list(st)


class Prepared:
    """visqol_hip_prepared: freed when collected or by free()."""

    def __init__(self, h, n: int):
        self._h = h
        self.count = n

    def free(self):
        if self._h:
            lib.visqol_hip_prepared_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """visqol_hip_server: continuous batching over a stream of requests.

    submit() copies the pair to the device and returns a ticket (blocking
    while the admission budget is full); wait(ticket) returns
    (status, result dict or None, latency_ms from submit to completion).
    Both may be called from several threads."""

    def __init__(self, precision: str | int = "fast", workers: int = 3, max_batch: int = 256,
                 slab_bytes: int = 16 << 30, max_inflight_bytes: int = 96 << 30, **opts):
        o = default_options()
        o.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.backend = BACKENDS["hip"]
        for k, v in opts.items():
            setattr(o, k, int(v))
        c = ServerConfig()
        lib.visqol_hip_server_config_default(ctypes.byref(c))
        c.workers, c.max_batch, c.slab_bytes, c.max_inflight_bytes = workers, max_batch, slab_bytes, max_inflight_bytes
        h = c_void_p()
        _check(lib.visqol_hip_server_create(ctypes.byref(o), ctypes.byref(c), ctypes.byref(h)))
        self._h = h

    def submit(self, mode: int, ref, deg, sample_rate: int) -> int:
        ref, deg = _f64(ref), _f64(deg)
        t = ctypes.c_uint64()
        _check(lib.visqol_hip_server_submit(self._h, mode, _dptr(ref), ref.size, _dptr(deg), deg.size, sample_rate,
                                            ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, convert: bool = True):
        out = POINTER(Result)()
        lat = c_double()
        rc = lib.visqol_hip_server_wait(self._h, ticket, ctypes.byref(out), ctypes.byref(lat))
        res = None
        if out:
            res = result_dict(out.contents) if convert else out.contents.moslqo
            lib.visqol_result_free(out)
        return rc, res, lat.value

    def wait_timed(self, ticket: int, convert: bool = True):
        """(status, result dict or MOS, timing dict): where the request's
        time went (admission, upload, queue, batch run, total ms) and the
        size of the batch it ran in."""
        out = POINTER(Result)()
        t = RequestTiming()
        rc = lib.visqol_hip_server_wait_timed(self._h, ticket, ctypes.byref(out), ctypes.byref(t))
        res = None
        if out:
            res = result_dict(out.contents) if convert else out.contents.moslqo
            lib.visqol_result_free(out)
        timing = {k: getattr(t, k) for k, _ in RequestTiming._fields_}
        return rc, res, timing

    def stats(self) -> dict:
        b, p, f = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib.visqol_hip_server_stats(self._h, ctypes.byref(b), ctypes.byref(p), ctypes.byref(f)))
        return dict(batches=b.value, pairs=p.value, inflight_bytes=f.value)

    def close(self):
        if self._h:
            lib.visqol_hip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def spectrogram(signal, sample_rate: int, mode: int = MODE_AUDIO, precision: str = "exact") -> np.ndarray:
    x = _f64(signal)
    nb = 32 if mode == 0 else 21
    hop = int(int(np.floor(sample_rate * 0.08 + 0.5)) * 0.25)
    out = np.zeros(nb * (x.size // max(1, hop) + 1))
    nf = c_size_t()
    _check(lib.visqol_hip_spectrogram(mode, sample_rate, PRECISIONS[precision], _dptr(x), x.size, _dptr(out),
                                      ctypes.byref(nf)))
    return out[: nb * nf.value].reshape(nb, nf.value)


def _check_count(rc: int) -> int:
    if rc < 0:
        _check(rc)
    return rc


def lattice_predict(model: bytes, fvnsim, fvnsim10, fstdnsim, fvdegenergy) -> np.ndarray:
    """The lattice mapping of [count, nb] feature arrays (host)."""
    f = [np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64) for x in (fvnsim, fvnsim10, fstdnsim, fvdegenergy)]
    count, nb = f[0].shape
    y = np.zeros(count)
    _check(lib.visqol_hip_lattice_predict(model, len(model), nb, count, *[_dptr(x) for x in f], _dptr(y)))
    return y


def frame_grad(mode: int, sample_rate: int, frame, g, lo: int = 0, hi: int | None = None) -> np.ndarray:
    """∂(Σ_b g_b · rms_b)/∂frame on the device for one spectrogram frame."""
    frame = _f64(frame)
    g = _f64(g)
    hi = frame.size if hi is None else hi
    out = np.zeros(frame.size)
    _check(lib.visqol_hip_frame_grad(mode, sample_rate, _dptr(frame), frame.size, _dptr(g), lo, hi, _dptr(out)))
    return out


def patch_similarity_grad(ref, deg, g_means=None, g_stddevs=None, g_energy=None):
    """(∂/∂ref cells, ∂/∂deg cells) of Σ g·(band means, stddevs, deg energy)
    of measurePatchSimilarity on the [nrows, ncols] patches."""
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    deg = np.ascontiguousarray(deg, dtype=np.float64)
    nrows, ncols = ref.shape
    opt = [None if g is None else _f64(g) for g in (g_means, g_stddevs, g_energy)]
    gr, gd = np.zeros_like(ref), np.zeros_like(deg)
    _check(lib.visqol_hip_patch_similarity_grad(_dptr(ref), _dptr(deg), nrows, ncols,
                                                *[None if g is None else _dptr(g) for g in opt], _dptr(gr), _dptr(gd)))
    return gr, gd


def prepare_grad(raw_ref, raw_deg, g_ref_prep, g_deg_prep):
    """prepareForComparison's slopes taken back to the raw [nb, nf] spectrograms."""
    raw_ref = np.ascontiguousarray(raw_ref, dtype=np.float64)
    raw_deg = np.ascontiguousarray(raw_deg, dtype=np.float64)
    nb, nfr = raw_ref.shape
    nfd = raw_deg.shape[1]
    g1, g2 = _f64(np.ascontiguousarray(g_ref_prep)), _f64(np.ascontiguousarray(g_deg_prep))
    gr, gd = np.zeros_like(raw_ref), np.zeros_like(raw_deg)
    _check(lib.visqol_hip_prepare_grad(_dptr(raw_ref), _dptr(raw_deg), nb, nfr, nfd, _dptr(g1), _dptr(g2), _dptr(gr),
                                       _dptr(gd)))
    return gr, gd


def gammatone_shapes() -> list[dict]:
    """The spectrogram kernel's block shapes (64 frames per block each)."""
    out = []
    w, b, c = c_int(), c_int(), c_int()
    n = _check_count(lib.visqol_hip_gammatone_shape(0, ctypes.byref(w), ctypes.byref(b), ctypes.byref(c)))
    for v in range(n):
        _check_count(lib.visqol_hip_gammatone_shape(v, ctypes.byref(w), ctypes.byref(b), ctypes.byref(c)))
        out.append({"waves": w.value, "bands_per_lane": b.value, "chunk": c.value, "frames_per_block": 64,
                    "bands_per_block": w.value * b.value})
    return out


def decode_wav(data: bytes) -> tuple[np.ndarray, int]:
    """visqol_hip_decode_wav: (mono f64 samples, sample rate) of WAV bytes;
    raises NativeError with the file entry points' code on a bad file."""
    n = c_size_t(0)
    sr = c_uint32(0)
    _check(lib.visqol_hip_decode_wav(data, len(data), None, 0, ctypes.byref(n), ctypes.byref(sr)))
    out = np.empty(n.value, dtype=np.float64)
    _check(lib.visqol_hip_decode_wav(data, len(data), _dptr(out), out.size, ctypes.byref(n), ctypes.byref(sr)))
    return out, sr.value


def filters(sample_rate: int, mode: int = MODE_AUDIO):
    nb = 32 if mode == 0 else 21
    b = np.zeros(8 * nb)
    cf = np.zeros(nb)
    _check(lib.visqol_hip_filters(mode, sample_rate, _dptr(b), _dptr(cf)))
    return b.reshape(nb, 8), cf


SUM_PLAIN, SUM_SQ, SUM_ALTERNATING, SUM_SQ_FROM_LAG = 0, 1, 2, 3
SUM_SERIAL, SUM_SPECULATIVE, SUM_AS_SCORED = 0, 1, 2


def reference_sum(kind: int, x, plain_sum: float = 0.0, lag: int = 0, half: int = 0, form: int = SUM_SERIAL) -> float:
    """visqol_hip_sum: one reference-order f64 sum of x on the device."""
    x = _f64(x)
    out = c_double()
    _check(lib.visqol_hip_sum(kind, _dptr(x), x.size, plain_sum, lag, half, form, ctypes.byref(out)))
    return out.value


def best_lag(a, b) -> int:
    a, b = _f64(a), _f64(b)
    lag = c_int64()
    _check(lib.visqol_hip_best_lag(_dptr(a), a.size, _dptr(b), b.size, ctypes.byref(lag)))
    return lag.value


def best_lag_noalloc(dev_a: "DeviceBuffer", dev_b: "DeviceBuffer") -> int:
    """visqol_hip_best_lag_noalloc on two device-resident signals."""
    lag = c_int64()
    _check(lib.visqol_hip_best_lag_noalloc(dev_a.ptr, dev_a.size, dev_b.ptr, dev_b.size, ctypes.byref(lag)))
    return lag.value


def xcorr(a, b) -> np.ndarray:
    """Re of the circular cross-correlation of the upper envelopes (device)."""
    a, b = _f64(a), _f64(b)
    biggest = max(a.size, b.size)
    n = 1
    while n < max(1, 2 * biggest - 1):
        n <<= 1
    out = np.zeros(n if biggest else 0)
    got = c_size_t()
    _check(lib.visqol_hip_xcorr(_dptr(a), a.size, _dptr(b), b.size, _dptr(out), out.size, ctypes.byref(got)))
    return out[: got.value]


def xcorr32(a, b) -> np.ndarray:
    """The throughput mode's f32 correlation of the upper envelopes (Re,
    times n); empty when n is outside 2^14..2^30."""
    a, b = _f64(a), _f64(b)
    biggest = max(a.size, b.size)
    n = 1
    while n < max(1, 2 * biggest - 1):
        n <<= 1
    out = np.zeros(n, dtype=np.float32)
    got = c_size_t()
    _check(lib.visqol_hip_xcorr32(_dptr(a), a.size, _dptr(b), b.size,
                                  out.ctypes.data_as(POINTER(ctypes.c_float)), out.size, ctypes.byref(got)))
    return out[: got.value]


def best_lag_fast(a, b, cap: int = -1):
    """(lag, nominated) of the throughput mode's lag search on raw signals."""
    a, b = _f64(a), _f64(b)
    lag, nom = c_int64(), c_int()
    _check(lib.visqol_hip_best_lag_fast(_dptr(a), a.size, _dptr(b), b.size, cap, ctypes.byref(lag), ctypes.byref(nom)))
    return lag.value, bool(nom.value)


def fft(z: np.ndarray, dit: bool = False) -> np.ndarray:
    """Device DIF forward (bit-reversed out) or DIT inverse of complex z, unscaled."""
    z = np.asarray(z, dtype=np.complex128)
    buf = np.ascontiguousarray(np.stack([z.real, z.imag], axis=1).ravel())
    _check(lib.visqol_hip_fft(_dptr(buf), z.size, int(dit)))
    return buf[0::2] + 1j * buf[1::2]


def math_probe(op: int, a, b=None) -> np.ndarray:
    a = _f64(a)
    bb = _f64(b) if b is not None else None
    out = np.zeros(a.size)
    _check(lib.visqol_hip_math_probe(op, _dptr(a), _dptr(bb) if bb is not None else None, _dptr(out), a.size))
    return out
