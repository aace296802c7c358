"""ctypes binding to the CPU f64 oracle (oracle/build/libvo.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The shipped package never
imports this module.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint32, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libvo.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


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


class Band(ctypes.Structure):
    _fields_ = [(n, c_double) for n in ("a0", "a11", "a12", "a13", "a14", "b1", "b2", "gain")]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        D = POINTER(c_double)
        S = POINTER(c_size_t)
        sig = {
            "vo_last_error": (c_char_p, []),
            "vo_options_default": (None, [POINTER(Options)]),
            "vo_create": (c_int, [POINTER(Options), POINTER(c_void_p)]),
            "vo_destroy": (None, [c_void_p]),
            "vo_measure_trace": (c_int, [c_void_p, D, c_size_t, D, c_size_t, c_uint32, POINTER(POINTER(Result)),
                                         POINTER(Trace), D, D, D]),
            "vo_measure_batch": (c_int, [c_void_p, c_size_t, POINTER(D), S, POINTER(D), S, c_uint32, c_uint32,
                                         POINTER(POINTER(Result)), POINTER(c_int32)]),
            "vo_result_free": (None, [POINTER(Result)]),
            "vo_build_result_from": (c_int, [c_void_p, c_size_t, D, D, D, D, D, c_double, POINTER(POINTER(Result))]),
            "vo_log10": (c_double, [c_double]),
            "vo_center_freqs_desc": (None, [c_size_t, c_double, c_double, D]),
            "vo_filters": (None, [c_uint32, c_size_t, c_double, c_double, POINTER(Band), D]),
            "vo_filter_band": (None, [POINTER(Band), D, c_size_t, D]),
            "vo_band_rms_fresh": (c_double, [POINTER(Band), D, c_size_t]),
            "vo_hann": (None, [c_size_t, D]),
            "vo_num_frames": (c_size_t, [c_size_t, c_size_t, c_size_t]),
            "vo_spectrogram": (c_int, [POINTER(Band), c_size_t, D, c_size_t, c_size_t, c_size_t, D, D]),
            "vo_db_floor": (None, [D, c_size_t, c_double]),
            "vo_raise_floor_per_frame": (None, [D, D, c_size_t, c_size_t, c_size_t, c_double]),
            "vo_subtract_joint_min": (None, [D, c_size_t, D, c_size_t]),
            "vo_prepare_for_comparison": (None, [D, D, c_size_t, c_size_t, c_size_t]),
            "vo_upper_envelope": (None, [D, c_size_t, D]),
            "vo_fft_size_for": (c_size_t, [c_size_t]),
            "vo_fft_dif_forward": (None, [D, c_size_t]),
            "vo_fft_dit_inverse": (None, [D, c_size_t]),
            "vo_fft_natural": (None, [D, c_size_t, c_int]),
            "vo_best_lag": (c_int64, [D, c_size_t, D, c_size_t]),
            "vo_xcorr_re": (c_int, [D, c_size_t, D, c_size_t, D]),
            "vo_best_lag_reference": (c_int64, [D, c_size_t, D, c_size_t]),
            "vo_best_lag_direct": (c_int64, [D, c_size_t, D, c_size_t]),
            "vo_spl": (c_double, [D, c_size_t]),
            "vo_scale_to_match_spl": (None, [D, c_size_t, D, c_size_t]),
            "vo_conv2d3x3": (None, [D, D, c_size_t, c_size_t]),
            "vo_patch_similarity": (None, [D, D, c_size_t, c_size_t, D, D, D, D]),
            "vo_similarity_scalar": (c_double, [D, D, c_size_t, c_size_t]),
            "vo_similarity_table": (None, [D, c_size_t, D, c_size_t, c_size_t, c_size_t, S, S, S, c_size_t, D]),
            "vo_ref_patch_indices": (c_size_t, [c_size_t, c_size_t, S, c_size_t]),
            "vo_calc_max_num_patches": (c_size_t, [S, c_size_t, c_size_t, c_size_t]),
            "vo_vad_patch_indices": (c_size_t, [D, c_size_t, c_size_t, c_size_t, c_size_t, c_double, S, c_size_t]),
            "vo_dp_search": (c_double, [D, c_size_t, c_size_t, S, S, S, POINTER(c_uint8)]),
            "vo_find_optimal_patches": (c_int, [D, S, c_size_t, D, c_size_t, c_size_t, c_size_t, c_size_t, c_double,
                                                c_int32, D]),
            "vo_slice_layout": (None, [c_size_t, c_uint32, c_double, c_double, S, S, S]),
            "vo_lag_samples_back": (c_int64, [c_int64, c_uint32]),
            "vo_trim_geometry": (None, [c_size_t, c_size_t, c_int64, c_uint32, S, S, POINTER(c_int64), D]),
            "vo_svr_predict": (c_int, [c_char_p, c_size_t, D, c_size_t, D]),
            "vo_lattice_predict": (c_int, [c_char_p, c_size_t, D, D, D, D, c_size_t, D]),
            "vo_lattice_op_count": (c_int, [c_char_p, c_size_t]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(POINTER(c_double))


def sptr(a: np.ndarray):
    assert a.dtype == np.uintp and a.flags.c_contiguous
    return a.ctypes.data_as(POINTER(c_size_t))


def f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def result_to_dict(r: Result) -> dict:
    nb, npch = r.num_bands, r.num_patches
    arr = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0)
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


class Oracle:
    """The oracle's visqol.h-shaped API (vo_create / vo_measure / vo_measure_batch)."""

    def __init__(self, mode: int = 0, use_lattice_model: bool = True, search_window_radius: int = 60,
                 disable_global_alignment: bool = False, disable_realignment: bool = False,
                 use_unscaled_speech_mos_mapping: bool = False, model_path: str | None = None):
        L = lib()
        o = Options()
        L.vo_options_default(ctypes.byref(o))
        o.mode = mode
        o.use_lattice_model = int(use_lattice_model)
        o.search_window_radius = search_window_radius
        o.disable_global_alignment = int(disable_global_alignment)
        o.disable_realignment = int(disable_realignment)
        o.use_unscaled_speech_mos_mapping = int(use_unscaled_speech_mos_mapping)
        o.model_path = model_path.encode() if model_path else None
        h = c_void_p()
        rc = L.vo_create(ctypes.byref(o), ctypes.byref(h))
        if rc:
            raise OracleError(rc, L.vo_last_error().decode())
        self._h = h
        self.mode = mode

    def close(self):
        if self._h:
            lib().vo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def measure(self, ref, deg, sample_rate: int, want_spec: bool = False):
        """Returns (result dict, trace dict); raises OracleError on failure."""
        L = lib()
        ref, deg = f64(ref), f64(deg)
        out = POINTER(Result)()
        tr = Trace()
        rs = ds = sims = None
        if want_spec:
            nb = 32 if self.mode == 0 else 21
            hop = int(int(np.floor(sample_rate * 0.08 + 0.5)) * 0.25)
            patch = 30 if self.mode == 0 else 20
            nfr = len(ref) // hop + 1
            nfd = (len(deg) + len(ref) // 2) // hop + 1
            rs = np.zeros(nb * nfr, np.float64)
            ds = np.zeros(nb * nfd, np.float64)
            sims = np.zeros((nfr // patch + 2) * nfd, np.float64)
        rc = L.vo_measure_trace(self._h, dptr(ref), len(ref), dptr(deg), len(deg), sample_rate, ctypes.byref(out),
                                ctypes.byref(tr), dptr(rs) if rs is not None else None,
                                dptr(ds) if ds is not None else None, dptr(sims) if sims is not None else None)
        if rc:
            raise OracleError(rc, L.vo_last_error().decode())
        res = result_to_dict(out.contents)
        L.vo_result_free(out)
        np_ = tr.num_patches
        trace = dict(lag=tr.lag, spl_factor=tr.spl_factor, nf_ref=tr.nf_ref, nf_deg=tr.nf_deg, num_patches=np_,
                     ref_indices=list(tr.ref_indices[:np_]), realign_lag=list(tr.realign_lag[:np_]),
                     realign_frames=list(tr.realign_frames[:np_]),
                     realign_accepted=list(tr.realign_accepted[:np_]), match_valid=list(tr.match_valid[:np_]),
                     dp_offset=list(tr.dp_offset[:np_]))
        if want_spec:
            nb = 32 if self.mode == 0 else 21
            trace["ref_spec"] = rs[: nb * tr.nf_ref].reshape(nb, tr.nf_ref).copy()
            trace["deg_spec"] = ds[: nb * tr.nf_deg].reshape(nb, tr.nf_deg).copy()
            trace["sims"] = sims[: np_ * tr.nf_deg].reshape(np_, tr.nf_deg).copy()
        return res, trace

    def build_result(self, means, stddevs, energy, ref_start, ref_end, frame_duration: float) -> dict:
        """buildResult + mapping over given matches (arrays [n][nb] and [n])."""
        L = lib()
        m, s, e = (np.ascontiguousarray(x, dtype=np.float64) for x in (means, stddevs, energy))
        r0, r1 = f64(ref_start), f64(ref_end)
        out = POINTER(Result)()
        rc = L.vo_build_result_from(self._h, m.shape[0], dptr(m), dptr(s), dptr(e), dptr(r0), dptr(r1),
                                    frame_duration, ctypes.byref(out))
        if rc:
            raise OracleError(rc, L.vo_last_error().decode())
        res = result_to_dict(out.contents)
        L.vo_result_free(out)
        return res

    def measure_batch(self, refs, degs, sample_rate: int, num_threads: int = 0):
        L = lib()
        refs = [f64(r) for r in refs]
        degs = [f64(d) for d in degs]
        n = len(refs)
        D = POINTER(c_double)
        rp = (D * n)(*[dptr(r) for r in refs])
        dp = (D * n)(*[dptr(d) for d in degs])
        rl = np.array([len(r) for r in refs], np.uintp)
        dl = np.array([len(d) for d in degs], np.uintp)
        out = (POINTER(Result) * n)()
        st = (c_int32 * n)()
        rc = L.vo_measure_batch(self._h, n, rp, sptr(rl), dp, sptr(dl), sample_rate, num_threads, out, st)
        results = []
        for i in range(n):
            if out[i]:
                results.append(result_to_dict(out[i].contents))
                L.vo_result_free(out[i])
            else:
                results.append(None)
        return rc, results, list(st)


def filters(sample_rate: int, nbands: int, low: float, high: float):
    bands = (Band * nbands)()
    cf = np.zeros(nbands)
    lib().vo_filters(sample_rate, nbands, low, high, bands, dptr(cf))
    return bands, cf


def hann(n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().vo_hann(n, dptr(out))
    return out


def spectrogram(signal, sample_rate: int, nbands: int, high: float | None = None):
    """Linear gammatone spectrogram [bands][frames] (accel.zig:116-124)."""
    sig = f64(signal)
    window = int(np.floor(sample_rate * 0.08 + 0.5))
    hop = int(window * 0.25)
    bands, _ = filters(sample_rate, nbands, 50.0, high if high else sample_rate / 2.0)
    h = hann(window)
    nf = lib().vo_num_frames(len(sig), window, hop)
    out = np.zeros(nbands * nf)
    rc = lib().vo_spectrogram(bands, nbands, dptr(sig), len(sig), window, hop, dptr(h), dptr(out))
    if rc:
        raise OracleError(rc, "spectrogram")
    return out.reshape(nbands, nf)
