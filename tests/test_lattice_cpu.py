"""The library's TF-Lattice interpreter against the oracle's, bit for bit.

The speech quality mapping (lattice.zig:49-71, tflite.zig) runs on the host
in both: a flatbuffer graph of ~5000 f32 ops per prediction.  The library's
interpreter (models.cpp) is checked here against the oracle's independent
one (oracle/vo_map.c) on random feature vectors spanning and exceeding the
ranges real signals produce, plus edge values, through the host-only probe
visqol_hip_lattice_predict.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

from oracle import vo
from oracle.vo import dptr, f64

MODEL = (Path(__file__).resolve().parents[1] / "akvisqol_amd" / "models" / "lattice_speech.tflite").read_bytes()
NB = 21


@pytest.fixture(scope="module")
def native():
    from akvisqol_amd import native as n

    return n


def oracle_predict(a, b, c, d):
    y = ctypes.c_double()
    assert vo.lib().vo_lattice_predict(MODEL, len(MODEL), dptr(f64(a)), dptr(f64(b)), dptr(f64(c)), dptr(f64(d)),
                                       NB, ctypes.byref(y)) == 0
    return y.value


def feature_sets(seed, count):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.2, 1.05, (count, NB))
    b = a * rng.uniform(0.3, 1.0, (count, NB))
    c = rng.uniform(0.0, 0.3, (count, NB))
    d = rng.uniform(-5.0, 60.0, (count, NB))
    # edge rows: all zero, all one, beyond the calibrators' ranges
    a[0], b[0], c[0], d[0] = 0.0, 0.0, 0.0, 0.0
    a[1], b[1], c[1], d[1] = 1.0, 1.0, 0.0, 20.0
    a[2], b[2], c[2], d[2] = 2.0, -1.0, 5.0, 500.0
    a[3], b[3], c[3], d[3] = -3.0, 3.0, -1.0, -100.0
    return a, b, c, d


def test_lattice_matches_oracle_bit_for_bit(native):
    a, b, c, d = feature_sets(11, 160)
    got = native.lattice_predict(MODEL, a, b, c, d)
    for k in range(a.shape[0]):
        assert got[k] == oracle_predict(a[k], b[k], c[k], d[k]), k


def test_lattice_batch_rows_independent(native):
    a, b, c, d = feature_sets(12, 24)
    whole = native.lattice_predict(MODEL, a, b, c, d)
    for k in (0, 5, 23):
        one = native.lattice_predict(MODEL, a[k], b[k], c[k], d[k])
        assert one[0] == whole[k]


def test_lattice_rejects_truncated_model(native):
    a, b, c, d = feature_sets(13, 4)
    with pytest.raises(native.NativeError):
        native.lattice_predict(MODEL[: len(MODEL) // 2], a, b, c, d)


def test_plan_equals_interpreter(native, monkeypatch):
    # the execution plan built at load (folded constants, one arena) and the
    # graph interpreter it replaces give the same bits
    a, b, c, d = feature_sets(14, 96)
    planned = native.lattice_predict(MODEL, a, b, c, d)
    monkeypatch.setenv("VISQOL_HIP_LATTICE_PLAN", "0")
    interpreted = native.lattice_predict(MODEL, a, b, c, d)
    assert np.array_equal(planned, interpreted)
