"""visqol_hip_best_lag_noalloc: the allocation-free lag search (GPU).

Called through the C ABI on device-resident signals, twice in a row with
the second pair smaller (the shared static scratch keeps the first call's
contents), and once past its size limit.  Each lag must equal the existing
entry point's, visqol_hip_best_lag on the same signals as host arrays (the
exact f64 search, itself bit-identical to the oracle), and the oracle's.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_lag(vo, a, b):
    ea, eb = np.zeros(a.size), np.zeros(b.size)
    vo.lib().vo_upper_envelope(vo.dptr(vo.f64(a)), a.size, vo.dptr(ea))
    vo.lib().vo_upper_envelope(vo.dptr(vo.f64(b)), b.size, vo.dptr(eb))
    return vo.lib().vo_best_lag(vo.dptr(ea), a.size, vo.dptr(eb), b.size)


def test_noalloc_lag_equals_best_lag_two_calls():
    from akvisqol_amd import native, synth
    from oracle import vo

    cases = [synth.audio_pair(31, duration=10.0, delay=1500, snr_db=12.0),  # n = 2^20, the limit
             synth.audio_pair(32, duration=1.5, delay=64, snr_db=30.0)]  # then a smaller transform
    for k, (a, b) in enumerate(cases):
        say(f"noalloc call {k}: {a.size} / {b.size} samples")
        da, db = native.DeviceBuffer(a), native.DeviceBuffer(b)
        try:
            got = native.best_lag_noalloc(da, db)
        finally:
            da.free()
            db.free()
        want = native.best_lag(a, b)
        say(f"  lag {got}, visqol_hip_best_lag {want}")
        assert got == want == _oracle_lag(vo, a, b)


def test_noalloc_rejects_oversized_signals():
    from akvisqol_amd import native

    x = np.zeros((1 << 19) + 1)
    d = native.DeviceBuffer(x)
    try:
        with pytest.raises(native.NativeError) as e:
            native.best_lag_noalloc(d, d)
    finally:
        d.free()
    assert e.value.code == native.E_INVALID_ARGUMENT
    assert "limit" in str(e.value)
