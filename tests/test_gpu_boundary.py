"""Reference callers reach the device unchanged (GPU).

tests/native/default_flow.c is written the way the reference's wrappers
call the C ABI (visqol_compat.cc:83-150, visqol_lib_py.py:76-77): default
options, backend VISQOL_BACKEND_CPU from an unset VISQOL_BACKEND.  It is
built here against this repository's include/visqol.h (the CPU suite proves
that header layout- and constant-identical to the reference's and builds
the same program against the reference header) and linked to libvisqol.so;
its scores must equal the CPU f64 oracle's bit for bit, since the CPU slot
runs the exact device path.
"""
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "akvisqol_amd" / "libvisqol.so"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def default_flow(tmp_path_factory):
    exe = tmp_path_factory.mktemp("flow") / "default_flow"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "native" / "default_flow.c"), "-o", str(exe), f"-L{LIB.parent}", "-lvisqol",
                    f"-Wl,-rpath,{LIB.parent}"], check=True)
    return exe


def _run(exe, tmp_path, mode, lattice, sr, ref, deg, backend_env=None):
    a, b = tmp_path / "ref.f64", tmp_path / "deg.f64"
    np.asarray(ref, "<f8").tofile(a)
    np.asarray(deg, "<f8").tofile(b)
    env = {k: v for k, v in os.environ.items() if k != "VISQOL_BACKEND"}
    if backend_env is not None:
        env["VISQOL_BACKEND"] = backend_env
    out = subprocess.run([str(exe), str(mode), str(int(lattice)), str(sr), str(a), str(b)], check=True,
                         capture_output=True, text=True, env=env, timeout=120).stdout
    return dict(line.split(" ", 1) for line in out.strip().splitlines())


@pytest.mark.parametrize("case", ["audio", "speech_lattice", "speech_exponential", "unknown_env"])
def test_default_flow_scores_equal_oracle(default_flow, tmp_path, case):
    from akvisqol_amd import synth
    from oracle import vo

    say(f"default flow (backend CPU, wrappers' flow): {case}")
    env = None
    if case in ("audio", "unknown_env"):
        ref, deg = synth.audio_pair(11, duration=4.0, delay=250, snr_db=18.0)
        mode, lattice, sr = 0, False, 48000
        env = "hip" if case == "unknown_env" else None  # names no reference backend: CPU
    else:
        ref, deg = synth.speech_pair(4, duration=5.0, snr_db=15.0, delay=90)
        mode, lattice, sr = 1, case == "speech_lattice", 16000
    got = _run(default_flow, tmp_path, mode, lattice, sr, ref, deg, env)
    assert int(got["backends"]) == (1 << 0) | (1 << 4)
    assert got["create"] == "0" and got["measure"] == "0", got
    want, _ = vo.Oracle(mode=mode, use_lattice_model=lattice).measure(ref, deg, sr)
    say(f"  MOS {float.fromhex(got['moslqo'])!r} oracle {want['moslqo']!r}")
    assert float.fromhex(got["moslqo"]) == want["moslqo"]
    assert float.fromhex(got["vnsim"]) == want["vnsim"]
    assert float.fromhex(got["lag"]) == want["alignment_lag_s"]
    assert [float.fromhex(v) for v in got["fvnsim"].split()] == list(want["fvnsim"])
    assert int(got["patches"]) == len(want["patch_sims"])


def test_cpu_backend_ignores_fast_precision():
    """VISQOL_BACKEND_CPU computes in f64 whatever the precision, as the
    reference's CPU backend does: a context asking for CPU + FAST scores
    bit-equal to the oracle."""
    from akvisqol_amd import native, synth
    from oracle import vo

    say("backend CPU with precision FAST -> exact scores")
    ref, deg = synth.audio_pair(3, duration=3.0, delay=40, snr_db=22.0)
    with native.Context(native.MODE_AUDIO, backend="cpu", precision="fast") as ctx:
        got = ctx.measure(ref, deg, 48000)
    want, _ = vo.Oracle(mode=0).measure(ref, deg, 48000)
    assert got["moslqo"] == want["moslqo"]
    assert np.array_equal(got["fvnsim"], want["fvnsim"])
