"""The visqol command-line tool on the device: its scores equal
visqol_measure_files on the same WAV files (single pairs printed to six
decimals, batch results CSV exactly), in both modes and precisions, and its
batch rules hold (one sample rate per batch)."""
import subprocess
import sys
import wave
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "akvisqol_amd" / "visqol"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _wav(path, x, sr):
    q = np.clip(np.round(np.asarray(x) * 32768.0), -32768, 32767).astype("<i2")
    with wave.open(str(path), "wb") as w:
        w.setnchannels(1)
        w.setsampwidth(2)
        w.setframerate(sr)
        w.writeframes(q.tobytes())


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    from akvisqol_amd import synth

    d = tmp_path_factory.mktemp("wav")
    out = {}
    for k in range(3):
        r, g = synth.audio_pair(60 + k, duration=3.0 + k, delay=40 * k, snr_db=15.0 + 5 * k)
        _wav(d / f"a{k}_ref.wav", r, 48000)
        _wav(d / f"a{k}_deg.wav", g, 48000)
        out[f"a{k}"] = (d / f"a{k}_ref.wav", d / f"a{k}_deg.wav")
    r, g = synth.speech_pair(8, duration=4.0, snr_db=18.0)
    _wav(d / "s_ref.wav", r, 16000)
    _wav(d / "s_deg.wav", g, 16000)
    out["s"] = (d / "s_ref.wav", d / "s_deg.wav")
    out["dir"] = d
    return out


def _run(*args):
    r = subprocess.run([str(TOOL), *map(str, args)], capture_output=True, text=True, timeout=300)
    return r


@pytest.mark.parametrize("case,precision", [("a0", "exact"), ("a1", "fast"), ("s", "exact"), ("s", "fast")])
def test_single_pair_equals_measure_files(files, case, precision):
    from akvisqol_amd import native

    ref, deg = files[case]
    mode = 1 if case == "s" else 0
    args = ["--reference_file", ref, "--degraded_file", deg, "--precision", precision, "--backend", "hip"]
    if mode:
        args.append("--use_speech_mode")
    r = _run(*args)
    say(f"cli {case} {precision}: rc {r.returncode} {r.stdout.strip().splitlines()[-1] if r.stdout else r.stderr}")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == f"Reference Filepath:\t{ref}" and lines[1] == f"Degraded Filepath:\t{deg}"
    want = native.Context(mode, precision=precision).measure_files(ref, deg)
    assert lines[2] == f"MOS-LQO:\t\t{want['moslqo']:.6f}"


def test_verbose_lists_bands_and_patches(files):
    from akvisqol_amd import native

    ref, deg = files["a2"]
    r = _run("--reference_file", ref, "--degraded_file", deg, "--verbose")
    assert r.returncode == 0, r.stderr
    want = native.Context(0, precision="exact").measure_files(ref, deg)
    text = r.stdout
    assert f"VNSIM:\t\t\t{want['vnsim']:.6f}" in text
    bands = [ln for ln in text.splitlines() if ln.startswith("  ") and len(ln.split()) == 6 and ln.split()[0].isdigit()]
    assert len(bands) == 32 + len(want["patch_sims"])
    first = bands[0].split()
    assert float(first[2]) == pytest.approx(want["fvnsim"][0], abs=5e-7)


def test_batch_csv_and_results(files, tmp_path):
    from akvisqol_amd import native

    csv = tmp_path / "pairs.csv"
    keys = ["a0", "a1", "a2", "a0"]
    csv.write_text("reference,degraded\n" + "".join(f"{files[k][0]},{files[k][1]}\n" for k in keys))
    res = tmp_path / "out.csv"
    r = _run("--batch_input_csv", csv, "--results_csv", res, "--precision", "fast", "--backend", "hip")
    assert r.returncode == 0, r.stderr
    rows = res.read_text().splitlines()
    assert rows[0] == "reference,degraded,moslqo" and len(rows) == 1 + len(keys)
    ctx = native.Context(0, precision="fast")
    for k, row, line in zip(keys, rows[1:], r.stdout.splitlines()):
        ref, deg, mos = row.split(",")
        want = ctx.measure_files(files[k][0], files[k][1])["moslqo"]
        assert (ref, deg) == (str(files[k][0]), str(files[k][1]))
        assert float(mos) == want
        assert line == f"{ref} {deg} MOS-LQO: {want:.6f}"


def test_cpu_backend_is_exact(files):
    """--backend cpu (the default) scores in f64 whatever --precision says,
    as the reference's CPU backend does."""
    from akvisqol_amd import native

    ref, deg = files["a1"]
    r = _run("--reference_file", ref, "--degraded_file", deg, "--precision", "fast")
    assert r.returncode == 0, r.stderr
    want = native.Context(0, precision="exact").measure_files(ref, deg)
    assert r.stdout.splitlines()[2] == f"MOS-LQO:\t\t{want['moslqo']:.6f}"


def test_batch_rejects_mixed_sample_rates(files, tmp_path):
    csv = tmp_path / "mixed.csv"
    csv.write_text(f"{files['a0'][0]},{files['a0'][1]}\n{files['s'][0]},{files['s'][1]}\n")
    r = _run("--batch_input_csv", csv)
    assert r.returncode == 1
    assert r.stdout.strip() == "error: batch mode requires every file to share one sample rate"


def test_missing_file_fails(files, tmp_path):
    r = _run("--reference_file", tmp_path / "nope.wav", "--degraded_file", files["a0"][1])
    assert r.returncode == 1 and "FileNotFound" in r.stderr
