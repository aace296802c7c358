"""The visqol command-line tool without a GPU: flags (main.zig:10-33,
61-104), the batch CSV (main.zig:142-157), the results-CSV number format
(main.zig:187-196: the shortest decimal that reads back to the score), and
the tool's exits.  The GPU test (test_gpu_cli.py) checks its scores."""
import os
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "akvisqol_amd" / "visqol"
pytestmark = pytest.mark.skipif(not TOOL.exists(), reason="the visqol tool is not built (run __graft_entry__.build())")


@pytest.fixture(scope="module")
def check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("cli") / "cli_args_check"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'akvisqol_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "cli_args_check.cpp"), "-o", str(exe)], check=True)
    return exe


def _args(check, *argv):
    out = subprocess.run([str(check), "args", *argv], check=True, capture_output=True, text=True).stdout
    return dict((line.split(" ", 1) + [""])[:2] for line in out.splitlines())


def test_defaults(check):
    a = _args(check)
    assert a["status"] == "0" and a["radius"] == "60" and a["lattice"] == "1" and a["speech"] == "0"
    assert a["backend"] == "0" and a["precision"] == "3" and a["threads"] == "0"


def test_every_flag(check):
    a = _args(check, "--reference_file", "r.wav", "--degraded_file", "d.wav", "--batch_input_csv", "p.csv",
              "--results_csv", "o.csv", "--similarity_to_quality_model", "m.txt", "--use_speech_mode",
              "--nouse_lattice_model", "--use_unscaled_speech_mos_mapping", "--search_window_radius", "17",
              "--disable_global_alignment", "--disable_realignment", "--verbose", "--backend", "hip",
              "--threads", "3", "--precision", "balanced")
    assert (a["ref"], a["deg"], a["csv"], a["results"], a["model"]) == ("r.wav", "d.wav", "p.csv", "o.csv", "m.txt")
    assert (a["speech"], a["lattice"], a["unscaled"], a["no_global"], a["no_realign"], a["verbose"]) == \
        ("1", "0", "1", "1", "1", "1")
    assert (a["radius"], a["backend"], a["threads"], a["precision"]) == ("17", "4", "3", "2")


@pytest.mark.parametrize("flag,want", [("--use_lattice_model", "1"), ("--use_lattice_model=true", "1"),
                                       ("--use_lattice_model=false", "0"), ("--nouse_lattice_model", "0")])
def test_lattice_spellings(check, flag, want):
    assert _args(check, "--nouse_lattice_model", flag)["lattice"] == want


@pytest.mark.parametrize("argv,status,err", [
    (["--search_window_radius", "x1"], "1", "InvalidCharacter"),
    (["--search_window_radius", "99999999999"], "1", "Overflow"),
    (["--threads"], "1", "MissingArgument"),
    (["--backend", "opencl"], "3", "unknown backend"),
    (["--precision", "double"], "3", "unknown precision"),
    (["--frobnicate"], "2", "--frobnicate"),
])
def test_bad_flags(check, argv, status, err):
    a = _args(check, *argv)
    assert a["status"] == status and a["err"] == err


def test_backend_names(check):
    got = [_args(check, "--backend", b)["backend"] for b in ("cpu", "metal", "cuda", "vulkan", "hip")]
    assert got == ["0", "1", "2", "3", "4"]


def test_batch_csv(check, tmp_path):
    p = tmp_path / "pairs.csv"
    p.write_bytes(b"reference,degraded\r\n a.wav , b.wav \r\n\r\nc.wav,d.wav,extra\nlonely\ne.wav,\n")
    out = subprocess.run([str(check), "csv", str(p)], check=True, capture_output=True, text=True).stdout
    assert out.splitlines() == ["[a.wav]|[b.wav]", "[c.wav]|[d.wav]", "[e.wav]|[]"]
    # a header only counts on the first row with two fields
    p.write_bytes(b"x.wav,y.wav\nreference,degraded\n")
    out = subprocess.run([str(check), "csv", str(p)], check=True, capture_output=True, text=True).stdout
    assert out.splitlines() == ["[x.wav]|[y.wav]", "[reference]|[degraded]"]
    p.write_bytes(b"no comma here\nreference,degraded\nr.wav,d.wav")
    out = subprocess.run([str(check), "csv", str(p)], check=True, capture_output=True, text=True).stdout
    assert out.splitlines() == ["[r.wav]|[d.wav]"]


def test_results_csv_numbers_read_back(check):
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(1.0, 5.0, 200), [1.0, 5.0, 4.30533, 2.5, 3.0000000000000004]])
    out = subprocess.run([str(check), "num", *[repr(float(x)) for x in xs]], check=True, capture_output=True,
                         text=True).stdout.split()
    for x, s in zip(xs, out):
        assert float(s) == x and "e" not in s
        # Python's repr is also the shortest round-trip text; whole numbers
        # print without a fraction ("5", as Zig's {d} does)
        assert s == repr(float(x)).removesuffix(".0")


def test_tool_exits_without_device(tmp_path):
    run = lambda *a: subprocess.run([str(TOOL), *a], capture_output=True, text=True)
    r = run("--help")
    assert r.returncode == 0 and r.stdout.startswith("usage: visqol --reference_file")
    r = run("--frobnicate")
    assert r.returncode == 2 and r.stdout.startswith("unknown flag: --frobnicate\n\nusage:")
    r = run("--backend", "opencl")
    assert r.returncode == 1 and r.stdout == "error: unknown backend\n"
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = run("--reference_file", "a.wav", "--degraded_file", "b.wav")
    assert r.returncode == 1 and "no HIP device" in r.stderr
