"""Builds akvisqol_amd/libvisqol.so for gfx950 with hipcc, and the visqol
command-line tool (akvisqol_amd/visqol) over it with the host compiler.

Every source compiles in parallel into _build/, then links into one shared
library next to this file, where the bundled models live.  Device code is
built with -ffp-contract=off (exact kernels round every operation as the
reference CPU path does; fast kernels spell out their fused multiply-adds)
and without packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libvisqol.so"
CLI = HERE / "visqol"  # the command-line tool (csrc/cli.cpp)
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["gammatone.hip", "signal.hip", "xcorr.hip", "nsim.hip", "adjoint.hip", "engine.cpp", "engine_grad.cpp",
           "host_plan.cpp", "models.cpp", "wav.cpp", "capi.cpp", "server.cpp", "grad.cpp",
           "noalloc.hip"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    # sin and cos of one angle stay two libm calls, as the reference's @cos
    # and @sin; fused into sincos they can differ in the last ulp.
    "-fno-builtin-sin", "-fno-builtin-cos",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src: str) -> Path:
    out = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", str(CSRC / src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    noise = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return out


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = min(len(SOURCES), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # the command-line tool, host code over the C ABI, next to the library
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", f"-I{HERE.parent / 'include'}", str(CSRC / "cli.cpp"), "-o", str(CLI) + ".tmp",
           str(LIB), "-Wl,-rpath,$ORIGIN", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the visqol tool failed:\n{r.stderr}")
    os.replace(str(CLI) + ".tmp", CLI)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
