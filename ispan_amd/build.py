"""Build libispan_hip.so and the drop-in `ispan` executable with hipcc (gfx950).

Outputs stay in the source tree (ispan_amd/lib, ispan_amd/bin) so they travel
with the repository snapshot.  Usage: python -m ispan_amd.build [--jobs N]
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIB_DIR, "libispan_hip.so")
CLI = os.path.join(BIN_DIR, "ispan")

ARCH = os.environ.get("ISPAN_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["kernels.hip", "giant.hip", "solver.hip", "gen.hip", "batch.hip", "server.hip"]
# -packed-fp32-ops: no packed FP32 instructions in any kernel.
HIPCC_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", f"-I{ROOT}/include", f"-I{CSRC}"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    tmp = LIB + ".tmp"
    subprocess.run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    cli_tmp = CLI + ".tmp"
    subprocess.run([hipcc(), "-O2", "-std=c++17", f"-I{ROOT}/include", os.path.join(CSRC, "ispan_main.cpp"),
                    "-o", cli_tmp, f"-L{LIB_DIR}", "-lispan_hip", "-Wl,-rpath,$ORIGIN/../lib"], check=True)
    os.replace(cli_tmp, CLI)
    if verbose:
        print(f"built {LIB} and {CLI}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args(argv)
    build(a.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
