"""Build recipe of the HIP library libsph_hip.so (gfx950) and the CPU oracle.

    python cudasphnew_amd/build.py            # library + oracle
    python cudasphnew_amd/build.py --lib-only

(run as a script: importing the package would load the library being built)

Objects go to cudasphnew_amd/_build/, the library to
cudasphnew_amd/libsph_hip.so (in-tree, so that it travels with the source
tree to the GPU machine).  hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsph_hip.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPH_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{os.path.join(ROOT, 'include')}",
]
# Host-only sources reproduce the reference's FP32 host arithmetic; no
# contraction into FMA.
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              f"-I{os.path.join(ROOT, 'include')}"]

SOURCES = [
    ("scene.cpp", "host"),
    ("radix_sort.hip", "device"),
    ("cell_sort.hip", "device"),
    ("kernels.hip", "device"),
    ("twoscale.hip", "device"),
    ("sim.hip", "device"),
    ("neural.hip", "device"),
    ("mesh.hip", "device"),
    ("noalloc.hip", "device"),
]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return p


def _compile(src: str, kind: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if kind == "host":
        cmd = ["g++", *HOST_FLAGS, "-c", path, "-o", obj]
    else:
        cmd = [HIPCC, *DEVICE_FLAGS, "-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_library(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, k) for s, k in SOURCES]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> str:
    _run(["make", "-s", "-C", ORACLE_DIR])
    return os.path.join(ORACLE_DIR, "liboracle_sph.so")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    if a.clean and os.path.isdir(BUILD):
        shutil.rmtree(BUILD)
    print(build_library(a.j))
    if not a.lib_only:
        print(build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
