"""Build libairhip.so (the engine's C ABI + HIP kernels) for gfx950, in-tree.

    python -m airtoolsii_amd.build          # or airtoolsii_amd.build.build()

Each HIP source is compiled to its own object (objects are independent: no
relocatable device code) and linked into ``airtoolsii_amd/lib/libairhip.so``.
The ray and circle tracers are compiled with ``-ffp-contract=off`` so that
their FP64 arithmetic rounds operation by operation, as the reference tracers
do; all
HIP code disables packed-FP32 instruction selection.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libairhip.so")

ARCH = os.environ.get("AIR_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = {
    "air_tracer.hip": ["-ffp-contract=off", "-fno-builtin-sin", "-fno-builtin-cos"],
    "air_circles.hip": ["-ffp-contract=off", "-fno-builtin-sin", "-fno-builtin-cos"],
    "air_fresnel.hip": ["-ffp-contract=off"],
    "air_fbp.hip": [],
    "air_art.hip": [],
    "air_sort.hip": [],
    "air_sirt.hip": [],
    "air_tiled.hip": [],
    "air_mlp.hip": [],
    "air_batch.hip": [],
    "air_multi.hip": [],
    "air_spmm.hip": [],
    "air_engine.hip": [],
}


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libairhip.so")
    return h


def _compile(src, extra):
    obj = os.path.join(OBJ_DIR, src.replace(".hip", ".o"))
    cmd = [_hipcc(), *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp, *objs,
           "-lhipfft"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
