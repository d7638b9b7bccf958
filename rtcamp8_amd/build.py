"""Builds the gfx950 library (rtcamp8_amd/librtc8x.so) with hipcc, in-tree.

Every kernel source is compiled with IEEE-exact single precision: no fast-math, no FP contraction,
correctly rounded fp32 division and square root, denormals kept.  The path tracer relies on this to
reproduce a scalar host run bit for bit.  Packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librtc8x.so"
RENDER = PKG / "rtc8x_render"
DRIVER_SOURCES = ["driver/scene_file.cpp", "driver/rtc8x_render.cpp"]

SOURCES = [
    "rtc8x_api.hip",
    "rtc8x_grid.hip",
    "rtc8x_env.hip",
    "rtc8x_sky.hip",
    "rtc8x_trace.hip",
    "rtc8x_nrc_data.hip",
    "rtc8x_nrc_mlp.hip",
    "rtc8x_measure.hip",
    "rtc8x_nvdb.hip",
    "rtc8x_post.hip",
    "rtc8x_noalloc.hip",
]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return exe


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build_library(jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-lz", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    return LIB


def build_driver() -> Path:
    """The batch renderer executable (host C++ over the C ABI, linked against librtc8x.so in place)."""
    cxx = shutil.which("g++") or "g++"
    tmp = RENDER.with_suffix(".tmp")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", *[str(CSRC / s) for s in DRIVER_SOURCES],
           f"-L{PKG}", "-lrtc8x", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, RENDER)
    return RENDER


def main() -> int:
    print(build_library())
    return 0


if __name__ == "__main__":
    sys.exit(main())
