"""Build the gfx950 shared library offlinerenderer_amd/lib/libwr_amd.so.

Compiles every source under csrc/ with hipcc for gfx950 and links one shared
object exporting the wrk_* kernel C ABI (include/wrk.h), the wr_c_* C ABI
(include/wr_c.h) and the wr:: C++ API (include/wr/*.h).

    python -m offlinerenderer_amd.build        (or build() from Python)
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
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libwr_amd.so"

SOURCES = [
    "bvh.hip",
    "sort.hip",
    "partition.hip",
    "wavefront.hip",
    "denoise.hip",
    "unet.cpp",
    "wrk_api.cpp",
    "engine.cpp",
    "wr_host.cpp",
    "wr_capi.cpp",
    "image_io.cpp",
]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -ffp-contract=off: path arithmetic rounds exactly as written (the CPU
# oracle is built the same way). packed-fp32-ops disabled: packed FP32
# instructions are avoided on this target.
COMMON = [
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"),
    "-Wno-unused-result",
]


def _compile(src: str) -> Path:
    out = OBJ / (src + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / src), "-o", str(out)]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
