"""Build the native library `librefrakt_amd.so` for gfx950.

Host-only C++ is compiled with g++, HIP translation units with hipcc, and the
two device runtime headers are embedded as string literals so the flame
compiler can hand them to hiprtc at run time.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "librefrakt_amd.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["json.cpp", "yaml_lite.cpp", "flang.cpp", "vardb.cpp", "flame.cpp", "flam3.cpp", "compiler.cpp"]
HIP_SOURCES = ["api.cpp", "kernels_static.hip", "queue_bin.hip", "unet.hip", "denoiser.cpp", "jpeg.hip"]
# hipcc flags: gfx950 only, packed FP32 instructions disabled
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++20", "-ffp-contract=fast-honor-pragmas",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++20", "-Wall", "-Wno-unused-function",
              f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]


def _embed_device_sources() -> None:
    parts = []
    for var, name in (("kFlameDeviceSrc", "flame_device.h"), ("kFlameKernelsSrc", "flame_kernels.h")):
        # the headers are concatenated into one hiprtc source: drop include guards
        text = "\n".join(l for l in (CSRC / "device" / name).read_text().splitlines() if l.strip() != "#pragma once")
        # split into chunks below the compilers' string literal limits
        chunks = [text[i:i + 12000] for i in range(0, len(text), 12000)]
        lits = "\n".join(f'R"RFKSRC({c})RFKSRC"' for c in chunks)
        parts.append(f"static const char {var}[] =\n{lits};\n")
    out = CSRC / "device_sources.inc"
    content = "// generated from csrc/device/*.h by refraktor_amd/build.py\n" + "\n".join(parts)
    if not out.exists() or out.read_text() != content:
        out.write_text(content)


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: str) -> Path:
    obj = OBJ / (src.replace(".", "_") + ".o")
    if src in HIP_SOURCES:
        cmd = ["hipcc", "-x", "hip", *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    else:
        cmd = ["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    _run(cmd)
    return obj


def build_native(jobs: int | None = None) -> Path:
    OBJ.mkdir(exist_ok=True)
    _embed_device_sources()
    srcs = HOST_SOURCES + HIP_SOURCES
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
          f"-L{ROCM}/lib", "-lhiprtc", "-lz", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_native())
