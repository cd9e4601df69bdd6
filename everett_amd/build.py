"""Builds everett_amd/_lib/libeverett.so from everett_amd/csrc (gfx950 only).

HIP sources (*.hip) go through hipcc for gfx950; host-only sources (*.cpp)
through g++ against the HIP runtime headers.  Objects land in
everett_amd/_build/.  Run as `python everett_amd/build.py` (a script:
importing the package would load the library it builds).

`python everett_amd/build.py --sanitize` builds _lib/libeverett_san.so
instead: the host code (every .cpp, and the host side of every .hip)
instrumented with AddressSanitizer and UndefinedBehaviorSanitizer, the GPU
code as usual.  Every source goes through clang (hipcc / amdclang++) so one
sanitizer runtime serves the whole library; a process loads it with clang's
ASan runtime preloaded (tests/test_sanitizers.py shows how).  It is for host
checks on the CPU; nothing runs it on a GPU.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libeverett.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",  # every fused multiply-add in the kernels is an explicit fma
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]
CXX_FLAGS = [
    "-O2",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    "-D__HIP_PLATFORM_AMD__",
    f"-I{ROCM / 'include'}",
    "-Wall",
    "-Wno-unused-function",
]


# Host-only sanitizer flags: each -fsanitize applies to the host compilation
# only (-Xarch_host), never to GPU code.
SAN_HOST = [
    "-Xarch_host", "-fsanitize=address",
    "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-sanitize-recover=undefined",
    "-Xarch_host", "-fno-omit-frame-pointer",
    "-Xarch_host", "-g",
]
SAN_BUILD = PKG / "_build_san"
SAN_LIB = LIBDIR / "libeverett_san.so"
CLANGXX = str(ROCM / "llvm" / "bin" / "clang++")


def _compile(src: Path, sanitize: bool = False) -> Path:
    obj = (SAN_BUILD if sanitize else BUILD) / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, *(SAN_HOST if sanitize else []), "-c", str(src), "-o", str(obj)]
    elif sanitize:
        cmd = [CLANGXX, *CXX_FLAGS, *SAN_HOST, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False, sanitize: bool = False) -> Path:
    """Compiles every source and links libeverett.so (always rebuilds), or
    with sanitize the host-sanitized libeverett_san.so."""
    out_dir, lib = (SAN_BUILD, SAN_LIB) if sanitize else (BUILD, LIB)
    out_dir.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    workers = min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda src: _compile(src, sanitize), sources))
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    shutil.move(str(tmp), str(lib))
    if verbose:
        print(f"built {lib} from {len(sources)} sources")
    return lib


if __name__ == "__main__":
    build(verbose=True, sanitize="--sanitize" in sys.argv[1:])
    sys.exit(0)
