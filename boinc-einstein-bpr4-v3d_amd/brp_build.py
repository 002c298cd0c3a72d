"""Build libbrp_hip.so (gfx950) in place.

Host science (csrc/science.cpp) is compiled by g++; the HIP sources by hipcc for
gfx950 only.  Every translation unit uses -ffp-contract=off: the resampling index
rule depends on each FP32 operation rounding as written.  Packed FP32 instructions
are disabled (-packed-fp32-ops) for all device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libbrp_hip.so"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter", f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
             f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_SOURCES = ["plain.hip", "fused.hip", "preprocess.hip", "toplist.hip", "engine.hip", "noalloc.hip"]
CLI_SOURCES = ["brp_bench.cpp", "brp_workunit.cpp"]
BIN = PKG / "bin"
CXX_SOURCES = ["science.cpp", "counters.cpp"]
# counters.cpp uses the rocprofiler-sdk headers (which include the HIP runtime headers) for its
# types only: the library loads rocprofiler-sdk itself at run time, when asked to.
CXX_EXTRA = {"counters.cpp": ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")


def build(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(BUILD / (src + ".o"))])
    for src in CXX_SOURCES:
        jobs.append(["g++", *CXX_FLAGS, *CXX_EXTRA.get(src, []), "-c", str(CSRC / src), "-o", str(BUILD / (src + ".o"))])
    workers = min(len(jobs), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        for cmd, fut in [(c, pool.submit(_run, c)) for c in jobs]:
            fut.result()
            if verbose:
                print(" ".join(cmd[:1] + cmd[-3:]))
    objects = [str(BUILD / (s + ".o")) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objects, "-lpthread", "-ldl"])
    os.replace(tmp, LIB)
    build_cli()
    return LIB


def build_cli() -> list[pathlib.Path]:
    """The command-line drivers (cli/*.cpp, the reference's brp_bench and brp_workunit)
    into bin/, linked against the library with an rpath to it."""
    BIN.mkdir(exist_ok=True)
    out = []
    for src in CLI_SOURCES:
        exe = BIN / pathlib.Path(src).stem
        tmp = exe.with_suffix(".tmp")
        _run(["g++", "-std=c++20", "-O2", "-ffp-contract=off", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", f"-I{PKG / 'cli'}",
              str(PKG / "cli" / src), "-o", str(tmp), f"-L{PKG}", "-lbrp_hip", "-lz", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
