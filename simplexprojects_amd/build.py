"""Build libblpg.so (HIP kernels + C ABI + C++ drop-in shim) in-tree for gfx950.

Usage: python -m simplexprojects_amd.build   (or simplexprojects_amd.build.build()).
Objects go to simplexprojects_amd/_build/, the library to simplexprojects_amd/libblpg.so.
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
LIB = PKG / "libblpg.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BLPG_ARCH", "gfx950")

# -ffp-contract=off: every FMA in the kernels is explicit, so tableau arithmetic matches
#   the CPU oracle operation for operation.
# packed-fp32-ops disabled: no v_pk_* FP32 instructions are generated.
HIPFLAGS = [
    "-mllvm", "-simplifycfg-sink-common=false",
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{ROOT / 'include'}",
]
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}"]

HIP_SOURCES = ["kernel_wave.hip", "kernel_stream.hip", "kernel_tile.hip", "kernel_verify.hip", "kernel_util.hip",
               "kernel_mlp.hip", "abi.cpp", "simplex_shim.cpp", "warmstart.cpp", "serve.cpp",
               "noalloc.cpp"]
SHIM_TEST = PKG / "test_shim"  # C++ drop-in test program (tests/cpp/test_shim.cpp)
CXX_SOURCES = ["gen.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    if src.suffix == ".hip" or src.name in ("abi.cpp", "simplex_shim.cpp", "warmstart.cpp",
                                            "serve.cpp", "noalloc.cpp"):
        cmd = [HIPCC, *HIPFLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXXFLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = [CSRC / s for s in HIP_SOURCES + CXX_SOURCES if (CSRC / s).exists()]
    jobs = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    shutil.move(str(tmp), str(LIB))
    # the C++ drop-in test program links the library like a reference caller would
    _run(["g++", "-O1", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests/cpp/test_shim.cpp"),
          str(LIB), "-Wl,-rpath,$ORIGIN", "-o", str(SHIM_TEST)])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
