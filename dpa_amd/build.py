"""Build recipe of libdpa_hip.so (HIP, gfx950) and of the C++ host tools.

Run as ``python -m dpa_amd.build`` or through ``__graft_entry__.build()``.
Objects are compiled in parallel into ``dpa_amd/_build`` and linked in-tree,
so the shared library travels with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libdpa_hip.so"

HIP_SOURCES = ["advect.hip", "compact.hip", "ftle.hip", "batch.hip", "serve.hip", "tune.hip",
               "graph.hip", "noalloc.hip", "capi.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DPAH_OFFLOAD_ARCH", "gfx950")

# Bit-level parity with dpa's float32 CPU arithmetic needs: no contraction into
# FMA, correctly rounded division and sqrt, IEEE denormals.  Packed FP32 ops are
# disabled for the device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build_library(jobs: int = 4) -> Path:
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs]])
    os.replace(tmp, LIB)
    return LIB


HOST_SOURCES = ["json.cpp", "stages.cpp", "io.cpp", "pipeline.cpp", "main.cpp"]
HOST = HERE / "host"
CLI = HERE / "bin" / "dpa_hip"
HOST_CXX = os.environ.get("DPAH_HOST_CXX", "/opt/rocm/llvm/bin/clang++")
HDF5_PREFIX = Path(os.environ.get("DPAH_HDF5_PREFIX", "/opt/conda"))


def build_cli() -> Path:
    """The dpa_hip command line (host C++ over libdpa_hip.so), with HDF5 I/O
    when a serial HDF5 is found."""
    CLI.parent.mkdir(exist_ok=True)
    # the same host compiler as hipcc, so both sides agree on the C++ runtime
    cmd = [HOST_CXX, "-std=c++17", "-O2", "-Wall", "-Wextra", "-fPIC",
           *[str(HOST / s) for s in HOST_SOURCES], "-o", str(CLI.with_suffix(".tmp")),
           f"-L{HERE}", "-ldpa_hip", "-Wl,-rpath,$ORIGIN/.."]
    if (HDF5_PREFIX / "include" / "hdf5.h").exists():
        cmd[4:4] = [f"-I{HDF5_PREFIX / 'include'}", "-DDPAH_WITH_HDF5"]
        # system libraries first: the HDF5 prefix may carry an older libstdc++
        cmd += [f"-L{HDF5_PREFIX / 'lib'}", "-lhdf5",
                f"-Wl,-rpath,/usr/lib/x86_64-linux-gnu:{HDF5_PREFIX / 'lib'}"]
    _run(cmd)
    os.replace(CLI.with_suffix(".tmp"), CLI)
    return CLI


def build_oracle() -> Path:
    """The CPU oracle is test infrastructure; it is built here so that tests,
    the smoke check and the bench's CPU baseline find it."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "libdpa_oracle.so"


def build_all() -> None:
    build_library()
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(LIB)
