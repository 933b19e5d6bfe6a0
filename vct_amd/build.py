"""Build the gfx950 HIP library (vct_amd/_lib/libvct.so) and the headless
application library (vct_amd/_lib/libvct_app.so).

Every .hip source under vct_amd/csrc is compiled with hipcc for gfx950 only;
the application layer (vct_amd/host) is plain C++ built with g++ and linked
against libvct.so, so it reaches the GPU only through the C ABI of
include/vct.h.  Objects go to vct_amd/_build (ignored by git); the shared
libraries are written in-tree so they travel with the repository snapshot.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent
REPO = ROOT.parent
CSRC = ROOT / "csrc"
HOST = ROOT / "host"
BUILD = ROOT / "_build"
LIB = ROOT / "_lib"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    # packed FP32 (v_pk_*_f32) is disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(REPO / "include"),
    "-I", str(CSRC),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    # surface real warnings, drop the host pass's note about the device-only feature
    noise = ("packed-fp32-ops", "warning: unknown target feature", "1 warning generated")
    for line in (proc.stdout + proc.stderr).splitlines():
        if line.strip() and not any(n in line for n in noise):
            sys.stderr.write(line + "\n")


def _compile_hip(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> dict:
    """Compile everything; returns the paths of the built libraries."""
    BUILD.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, sources))
    libvct = LIB / "libvct.so"
    tmp = LIB / "libvct.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, libvct)
    out = {"libvct": str(libvct)}
    host_sources = sorted(HOST.glob("*.cpp"))
    if host_sources:
        libapp = LIB / "libvct_app.so"
        tmp = LIB / "libvct_app.so.tmp"
        _run([CXX, "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall", "-ffp-contract=off",
              "-I", str(REPO / "include"), "-I", str(HOST),
              "-o", str(tmp), *map(str, host_sources),
              "-L", str(LIB), "-lvct", "-lz", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, libapp)
        out["libvct_app"] = str(libapp)
    if verbose:
        print(out)
    return out


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(verbose=True)
