"""Builds liblss.so (the HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Usage: ``python -m livespikesorter_amd.build`` or ``livespikesorter_amd.build.build()``.
Objects go to ``_build/`` at the repository root; the library to ``livespikesorter_amd/liblss.so``.
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
LIB = PKG / "liblss.so"
ARCH = os.environ.get("LSS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.cpp", "npy.cpp", "output.cpp", "preprocess.hip", "gemm.hip", "project.hip",
           "pursuit.hip", "tables.hip", "adjoint.hip", "train.hip", "train_api.cpp", "assign.hip"]

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build liblss")
    return p


def _compile(src: str, obj: Path) -> tuple[str, int, str]:
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
           "-Wno-unused-result", "-I", str(ROOT / "include")] + DEVICE_FLAGS + [
           "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, (r.stdout + r.stderr)


def build(verbose: bool = False, jobs: int = 8) -> Path:
    out = ROOT / "_build"
    out.mkdir(exist_ok=True)
    objs = [out / (Path(s).stem + ".o") for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, SOURCES, objs))
    for src, rc, log in results:
        msg = "\n".join(l for l in log.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if rc != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{log}")
        if verbose and msg:
            print(f"[{src}]\n{msg}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    build_facade_check()
    return LIB


def build_facade_check() -> Path | None:
    """The C++ program that drives include/lss_online_spikes.hpp in the tests (tests/ next to its
    source; it finds liblss.so through its rpath)."""
    src = ROOT / "tests" / "online_facade_check.cpp"
    if not src.exists():
        return None
    exe = src.with_suffix("")
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("no host C++ compiler for the facade check program")
    cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
           "-L", str(PKG), "-llss", "-Wl,-rpath,$ORIGIN/../livespikesorter_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade check build failed:\n{r.stdout}{r.stderr}")
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
