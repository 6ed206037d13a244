"""Builds the MI355X energy-path library (mcgpu_amd/libmcgpu.so) in-tree.

HIP sources (csrc/*.hip) are compiled by hipcc for gfx950 only; host sources
(host/*.cpp) by g++ without FMA contraction (the plain path's move
construction must produce the reference's doubles).  Packed-FP32 VALU ops are
disabled for device code.  Run as ``python -m mcgpu_amd.build``.
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
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libmcgpu.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
             f"-I{INCLUDE}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    # hipcc's host pass warns about the device-only target feature: harmless
    noise = "'-packed-fp32-ops' is not a recognized feature"
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l)
    if out.strip():
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in sorted((PKG / "csrc").glob("*.hip")):
        obj = BUILD / (src.stem + ".hip.o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted((PKG / "host").glob("*.cpp")):
        if src.name == "metrosim.cpp":
            continue
        obj = BUILD / (src.stem + ".o")
        jobs.append([CXX, *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *objs])
    os.replace(tmp, LIB)
    metrosim = PKG / "host" / "metrosim.cpp"
    if metrosim.exists():
        BIN.mkdir(exist_ok=True)
        _run([CXX, *CXX_FLAGS, str(metrosim), "-o", str(BIN / "metrosim"),
              f"-L{PKG}", "-lmcgpu", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle() -> None:
    """Test infrastructure: the C oracle, and the reference's CPU path when
    the reference tree is present in this container."""
    _run(["make", "-s", "-C", str(ROOT / "oracle"), "oracle"])
    if Path("/root/reference/src").is_dir() and shutil.which("g++"):
        _run(["make", "-s", "-j8", "-C", str(ROOT / "oracle"), "ref"])


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
