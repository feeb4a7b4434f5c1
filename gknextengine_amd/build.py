"""Builds the HIP library libgknx_rt.so for gfx950, in-tree.

Every source under csrc/ is compiled with hipcc (-O3, no FP contraction, packed-FP32
instructions disabled) and linked into gknextengine_amd/_lib/libgknx_rt.so. Run as
``python -m gknextengine_amd.build`` or through ``__graft_entry__.build()``.
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
OUT_DIR = PKG / "_lib"
BUILD_DIR = PKG / "_build"
LIB = OUT_DIR / "libgknx_rt.so"

ARCH = os.environ.get("GKNX_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-command-line-argument",
    f"-I{PKG.parent / 'include'}",
]


# per-source additions: the U-Net's fixed-shape convolutions unroll K loops of up to 64
# steps (each step's operand addresses are compile-time constants only when unrolled)
EXTRA = {"unet.hip": ["-mllvm", "-pragma-unroll-threshold=262144"]}


def sources() -> list[Path]:
    return sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])


def _compile(src: Path) -> Path:
    obj = BUILD_DIR / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, *EXTRA.get(src.name, []), "-x", "hip", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    OUT_DIR.mkdir(exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lz", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    build_host_demo()
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB) from {len(srcs)} sources, and {DEMO.name}")
    return LIB


DEMO = OUT_DIR / "gk_cornell_demo"


def build_host_demo() -> Path:
    """The C++ adapter demo (host/), a plain g++ build against the C ABI."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{PKG.parent / 'include'}", str(PKG / "host" / "cornell_demo.cpp"),
           f"-L{OUT_DIR}", "-lgknx_rt", "-Wl,-rpath,$ORIGIN", "-o", str(DEMO)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for the host demo:\n{res.stdout}\n{res.stderr}")
    return DEMO


if __name__ == "__main__":
    build()
    sys.exit(0)
