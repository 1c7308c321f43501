"""Builds libpbd.so (HIP for gfx950 + C++ host code) in-tree.

Usage: python -m positionbaseddynamicsold_amd.build   (or build() from Python)

Every source under csrc/ is compiled by hipcc into _build/ objects in
parallel and linked into positionbaseddynamicsold_amd/libpbd.so; then the
C++ example caller tests/cpp_driver.cpp is linked against it, and the
exhaustive rounding check tests/rounding_check.hip is built beside it.

-ffp-contract=off keeps every float expression a sequence of IEEE operations
in source order (no FMA contraction), which is what lets the device results
equal a plain CPU evaluation bit-for-bit.  Packed FP32 instructions are
disabled on purpose (-packed-fp32-ops).
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
BUILD = ROOT / "_build"
LIB = PKG / "libpbd.so"

ARCH = os.environ.get("PBD_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found")
    return exe


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [hipcc(), "-c", "-x", "hip", str(src), "-o", str(obj)] + FLAGS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn:
        sys.stderr.write("\n".join(warn) + "\n")
    return obj


DRIVER_SRC = ROOT / "tests" / "cpp_driver.cpp"
DRIVER = ROOT / "tests" / "cpp_driver"


def build_driver(verbose: bool = True) -> Path:
    """The C++ caller of the C ABI (tests/cpp_driver.cpp), linked against
    libpbd.so with an rpath relative to its own location."""
    tmp = DRIVER.with_name(DRIVER.name + ".tmp")
    cmd = [hipcc(), "-x", "c++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(DRIVER_SRC), "-o",
           str(tmp), f"-L{PKG}", "-lpbd", "-Wl,-rpath,$ORIGIN/../positionbaseddynamicsold_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stderr}")
    os.replace(tmp, DRIVER)
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


ROUNDING_SRC = ROOT / "tests" / "rounding_check.hip"
ROUNDING = ROOT / "tests" / "rounding_check"


def build_rounding_check(verbose: bool = True) -> Path:
    """The exhaustive check of the short correctly rounded square root and
    reciprocal in csrc/pbd_math.hpp (tests/rounding_check.hip), built with
    the library's flags."""
    tmp = ROUNDING.with_name(ROUNDING.name + ".tmp")
    flags = [f for f in FLAGS if f not in ("-fPIC", "-fvisibility=hidden")]
    cmd = [hipcc(), "-x", "hip", str(ROUNDING_SRC), "-o", str(tmp)] + flags
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"rounding check build failed:\n{r.stderr}")
    os.replace(tmp, ROUNDING)
    if verbose:
        print(f"built {ROUNDING}")
    return ROUNDING


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    workers = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_driver(verbose)
    build_rounding_check(verbose)
    return LIB


if __name__ == "__main__":
    build()
