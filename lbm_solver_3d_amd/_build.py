"""Build liblbm_hip.so (HIP, gfx950) in-tree with hipcc.

Four translation units:
  kernels_fp64.hip  FP64 step kernels, squirmer/remap and auxiliary kernels,
                    -ffp-contract=off (reference rounding)
  kernels_fast.hip  FP32 / FP16 / BF16 step kernels, -ffp-contract=off too
                    (their multiply-adds are written as fmaf)
  lbm_hip.hip       the C ABI (include/lbm_hip.h)
  mlp.hip           the learned velocity surrogate (include/lbm_mlp.h)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "liblbm_hip.so"
ARCH = os.environ.get("LBM_HIP_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
UNITS = {
    "kernels_fp64": ["-ffp-contract=off"],
    "kernels_fast": ["-ffp-contract=off"],  # multiply-adds are explicit fmaf (same rounding in every kernel)
    "lbm_hip": ["-ffp-contract=off"],
    "mlp": ["-ffp-contract=fast"],
}


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (Path(c).exists() or c == "hipcc"):
            return c
    raise RuntimeError("hipcc not found")


def _compile(name: str, extra: list[str]) -> Path:
    obj = BUILD / f"{name}.o"
    cmd = [hipcc(), *COMMON, *extra, "-c", str(CSRC / f"{name}.hip"), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile {name} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


HOST = PKG / "host"
FACADE = PKG / "liblbm_facade.so"
DRIVER = PKG / "bin" / "lbm_squirmer"
HOST_SRC = ["params.cpp", "squirmer.cpp", "lbm.cpp", "io.cpp", "analytic.cpp", "lbm_io.cpp"]


def build_host(verbose: bool = True) -> None:
    """The drop-in C++ facade (class LBM & co. over the C ABI) and the
    lbm_squirmer driver, compiled with the host compiler."""
    cxx = os.environ.get("CXX", "g++")
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", f"-I{HOST}"]
    link = [f"-L{PKG}", "-llbm_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."]
    cmd = [cxx, *flags, "-shared", *[str(HOST / f) for f in HOST_SRC], "-o", str(FACADE), *link]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"facade build failed:\n{' '.join(cmd)}\n{r.stderr}")
    DRIVER.parent.mkdir(exist_ok=True)
    cmd = [cxx, *flags, str(HOST / "main.cpp"), "-o", str(DRIVER), f"-L{PKG}", "-llbm_facade", "-llbm_hip",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print(f"built {FACADE} and {DRIVER}")


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=3) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), UNITS.items()))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
