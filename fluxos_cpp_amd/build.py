"""Build the native parts of fluxos_cpp_amd in-tree.

* ``fluxos_cpp_amd/libfluxos_hip.so`` -- HIP kernels for gfx950 + the C ABI
  declared in ``include/fluxos_hip.h`` (hipcc), plus the host input layer (g++).
* ``fluxos_cpp_amd/fluxos_hip`` -- the command-line driver (a restatement of the
  reference's ``main.cpp`` time loop over the C ABI).

Every HIP source is compiled with ``-ffp-contract=off`` (the FP64 parity
kernels must not fuse multiply-adds: the reference's x86-64 build has none)
and with packed FP32 instructions disabled.

Usage: ``python -m fluxos_cpp_amd.build`` (or ``build_all()``).
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
HOST = PKG / "host"
OBJ = PKG / "_build"
LIB = PKG / "libfluxos_hip.so"
CLI = PKG / "fluxos_hip"

ARCH = os.environ.get("FLUXOS_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]
CXX_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-variable"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


# sources of the FP32-mixed path: FP32 division and square root may be the
# hardware's approximate forms (<= 1 ulp) and FP32 denormals (< 1.2e-38) are
# flushed; FP64 arithmetic is unaffected
FAST_FP32 = {"fx_fp32.hip", "fx_ade32.hip"}


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        extra = (["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero"]
                 if src.name in FAST_FP32 else [])
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-x", "hip", "-c", str(src), "-o", str(out)]
    else:
        cmd = [CXX, *CXX_FLAGS, "-I/opt/rocm/include", "-c", str(src), "-o", str(out)]
    _run(cmd)
    return out


def build_lib(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> Path | None:
    srcs = sorted(HOST.glob("*.cpp"))
    if not srcs:
        return None
    tmp = CLI.with_name(CLI.name + ".tmp")
    _run([CXX, *CXX_FLAGS, "-I", str(ROOT / "include"), *map(str, srcs), "-o", str(tmp),
          f"-L{PKG}", "-lfluxos_hip", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liborc_fluxos.so"


def build_all(jobs: int = 8) -> None:
    build_lib(jobs)
    build_cli()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB, CLI if CLI.exists() else "(no CLI)", file=sys.stderr)
