"""Build script for libvoxelrt_hip.so (gfx950) and the oracle library.

Run ``python -m voxelrt_amd.build`` (or ``__graft_entry__.build()``).  Every
call recompiles: sources are few and hipcc takes seconds per file.
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
BUILD = PKG / "_build"
LIB = PKG / "libvoxelrt_hip.so"
ORACLE_DIR = ROOT / "oracle"

SOURCES = ["world.hip", "cast.hip", "rays.hip", "frame.hip", "shader.hip", "render.hip", "scene.hip", "cvox.hip", "batch.hip", "diff.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VRT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++20",
    # FP contract: only explicit fmaf fuses; IEEE division and sqrt (hipcc default)
    "-ffp-contract=off",
    "-fno-fast-math",
    # never emit packed-FP32 VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall",
    "-Wno-unused-function",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build_library(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    workers = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread", "-ldl"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(objs)} sources", flush=True)
    return LIB


TORCH_EXT = PKG / "_vrt_torch.so"


def build_torch_extension(verbose: bool = True) -> Path:
    """libtorch autograd wrapper (csrc/torch_render.cpp), a Python extension linked
    against libvoxelrt_hip.so (found next to it through $ORIGIN)."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension

    cxx = shutil.which("g++") or "g++"
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = [cxx, "-O2", "-fPIC", "-shared", "-std=c++17", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DTORCH_EXTENSION_NAME=_vrt_torch", "-DTORCH_API_INCLUDE_EXTENSION_H",
           f"-I{ROOT / 'include'}", f"-I{sysconfig.get_paths()['include']}"]
    cmd += [f"-I{p}" for p in cpp_extension.include_paths()]
    cmd += [str(CSRC / "torch_render.cpp"), "-o", str(TORCH_EXT) + ".tmp"]
    cmd += [f"-L{p}" for p in cpp_extension.library_paths()]
    cmd += ["-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", f"-L{PKG}", "-lvoxelrt_hip", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"torch extension build failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(str(TORCH_EXT) + ".tmp", TORCH_EXT)
    if verbose:
        print(f"built {TORCH_EXT.relative_to(ROOT)}", flush=True)
    return TORCH_EXT


def build_oracle(verbose: bool = True) -> Path:
    """The oracle is test infrastructure (the checker and CPU baseline)."""
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    res = subprocess.run([make, "-B", "-C", str(ORACLE_DIR)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print("built oracle/liboracle.so", flush=True)
    return ORACLE_DIR / "liboracle.so"


def build_all(verbose: bool = True) -> None:
    build_library(verbose)
    build_torch_extension(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all()
    sys.exit(0)
