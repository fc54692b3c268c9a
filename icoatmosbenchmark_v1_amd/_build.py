"""Builds libnicam_tt.so (the gfx950 HIP kernels behind include/nicam_tt.h).

The library is compiled in-tree with hipcc for gfx950 only.  Every object is
built with -ffp-contract=off: the FP64 path must reproduce the reference's
separately rounded multiplies and adds, and the packed-FP32 instructions are
disabled for the reasons given in the package documentation.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
INCLUDE = REPO_DIR / "include"
LIB_NAME = "libnicam_tt.so"
LIB_PATH = PKG_DIR / LIB_NAME
OBJ_DIR = PKG_DIR / "_build"

SOURCES = ["ctx.cpp", "comm.cpp", "grid.cpp", "gmtr.cpp", "oprt_setup.cpp", "hflux.hip", "hlim.hip", "vlim.hip", "vi.hip", "comm.hip", "oprt.hip", "glue.hip", "vglue.hip", "f32.hip", "adjoint.hip", "mass.hip", "conserve.hip", "group.cpp", "glue_adjoint.hip"]

HIPCC_FLAGS = [
    "-x", "hip",
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libnicam_tt.so)")


def build_library(verbose: bool = False, jobs: int = 4) -> Path:
    """Compiles every source and links LIB_PATH; always rebuilds."""
    hipcc = hipcc_path()
    OBJ_DIR.mkdir(exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = OBJ_DIR / (src.replace(".", "_") + ".o")  # comm.cpp and comm.hip both exist
        cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed:\n{out.stdout}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def _wait(item) -> None:
    src, proc = item
    out, _ = proc.communicate()
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{out}")
    if out and out.strip():
        # keep warnings visible, minus hipcc's host-pass note about the GPU-only feature flag
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            print(f"[{src}]\n" + "\n".join(lines), flush=True)


if __name__ == "__main__":
    print(build_library(verbose=True))
