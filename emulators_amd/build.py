"""Build the gfx950 shared library ``emulators_amd/_lib/libemu_sv_hip.so``.

Usage: ``python -m emulators_amd.build`` (or ``emulators_amd.build.build()``).
The library is built in-tree so it travels with the source tree; nothing is
installed.  hipcc cross-compiles for gfx950 without a GPU present.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
LIB_DIR = PKG_DIR / "_lib"
LIB_NAME = "libemu_sv_hip.so"
LIB_PATH = LIB_DIR / LIB_NAME

SOURCES = ["emu_sv_kernels.hip", "emu_sv_host.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> Path:
    """Compile every source into the shared library; returns its path."""
    LIB_DIR.mkdir(exist_ok=True)
    missing = [s for s in SOURCES if not (CSRC / s).exists()]
    if missing:
        raise RuntimeError(f"missing sources in {CSRC}: {missing}")
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB_DIR / (LIB_NAME + ".tmp")
    cmd = [hipcc_path(), *HIPCC_FLAGS, "-I", str(REPO_DIR / "include"), "-o", str(tmp), *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
