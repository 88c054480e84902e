"""Builds libqhmat_hip.so in-tree with hipcc for gfx950.

Usage: python qhmat_amd/build_lib.py  (or __graft_entry__.build()).  Run it as
a script, not with -m: importing the package loads the library being rebuilt.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
LIBRARY = PACKAGE / "libqhmat_hip.so"
SOURCES = ["device_store.hip", "wide_passes.hip", "server.hip", "host_store.cpp", "capi.cpp"]
ARCH = os.environ.get("QHMAT_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def command(output: Path) -> list[str]:
    return [
        hipcc(),
        "-O3",
        "-std=c++17",
        f"--offload-arch={ARCH}",
        # code objects the HIP runtime bundled with PyTorch-ROCm loads too
        "-mcode-object-version=5",
        "-fPIC",
        "-shared",
        "-ffp-contract=off",
        # Packed FP32 (v_pk_*_f32) stays off: only scalar FP32 instructions.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I", str(PACKAGE.parent / "include"),
        *[str(CSRC / s) for s in SOURCES],
        "-o", str(output),
        "-lpthread",
    ]


def build(verbose: bool = False) -> Path:
    """Compiles every source into a fresh library and swaps it in atomically."""
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError("hipcc failed to build libqhmat_hip.so")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build(verbose=True))
