"""Build the gfx950 HIP library in-tree (``weno_sr_publishing_amd/libwsr_hip.so``).

``python -m weno_sr_publishing_amd.build`` compiles it with hipcc; it runs on
a machine without a GPU (cross-compilation for gfx950).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwsr_hip.so"
SOURCES = [CSRC / "wsr_euler2d.hip", CSRC / "wsr_euler7.hip", CSRC / "wsr_euler3.hip", CSRC / "wsr_adv1d.hip", CSRC / "wsr_train1d.hip"]
HEADERS = [CSRC / "wsr_math.hpp", CSRC / "wsr_loop.hpp", CSRC / "wsr_weno7.hpp", CSRC / "wsr_adv1d.hpp", CSRC / "wsr_adjoint.hpp", PKG.parent / "include" / "wsr.h"]
ARCH = os.environ.get("WSR_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, *(extra or []), "-o", str(tmp), *map(str, SOURCES)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def stale() -> bool:
    if not LIB.exists():
        return True
    t = LIB.stat().st_mtime
    return any(p.stat().st_mtime > t for p in SOURCES + HEADERS)


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
    print(LIB)
