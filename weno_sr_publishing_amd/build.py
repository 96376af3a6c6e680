"""Build the gfx950 HIP library in-tree (``weno_sr_publishing_amd/libwsr_hip.so``).

``python -m weno_sr_publishing_amd.build`` compiles it with hipcc; it runs on
a machine without a GPU (cross-compilation for gfx950). Each translation unit
is compiled to its own object in parallel (``_build/``), then the objects are
linked into the shared library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwsr_hip.so"
OBJ_DIR = PKG / "_build"
SOURCES = [CSRC / "wsr_euler2d.hip", CSRC / "wsr_euler7.hip", CSRC / "wsr_euler3.hip", CSRC / "wsr_adv1d.hip",
           CSRC / "wsr_train1d.hip", CSRC / "wsr_init.hip", CSRC / "wsr_beta7.hip"]
# the initial-state kernels reproduce the host builders' roundings: no contraction
SOURCE_FLAGS = {"wsr_init.hip": ["-ffp-contract=off"],
                # the FP32 WENO7 closure's matrix-core results stay in VGPRs (no AGPR reads)
                "wsr_euler7.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
HEADERS = [CSRC / "wsr_math.hpp", CSRC / "wsr_loop.hpp", CSRC / "wsr_weno7.hpp", CSRC / "wsr_adv1d.hpp",
           CSRC / "wsr_adjoint.hpp", PKG.parent / "include" / "wsr.h"]
ARCH = os.environ.get("WSR_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
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


def _jobs() -> int:
    try:
        n = int(os.environ.get("MAX_JOBS", "0"))
    except ValueError:
        n = 0
    return max(1, min(len(SOURCES), n or (os.cpu_count() or 1)))


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    cc = hipcc()
    extra = list(extra or [])

    def compile_one(src: Path) -> Path:
        obj = OBJ_DIR / (src.stem + ".o")
        cmd = [cc, *HIPCC_FLAGS, *SOURCE_FLAGS.get(src.name, []), *extra, "-c", "-o", str(obj), str(src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    with ThreadPoolExecutor(max_workers=_jobs()) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
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
