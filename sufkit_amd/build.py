"""Build the in-tree native libraries of sufkit_amd.

    python -m sufkit_amd.build          # or sufkit_amd.build.build_all()

Outputs (git-ignored, shipped in-tree next to their sources):
    sufkit_amd/lib/libsufkit_hip.so   HIP/gfx950 construction backend + C ABI
                                      (include/sufkit_hip.h)
    sufkit_amd/lib/libsufkit.so       C++17 host library (reference model,
                                      SuffixArray build/save/load, .sufidx)
                                      linked against libsufkit_hip.so
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("SUFKIT_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libsufkit_hip.so"
    tmp = LIB / "libsufkit_hip.so.tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", str(tmp), *map(str, sorted(CSRC.glob("*.hip")))])
    os.replace(tmp, out)
    return out


# Host-code sanitizers (AddressSanitizer + UndefinedBehaviorSanitizer, every
# report fatal), for a CPU build of libsufkit.so only; the HIP library is
# never built with them.
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=all"]


def build_host(out_dir: Path | None = None, sanitize: bool = False) -> Path | None:
    """libsufkit.so into sufkit_amd/lib (or out_dir, which then also gets a
    link to the HIP library it loads).  sanitize: ASan + UBSan build."""
    sources = sorted(HOST.glob("*.cpp")) if HOST.exists() else []
    if not sources:
        return None
    out_dir = Path(out_dir) if out_dir else LIB
    out_dir.mkdir(parents=True, exist_ok=True)
    if out_dir != LIB:
        hip_lib = out_dir / "libsufkit_hip.so"
        if not hip_lib.exists():
            hip_lib.symlink_to(LIB / "libsufkit_hip.so")
    out = out_dir / "libsufkit.so"
    tmp = out_dir / "libsufkit.so.tmp"
    opt = SANITIZE_FLAGS if sanitize else ["-O2"]
    _run([
        CXX, *opt, "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
        f"-I{INCLUDE}", f"-I{HOST}",
        *map(str, sources),
        f"-L{LIB}", "-lsufkit_hip", "-Wl,-rpath,$ORIGIN", "-lz", "-lpthread",
        "-o", str(tmp),
    ])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_hip()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
