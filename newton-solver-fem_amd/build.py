"""Build libfem_hip.so for gfx950 in-tree (hipcc, no JIT cache).

    python newton-solver-fem_amd/build.py
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent
REPO = PKG_ROOT.parent
CSRC = PKG_ROOT / "csrc"
LIB = PKG_ROOT / "newton_solver_fem_amd" / "libfem_hip.so"
SOURCES = [CSRC / "fem_hip.hip", CSRC / "fem_contact.hip", CSRC / "fem_krylov.hip", CSRC / "fem_pattern.cpp"]
HEADERS = [CSRC / "fem_math.hpp", CSRC / "fem_pattern.hpp", CSRC / "fem_contact.hpp", CSRC / "fem_krylov.hpp",
           REPO / "include" / "fem_hip.h"]
ARCH = os.environ.get("FEM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def command(out: Path = LIB) -> list[str]:
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # Packed fp32 VALU forms are kept out of the generated code.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{REPO / 'include'}",
        f"-I{CSRC}",
        *[str(s) for s in SOURCES],
        "-o", str(out),
    ]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode})")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
