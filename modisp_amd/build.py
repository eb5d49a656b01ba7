"""Compile the gfx950 backend library in-tree: modisp_amd/libmodisp_rocm.so."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
LIBRARY = HERE / "libmodisp_rocm.so"
SOURCES = [HERE / "csrc" / "kernels.hip", HERE / "csrc" / "adjoint.hip", HERE / "csrc" / "solver.cpp",
           HERE / "csrc" / "bounds.cpp", HERE / "csrc" / "results.cpp"]
HEADERS = [HERE / "csrc" / "secular.hpp", HERE / "csrc" / "pipeline.hpp", HERE / "csrc" / "adjoint_root.hpp",
           ROOT / "include" / "modisp_rocm.h"]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = Path(rocm) / "bin" / "hipcc"
    return str(path) if path.exists() else "hipcc"


def command(output: Path = LIBRARY, arch: str = "gfx950") -> list[str]:
    return [hipcc(), f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
            # Packed FP32 instructions are disabled for every HIP translation unit.
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", str(ROOT / "include"), "-o", str(output),
            str(SOURCES[0]), str(SOURCES[1]), "-x", "hip", *[str(s) for s in SOURCES[2:]]]


def build(force: bool = True, quiet: bool = True) -> Path:
    """Compile the library (always when force, else only if sources are newer)."""
    if not force and LIBRARY.exists():
        newest = max(p.stat().st_mtime for p in SOURCES + HEADERS)
        if LIBRARY.stat().st_mtime >= newest:
            return LIBRARY
    tmp = LIBRARY.with_suffix(".so.tmp")
    proc = subprocess.run(command(tmp), capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + proc.stderr[-4000:])
    os.replace(tmp, LIBRARY)
    if not quiet:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(force=True, quiet=False)
