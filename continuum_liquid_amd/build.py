"""Build the gfx950 HIP library libcliquid_hip.so in-tree.

Usage: python -m continuum_liquid_amd.build   (or __graft_entry__.build()).

Every source is compiled for gfx950 only, with -ffp-contract=off (fused
multiply-adds appear only where the sources call fmaf()) and with packed FP32
instructions disabled (the -packed-fp32-ops target feature).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
CSRC = PACKAGE_DIR / "csrc"
LIBRARY = PACKAGE_DIR / "libcliquid_hip.so"
SOURCES = (
    "context.hip", "binning.hip", "p2g.hip", "grid.hip", "multigrid.hip", "g2p.hip", "voxelise.hip", "fcg.hip", "unet.hip", "sources.hip", "mesh.hip",
)

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    candidate = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(candidate).exists():
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return candidate


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile each source to an object (in parallel) and link the library."""
    out_dir = PACKAGE_DIR / "_build"
    out_dir.mkdir(exist_ok=True)
    compiler = hipcc()
    procs = []
    objects = []
    for name in SOURCES:
        src = CSRC / name
        obj = out_dir / (src.stem + ".o")
        objects.append(obj)
        cmd = [compiler, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for name, proc in procs:
        out, _ = proc.communicate()
        if proc.returncode != 0:
            failed.append(f"--- {name}\n{out}")
        elif verbose and out.strip():
            print(out, flush=True)
    if failed:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(failed))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [compiler, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("HIP link failed:\n" + res.stdout)
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
