"""Build the gfx950 HIP library ``cutcell/libcutcell_hip.so`` in-tree.

Usage: ``python gpu-cutcell-navier-stokes_amd/build.py`` (or ``build()`` from
Python). hipcc cross-compiles for gfx950 without a GPU present. Packed FP32
instructions are disabled (``-packed-fp32-ops``) for every HIP source.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
OUT = PKG / "cutcell" / "libcutcell_hip.so"

SOURCES = ["pcg.hip", "rowblock.hip", "batched.hip", "sumfact.hip", "cnn.hip", "fcg.hip", "mg.hip", "ekte.hip", "gmres.hip", "asm2d.hip", "netpass.hip", "runs.hip", "cut3d.hip"]
ARCH = os.environ.get("CUTCELL_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the HIP library)")


def _flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
    ]


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source to an object in parallel (one hipcc per file),
    then link the shared library. Objects go to a scratch folder beside the
    sources that is removed afterwards."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()

    def compile_one(src: Path):
        obj = objdir / (src.stem + ".o")
        cmd = [cc, *_flags(), "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        return obj, subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or min(len(srcs), os.cpu_count() or 1, 16)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = [r for _, r in results if r.returncode != 0]
    if failed:
        for r in failed:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {len(failed)} source(s)")
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for o, _ in results], "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed with status {res.returncode}")
    os.replace(tmp, OUT)
    shutil.rmtree(objdir, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
