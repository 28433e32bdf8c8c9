"""Build recipe for libsphx.so (HIP, gfx950) — the product library.

Invoked by ``__graft_entry__.build()``; runnable on its own as
``python phys-rs_amd/build_lib.py``.  Cross-compiles on a CPU-only host.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libsphx.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: the neighbour predicate must round exactly like the
# reference's uncontracted sqrt((dx*dx + dy*dy) + dz*dz); fused multiply-adds
# elsewhere are written out with fma().
# -packed-fp32-ops: packed FP32 instructions are disabled on this target.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["sphx_sort.hip", "sphx_pairs.hip", "sphx_particles.hip", "sphx_adjoint.hip", "sphx_runtime.hip"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines()
        if "packed-fp32-ops" not in line and "is not a recognized feature" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{proc.stdout}")
    if out.strip():
        print(out)


def build(verbose: bool = False) -> Path:
    """Compile every HIP source for gfx950 and link libsphx.so (always rebuilds)."""
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
