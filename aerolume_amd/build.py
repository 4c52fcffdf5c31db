"""Build ``libaerolume_amd.so`` in-tree with hipcc for gfx950.

``python -m aerolume_amd.build`` (or ``__graft_entry__.build()``) compiles every HIP source
of ``aerolume_amd/csrc`` into one shared library next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
SOURCES = ("ns2d.hip", "ns3d.hip", "cnn3d.hip", "geometry.hip", "particles.hip")
LIB = HERE / "libaerolume_amd.so"
ARCH = os.environ.get("AEROLUME_AMD_ARCH", "gfx950")
# -ffp-contract=on: a * b + c becomes one fused multiply-add only inside one source expression, so
# a kernel body gives the same bits whatever code surrounds it (the plan views of the tile
# kernels and the batched ensemble's generic kernels share bodies and must agree bit for bit)
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=on", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build aerolume_amd)")


def build(verbose: bool = False) -> Path:
    """Compile all sources (one object per source, in parallel) and link the library."""
    cc = hipcc()
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        objs.append(obj)
        cmd = [cc, f"--offload-arch={ARCH}", *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for src, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n" + "\n".join(lines))
        if verbose and lines:
            print("\n".join(lines))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[str(o) for o in objs], "-o", str(tmp)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
