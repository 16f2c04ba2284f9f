"""Builds libcsipm.so (gfx950) and the host-only geometry check library in-tree.

    python -m csipm.build        (from compressedsensingipm.jl_amd/)
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                     # compressedsensingipm.jl_amd
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INC = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
          # packed FP32 instructions are kept out of every kernel
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{INC}", f"-I{CSRC}"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError(f"build failed: {cmd[-1]}")


def build(verbose=True):
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_radix.py")], check=True)
    out = os.path.join(PKG, "libcsipm.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-O3", *COMMON, "-o", tmp, os.path.join(CSRC, "csipm.hip")])
    os.replace(tmp, out)
    chk = os.path.join(PKG, "libcsipm_geomcheck.so")
    _run([HIPCC, "-O2", *COMMON, "-o", chk + ".tmp", os.path.join(CSRC, "geom_check.hip")])
    os.replace(chk + ".tmp", chk)
    return out


if __name__ == "__main__":
    build()
