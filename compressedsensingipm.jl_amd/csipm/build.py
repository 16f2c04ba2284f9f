"""Builds libcsipm.so (gfx950) and the host-only geometry check library in-tree.

    python -m csipm.build        (from compressedsensingipm.jl_amd/)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                     # compressedsensingipm.jl_amd
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
INC = os.path.join(REPO, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_C = ["-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
            # packed FP32 instructions are kept out of every kernel
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            f"-I{INC}", f"-I{CSRC}"]
COMMON = COMMON_C + ["-shared"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(err + "\n")
        raise RuntimeError(f"build failed: {cmd[-1]}")


OBJ = os.path.join(PKG, "_obj")
NUNITS = 7          # translation units of FFT pass kernels, compiled in parallel


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=True, incremental=False):
    """Compile libcsipm.so: the FFT pass kernels in NUNITS translation units
    plus csipm.hip, in parallel, then link.  incremental=True recompiles only
    units whose sources changed (timestamps)."""
    subprocess.run([sys.executable, os.path.join(CSRC, "gen_radix.py")], check=True)
    sys.path.insert(0, CSRC)
    import fft_instances
    sys.path.pop(0)
    units = fft_instances.write(OBJ, NUNITS)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    headers.append(os.path.join(INC, "csipm.h"))
    # the kernel units include csipm_kernels.hpp and what it includes
    kheaders = [os.path.join(CSRC, f) for f in ("csipm_kernels.hpp", "csipm_device.hpp",
                                                "radix_consts.hpp")]
    jobs = []
    for u in units:
        o = u[:-4] + ".o"
        jobs.append((u, o, kheaders + [u]))
    main_o = os.path.join(OBJ, "csipm.o")
    jobs.append((os.path.join(CSRC, "csipm.hip"), main_o,
                 headers + [os.path.join(CSRC, "csipm.hip"), os.path.join(OBJ, "fft_extern.inc")]))

    def one(job):
        src, o, deps = job
        if incremental and not _stale(o, deps):
            return
        _run([HIPCC, "-O3", *COMMON_C, f"-I{OBJ}", "-c", "-o", o, src])

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(one, jobs))
    out = os.path.join(PKG, "libcsipm.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
          *[j[1] for j in jobs]])
    os.replace(tmp, out)
    chk = os.path.join(PKG, "libcsipm_geomcheck.so")
    _run([HIPCC, "-O2", *COMMON, "-o", chk + ".tmp", os.path.join(CSRC, "geom_check.hip")])
    os.replace(chk + ".tmp", chk)
    return out


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
