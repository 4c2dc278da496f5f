"""In-tree build of libmspms_gpu.so (HIP, gfx950) with hipcc.

    python -m mspms2_amd.build          # or mspms2_amd.build.build()

Each .hip / .cpp source of mspms2_amd/csrc is compiled to an object under
mspms2_amd/_build/ and linked into mspms2_amd/libmspms_gpu.so.  The drop-in
static library libmspms_dropin.a (frclong / frcshort over the reference's
globals) is built next to it.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmspms_gpu.so")
DROPIN = os.path.join(HERE, "libmspms_dropin.a")
ARCH = os.environ.get("MSPMS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["k_cluster.hip", "k_pair.hip", "k_intra.hip", "k_recip.hip", "k_integrate.hip",
               "k_sf.hip", "k_hmc.hip", "k_nvt.hip", "k_npt.hip", "k_sfgrid.hip", "k_sfgrid_noalloc.hip", "k_myinterp.hip",
               "engine.cpp", "terms_host.cpp"]
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        cmd = ["hipcc", "-x", "hip"] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # drop-in objects: plain C over the reference's globals
    # (dropin_globals.c: the globals read into the C ABI's tables;
    # dropin_terms.c: the term-level functions, an archive member of its own)
    dobjs = []
    for name in ("dropin.c", "dropin_globals.c", "dropin_terms.c"):
        dobj = os.path.join(BUILD, os.path.splitext(name)[0] + ".o")
        _run(["gcc", "-std=c99", "-O2", "-fPIC", "-fcommon", "-Wall", "-c",
              os.path.join(CSRC, name), "-o", dobj])
        dobjs.append(dobj)
    if os.path.exists(DROPIN):
        os.remove(DROPIN)
    _run(["ar", "rcs", DROPIN] + dobjs)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
