"""Build libocfd_hip.so (the gfx950 HIP library) in-tree with hipcc.

    python opencfd-sc_amd/build.py            # build (parallel per-file compile)
    python opencfd-sc_amd/build.py --clean

All kernels are compiled with -ffp-contract=off: the FP64 ones so that every
floating-point expression rounds exactly as the reference's (no fused
multiply-add), the FP32 ones of the mixed-precision form so that their
rounding does not depend on the code around them (a box of a batch equals the
same box stepped alone bit for bit).  Packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libocfd_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["ocfd_inviscid.hip", "ocfd_inviscid_mixed.hip", "ocfd_hybrid.hip", "ocfd_viscous.hip",
           "ocfd_viscous_mixed.hip", "ocfd_misc.hip", "ocfd_batchq.hip", "ocfd_batchq_mixed.hip",
           "ocfd_api.hip"]

FLAGS = [
    "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, msgs[-6000:]))
    if msgs and verbose:
        print(msgs[-3000:], flush=True)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=%s" % ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % (r.stdout + r.stderr)[-4000:])
    os.replace(tmp, LIB)
    return LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    else:
        print(build())
