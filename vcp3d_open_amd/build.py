"""Builds the in-tree gfx950 libraries with hipcc: vcp3d_open_amd/libvcp3d_amd.so
(the path) and vcp3d_open_amd/libvcp3d_ab.so (the VALU / MFMA interpolation
measurement that bench.py --workload interpolation runs)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libvcp3d_amd.so")
AB_LIB = os.path.join(HERE, "libvcp3d_ab.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "geometry.hip", "setup.cpp", "capi.cpp", "batch.cpp", "contraction_ab.hip"]
AB_SOURCES = ["contraction_ab.hip"]
# per-source flags: the device metric terms repeat the host set-up's
# arithmetic exactly, so neither side contracts multiply-adds
EXTRA = {"geometry.hip": ["-ffp-contract=off"], "setup.cpp": ["-ffp-contract=off"]}
COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=gfx950",
          # no packed FP32 instructions anywhere in device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *COMMON, *EXTRA.get(src, []), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(out)
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("hipcc failed for " + ", ".join(failed))
    for lib, srcs in ((LIB, [s for s in SOURCES if s not in AB_SOURCES]), (AB_LIB, AB_SOURCES)):
        tmp = lib + ".tmp"
        objs_l = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in srcs]
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs_l]
        subprocess.run(cmd, check=True)
        os.replace(tmp, lib)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
