"""Build recipe for the MI355X engine library (fevc_amd/libvckss_mi355x.so).

All HIP and host sources are compiled by hipcc for gfx950 into one shared
library exporting the C ABI declared in include/vckss_mi355x.h and
include/vckss_mi355x_ext.h.  FP contraction is off so that the FP64 level
products keep the reference's per-row summation order; packed FP32 ops are
disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["device.hip", "host_prepare.cpp", "host_cmg.cpp", "vckss_abi.cpp", "vckss_abi_refusals.cpp"]
OUTPUT = os.path.join(HERE, "libvckss_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build(verbose: bool = False) -> str:
    objs = []
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    common = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              f"-I{os.path.join(os.path.dirname(HERE), 'include')}"]
    for src in SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        cmd = [HIPCC, *common, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    link = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", OUTPUT + ".tmp"]
    if verbose:
        print(" ".join(link), file=sys.stderr)
    subprocess.run(link, check=True)
    os.replace(OUTPUT + ".tmp", OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
