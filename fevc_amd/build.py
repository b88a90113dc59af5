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
SOURCES = ["device.hip", "smoother.hip", "exact.hip", "host_prepare.cpp", "host_cmg.cpp", "numerical.cpp", "vckss_abi.cpp",
           "vckss_abi_refusals.cpp"]
PARAMS = os.path.join(HERE, "smoother_params.json")
OUTPUT = os.path.join(HERE, "libvckss_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def write_params(obj_dir: str) -> None:
    """The built-in trained smoother weights (fevc_amd/smoother_params.json,
    written by fevc_amd/train_smoother.py) as a C array; none if absent."""
    import json

    values = []
    if os.path.exists(PARAMS):
        with open(PARAMS) as fh:
            values = [float(v) for v in json.load(fh)["params"]]
    body = ", ".join(repr(v) for v in values) if values else "0.0"
    text = ("// generated from fevc_amd/smoother_params.json by fevc_amd/build.py\n"
            "#pragma once\n#include <cstddef>\n"
            f"static const double kBuiltinSmoother[] = {{{body}}};\n"
            f"static const size_t kBuiltinSmootherCount = {len(values)};\n")
    path = os.path.join(obj_dir, "smoother_params.inc")
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as fh:
            fh.write(text)


def build(verbose: bool = False) -> str:
    objs = []
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    write_params(obj_dir)
    common = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
              f"-I{os.path.join(os.path.dirname(HERE), 'include')}", f"-I{obj_dir}"]
    for src in SOURCES:
        obj = os.path.join(obj_dir, src + ".o")
        cmd = [HIPCC, *common, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "c++"]  # host code only
        if src.endswith(".hip"):
            cmd[1:1] = ["-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
        if src.endswith(".hip"):
            cmd.append("-Wno-unused-function")  # kernels of the shared headers unused in one unit
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
