"""Build recipe for the CPU oracle (test infrastructure only).

Compiles the plain-C restatement under oracle/ into oracle/_build/liborc.so
with gcc.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg load the result.  The reference Rust engine cannot be compiled here
(no cargo/rustc; rayon is not vendored), so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["rng.c", "sample.c", "cmg.c", "engine.c", "numerical.c"]
OUTPUT = os.path.join(HERE, "_build", "liborc.so")


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUTPUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-pthread",
           "-fno-fast-math", "-ffp-contract=off", "-Wall",
           *[os.path.join(HERE, s) for s in SOURCES], "-lm", "-o", OUTPUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(OUTPUT + ".tmp", OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
