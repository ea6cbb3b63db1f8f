"""Build the HIP library in-tree: baccolonygrowth_amd/libbcg.so (gfx950).

    python -m baccolonygrowth_amd.build

Flags: -ffp-contract=off keeps the mechanics and environment kernels free of
FMA contraction (every +,-,*,/ rounds as in the reference's x86-64 build);
packed FP32 ops are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbcg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# (source, extra flags) - each compiled to an object, then linked
SOURCES = [
    ("mech_kernels.hip", ["-ffp-contract=off"]),
    ("bcg_api.hip", ["-ffp-contract=off"]),
    ("mg_kernels.hip", []),
    ("mg_api.hip", []),
    ("env_kernels.hip", ["-ffp-contract=off"]),
    ("env_api.hip", ["-ffp-contract=off"]),
]

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src, extra in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [HIPCC] + COMMON + extra + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
