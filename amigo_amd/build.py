"""Build the HIP engine (libkktldl.so) for gfx950 in-tree.

    python -m amigo_amd.build

Compiles amigo_amd/csrc/{symbolic.cpp,kkt_ldl.hip,assembly.hip,ipm.hip} with
hipcc into amigo_amd/libkktldl.so.  Every HIP translation unit is built with
packed FP32 arithmetic disabled (-Xclang -target-feature -Xclang
-packed-fp32-ops).  ipm.hip is built without floating-point contraction, so
that its elementwise results are the reference's expressions evaluated
operation by operation.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libkktldl.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("KKT_LDL_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs = []
    for src in ("symbolic.cpp", "kkt_ldl.hip", "assembly.hip", "ipm.hip"):
        obj = os.path.join(build_dir, src + ".o")
        extra = ["-x", "hip"] if src.endswith(".hip") else []
        if src == "ipm.hip":
            extra += ["-ffp-contract=off"]
        _run([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
