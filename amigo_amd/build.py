"""Build the HIP engine (libkktldl.so) for gfx950 in-tree.

    python -m amigo_amd.build [--incremental]

Compiles amigo_amd/csrc/{symbolic.cpp, nd_ordering.cpp, kkt_ldl.hip, kkt_front.hip, kkt_panel.hip (once per
front element type), assembly.hip, ipm.hip} with hipcc, in parallel, into
amigo_amd/libkktldl.so.  Every HIP translation unit is built with packed FP32
arithmetic disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
ipm.hip is built without floating-point contraction, so that its elementwise
results are the reference's expressions evaluated operation by operation.

--incremental recompiles only objects older than their source or a header of
csrc/ or include/; the default recompiles everything.
"""

import glob
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

# (source, object name, extra flags)
UNITS = (
    ("symbolic.cpp", "symbolic.o", []),
    ("nd_ordering.cpp", "nd_ordering.o", []),
    ("kkt_ldl.hip", "kkt_ldl.o", ["-x", "hip"]),
    ("kkt_front.hip", "kkt_front.o", ["-x", "hip"]),
    ("kkt_panel.hip", "kkt_panel_f32.o", ["-x", "hip", "-DKKT_PANEL_T=float"]),
    ("kkt_panel.hip", "kkt_panel_f64.o", ["-x", "hip", "-DKKT_PANEL_T=double"]),
    ("assembly.hip", "assembly.o", ["-x", "hip"]),
    ("ipm.hip", "ipm.o", ["-x", "hip", "-ffp-contract=off"]),
)


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + glob.glob(os.path.join(CSRC, "*.hpp")) + \
        glob.glob(os.path.join(ROOT, "include", "*.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=True, incremental=False):
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs, procs = [], []
    for src, name, extra in UNITS:
        obj = os.path.join(build_dir, name)
        objs.append(obj)
        path = os.path.join(CSRC, src)
        if incremental and not _stale(path, obj):
            continue
        cmd = [HIPCC] + HIP_FLAGS + extra + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("compilation failed: " + " | ".join(" ".join(c) for c in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
