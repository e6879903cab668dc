"""Shared test setup: repo on sys.path, the `gpu` marker, built libraries."""

import glob
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    # Build the oracle (the checker) and the engine only if nothing is built.
    if not os.path.exists(os.path.join(ROOT, "oracle", "_build", "libldl_oracle.so")):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(ROOT, "amigo_amd", "libkktldl.so")):
        subprocess.run([sys.executable, "-m", "amigo_amd.build"], cwd=ROOT, check=True)


def golden_cases():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))):
        z = np.load(path, allow_pickle=False)
        out.append({k: z[k] for k in z.files})
    return out


@pytest.fixture(scope="session")
def golden():
    return golden_cases()


def engine_order(c):
    """Engine ordering code of a golden fixture.  The fixtures record the
    oracle's numbering (0 natural, 1 AMD), the engine takes the reference's
    OrderingType numbering (ND 0, AMD 1, MULTI_COLOR 2, NATURAL 3)."""
    from amigo_amd.ldl import OrderingType
    return {0: OrderingType.NATURAL, 1: OrderingType.AMD}[int(c["order"])]


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
