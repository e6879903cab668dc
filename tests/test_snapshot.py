"""The KKT snapshot wire format (amigo_amd/snapshot.py): the golden KKT
fixtures (the reference's own factorization outputs, tests/golden/kkt_*.npz)
and the workload generators round-trip bit for bit; the variable types give
the reference's primal / constraint index lists; damaged files are rejected."""

import os

import numpy as np
import pytest

from amigo_amd import snapshot as S
from amigo_amd import workloads as W

from .conftest import GOLDEN

KKT = sorted(f for f in os.listdir(GOLDEN) if f.startswith("kkt_"))


@pytest.mark.parametrize("name", KKT)
def test_golden_kkt_round_trip(tmp_path, name):
    c = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    snap = S.Snapshot(c["rowp"], c["cols"], c["vals"],
                      c["diag"] if bool(c["has_diag"]) else None, c["rhs"],
                      meta={"fixture": name, "inertia": [int(v) for v in c["inertia"]]})
    path = tmp_path / "k.kkt"
    S.save(path, snap)
    back = S.load(path)
    for k in ("rowp", "cols", "vals"):
        assert np.array_equal(getattr(back, k), c[k]), k
        assert getattr(back, k).dtype == getattr(snap, k).dtype
    if bool(c["has_diag"]):
        assert np.array_equal(back.diag, c["diag"])
    assert np.array_equal(back.rhs, c["rhs"])
    assert back.meta == {"fixture": name, "inertia": [int(v) for v in c["inertia"]]}


@pytest.mark.parametrize("gen", [W.rosenbrock, lambda: W.cartpole(N=200),
                                 lambda: W.poisson2d(N=16), lambda: W.elasticity3d(N=3)],
                         ids=["rosenbrock", "cartpole", "poisson2d", "elasticity3d"])
def test_workload_round_trip_and_types(tmp_path, gen):
    P = gen()
    snap = S.from_kkt(P, {"iteration": 7})
    path = tmp_path / "w.kkt"
    S.save(path, snap)
    back = S.load(path)
    assert np.array_equal(back.vals, P.vals) and np.array_equal(back.diag, P.diag)
    assert np.array_equal(back.rhs, P.rhs)
    # the reference's index lists (optimization_problem.h:1370-1388)
    assert set(back.primal_indices()) | set(back.constraint_indices()) == set(range(P.n))
    assert not set(back.primal_indices()) & set(back.constraint_indices())
    fixed = set(back.fixed.tolist())
    assert set(back.constraint_indices()) == set(np.asarray(P.dual_idx).tolist()) - fixed
    # fixed rows are identity rows (their pattern kept, off-diagonal values 0)
    K = back.matrix(with_diag=False)
    for i in back.fixed[:20]:
        row = K.getrow(int(i))
        assert np.array_equal(row.toarray().ravel(), np.eye(P.n)[i])
    assert back.meta["iteration"] == 7


def test_multiple_rhs_and_reference_objects(tmp_path):
    """from_problem takes the reference's objects (CSRMat, Vectors) and calls
    their host-sync methods as AmigoSolver does."""
    from amigo_amd.ldl import CSRMat, Vector
    P = W.poisson2d(N=8)
    calls = []

    class H(CSRMat):
        def copy_data_device_to_host(self):
            calls.append("hessian")

    class V(Vector):
        def copy_device_to_host(self):
            calls.append("vector")

    d = V(P.n)
    d.get_array()[:] = P.diag
    snap = S.from_problem(H(P.n, P.n, P.rowp, P.cols, P.vals), d,
                          rhs=np.column_stack([P.rhs, 2 * P.rhs]), var_types=np.full(P.n, 4))
    assert calls == ["hessian", "vector"]
    S.save(tmp_path / "m.kkt", snap)
    back = S.load(tmp_path / "m.kkt")
    assert back.rhs.shape == (P.n, 2) and np.array_equal(back.rhs[:, 1], 2 * P.rhs)
    assert np.array_equal(back.diag, P.diag)


def test_damaged_files_are_rejected(tmp_path):
    P = W.poisson2d(N=8)
    path = tmp_path / "d.kkt"
    S.save(path, S.from_kkt(P))
    raw = bytearray(path.read_bytes())
    cases = {
        "magic": bytes(b"XXXXXXXX") + bytes(raw[8:]),
        "version": bytes(raw[:8]) + (7).to_bytes(4, "little") + bytes(raw[12:]),
        "truncated": bytes(raw[:-20]),
        "flipped": bytes(raw[:100]) + bytes([raw[100] ^ 1]) + bytes(raw[101:]),
    }
    for what, data in cases.items():
        bad = tmp_path / f"{what}.kkt"
        bad.write_bytes(data)
        with pytest.raises(ValueError):
            S.load(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("name", KKT)
def test_replay_snapshot_on_the_engine(tmp_path, name):
    """A golden KKT written and read back as a snapshot, replayed on the GPU
    engine: the reference's inertia, residual <= 1e-10."""
    from amigo_amd.ldl import CSRMat, OrderingType, SparseLDL
    c = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    S.save(tmp_path / "r.kkt", S.Snapshot(c["rowp"], c["cols"], c["vals"],
                                          c["diag"] if bool(c["has_diag"]) else None, c["rhs"]))
    snap = S.load(tmp_path / "r.kkt")
    s = SparseLDL(CSRMat(snap.n, snap.n, snap.rowp, snap.cols, snap.vals),
                  order=OrderingType.AMD)
    assert s.factor(snap.diag) == 0
    assert s.get_inertia() == tuple(int(v) for v in c["inertia"])
    x = snap.rhs.copy()
    assert s.solve(x, refine=True) == 0
    K = snap.matrix()
    assert np.linalg.norm(K @ x - snap.rhs) / np.linalg.norm(snap.rhs) <= 1e-10
