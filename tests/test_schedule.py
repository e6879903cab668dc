"""The factorization's schedule.

By default a front starts at the tick after its last child's last panel, and
every tick runs one panel of each active large front; the panel launches then
follow the tree's longest path instead of the sum over tree levels of each
level's longest panel chain.  KKT_LDL_LEVEL_SCHEDULE=1 (read when the handle
is created) selects the level schedule.  The two run the same kernels on each
front in the same order, so they must give the same factorization bit for
bit: pivots, delays, 2x2 blocks, inertia and the solution.

Cases: the Poisson control KKT under nested dissection (also at its full
256 x 256 size in FP32), 3D elasticity, a forest with two roots, random
indefinite matrices whose delays overflow the planned front slack (the
re-planning path, which rebuilds the schedule), and a Cholesky factorization.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from amigo_amd import workloads as W
from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SolverType, SparseLDL

from .test_panel_speculation import _case as _spec_case


def _laplacian(N):
    """5-point Laplacian plus identity (SPD) on an N x N grid."""
    T = sp.diags([-np.ones(N - 1), 2 * np.ones(N), -np.ones(N - 1)], [-1, 0, 1])
    A = (sp.kron(sp.eye(N), T) + sp.kron(T, sp.eye(N)) + sp.eye(N * N)).tocsr()
    A.sort_indices()
    n = A.shape[0]
    return (n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy(),
            np.zeros(n), np.random.default_rng(0).standard_normal(n))


def _case(name):
    if name == "poisson2d_256":
        P = W.poisson2d()
        return P.n, P.rowp, P.cols, P.vals, P.diag, P.rhs, OrderingType.NESTED_DISSECTION
    if name == "cholesky":
        return _laplacian(120) + (OrderingType.NESTED_DISSECTION,)
    return _spec_case(name)


def _run(monkeypatch, case, precision, level, **kw):
    n, rowp, cols, vals, diag, rhs, order = case
    monkeypatch.setenv("KKT_LDL_LEVEL_SCHEDULE", level)
    s = SparseLDL(CSRMat(n, n, rowp, cols, vals), order=order, precision=precision, **kw)
    s.factor(diag)
    x = rhs.copy()
    s.solve(x)
    return s, x


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", ["poisson2d", "elasticity3d", "forest", "random_1", "random_2",
                                  "cholesky", "poisson2d_256"])
def test_tick_schedule_equals_level_schedule(monkeypatch, name, precision):
    if name == "poisson2d_256" and precision == FP64:
        pytest.skip("the full-size case runs in the benchmark's precision only")
    print("case", name, precision, flush=True)
    case = _case(name)
    kw = dict(delay_growth=1.0) if name.startswith("random") else {}
    if name == "cholesky":
        kw["solver_type"] = SolverType.CHOLESKY
    a, xa = _run(monkeypatch, case, precision, "0", **kw)
    b, xb = _run(monkeypatch, case, precision, "1", **kw)
    sa, sb = a.stats(), b.stats()
    assert sb["panel_launches"] == sb["level_panel_sum"]
    assert sa["panel_launches"] <= sb["panel_launches"]
    if name.startswith("random"):
        print("replans", sa["replans"], sb["replans"])
    assert a.get_inertia() == b.get_inertia()
    na, nb = a.node_info(), b.node_info()
    for k in na:
        assert np.array_equal(na[k], nb[k]), k
    assert np.array_equal(xa, xb)
    if name == "cholesky":
        # no inertia in Cholesky mode (sparse_ldl.h:205-213); the solution
        # is a solution
        assert a.get_inertia() == (0, 0)
        n, rowp, cols, vals, diag, rhs, _ = case
        K = sp.csr_matrix((vals, cols, rowp), shape=(n, n))
        r = np.linalg.norm(K @ xa - rhs) / np.linalg.norm(rhs)
        assert r < (1e-12 if precision == FP64 else 1e-5)
    else:
        assert sum(a.get_inertia()) == case[0]


def test_tick_schedule_plan():
    """On the CPU (analysis only): on the Poisson control KKT under nested
    dissection the tick schedule needs far fewer panel launches than the
    level schedule, and never more than the level schedule on any tree."""
    P = W.poisson2d(N=128)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.NESTED_DISSECTION)
    st = s.stats()
    print("levels", st["num_levels"], "level panel sum", st["level_panel_sum"],
          "panel launches", st["panel_launches"])
    assert 0 < st["panel_launches"] <= 0.8 * st["level_panel_sum"]
    C = W.cartpole(N=500)
    s = SparseLDL(CSRMat(C.n, C.n, C.rowp, C.cols, C.vals), order=OrderingType.AMD)
    st = s.stats()
    assert st["panel_launches"] <= st["level_panel_sum"]


def test_level_schedule_switch(monkeypatch):
    monkeypatch.setenv("KKT_LDL_LEVEL_SCHEDULE", "1")
    P = W.poisson2d(N=64)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.NESTED_DISSECTION)
    st = s.stats()
    assert st["panel_launches"] == st["level_panel_sum"] > 0
