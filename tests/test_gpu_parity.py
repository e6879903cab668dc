"""GPU parity: the HIP engine (through the C ABI) against the reference's
golden fixtures and the oracle.

Bar: inertia identical; per-node pivot outcome (pivots, delayed, 2x2)
identical below the root; relative KKT residual of the solution within the
reference's own and <= 1e-10 after FP64 refinement; multi-RHS error < 1e-11
on the reference's Poisson tests; bitwise-repeatable factorizations.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from amigo_amd import workloads as W
from amigo_amd.ldl import CSRMat, OrderingType, SolverType, SparseLDL, Vector
from oracle.pyoracle import AMD, OracleLDL

from .conftest import engine_order, golden_cases

pytestmark = pytest.mark.gpu

CASES = golden_cases()


def _engine(c, **kw):
    n = int(c["n"])
    return SparseLDL(CSRMat(n, n, c["rowp"], c["cols"], c["vals"]),
                     solver_type=int(c["solver_type"]), ustab=float(c["ustab"]),
                     pivot_tol=float(c["pivot_tol"]), order=engine_order(c), **kw)


def _K(c):
    n = int(c["n"])
    K = sp.csr_matrix((c["vals"], c["cols"], c["rowp"]), shape=(n, n))
    if bool(c["has_diag"]):
        K = K + sp.diags(c["diag"])
    return K


def _relres(K, x, b):
    n = K.shape[0]
    b = b.reshape(n, -1)
    return np.linalg.norm(K @ x.reshape(n, -1) - b) / np.linalg.norm(b)


@pytest.mark.parametrize("c", CASES, ids=[str(c["name"]) for c in CASES])
def test_factor_solve_inertia_vs_reference(c):
    print("case", c["name"], flush=True)
    s = _engine(c)
    diag = c["diag"] if bool(c["has_diag"]) else None
    assert s.factor(diag) == 0
    if int(c["solver_type"]) == SolverType.LDL:
        assert s.get_inertia() == tuple(int(v) for v in c["inertia"])
    info = s.node_info()
    nonroot = c["ni_root"] == 0
    assert np.array_equal(info["npiv"], c["ni_npiv"])
    for k in ("ndel", "n2x2"):
        assert np.array_equal(info[k][nonroot], c["ni_" + k][nonroot]), k
    x = np.array(c["rhs"], dtype=np.float64, order="F", copy=True)
    s.solve(x)
    K = _K(c)
    r_ref = _relres(K, c["x"], c["rhs"])
    assert _relres(K, x, c["rhs"]) <= max(1e-10, 10 * r_ref)
    assert np.abs(x - c["x"]).max() <= 1e-6 * np.abs(c["x"]).max()
    if r_ref <= 1e-10:
        y = np.array(c["rhs"], dtype=np.float64, order="F", copy=True)
        s.solve(y, refine=True)
        assert _relres(K, y, c["rhs"]) <= 1e-10


POIS = [c for c in CASES if str(c["name"]).startswith("poisson")]


@pytest.mark.parametrize("c", POIS, ids=[str(c["name"]) for c in POIS])
def test_multi_rhs_reference_assertion(c):
    """tests/unit/test_ldl.py:230-314: 7 right-hand sides, error < 1e-11."""
    s = _engine(c)
    s.factor()
    if int(c["solver_type"]) == SolverType.LDL:
        assert s.get_inertia() == (int(c["n"]), 0)
    X = np.array(c["rhs"], order="F", copy=True)
    s.solve(X)
    err = max(np.linalg.norm(X[:, k] - c["exact"]) for k in range(X.shape[1]))
    assert err < 1e-11
    # C-ordered input goes through a Fortran copy and back
    Xc = np.ascontiguousarray(c["rhs"])
    s.solve(Xc)
    assert np.allclose(Xc, X)


def test_vector_and_device_tensor_paths_agree():
    import torch
    P = W.poisson2d(N=24)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD)
    v = Vector(P.n)
    v.get_array()[:] = P.rhs
    s.factor(P.diag)
    s.solve(v)
    vals = torch.tensor(P.vals, device="cuda")
    diag = torch.tensor(P.diag, device="cuda")
    sd = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, vals), order=OrderingType.AMD)
    sd.factor(diag)
    xd = torch.tensor(P.rhs, device="cuda")
    sd.solve(xd)
    torch.cuda.synchronize()
    assert np.array_equal(xd.cpu().numpy(), v.get_array())
    assert sd.get_inertia() == s.get_inertia()


def test_factorization_is_bitwise_repeatable():
    P = W.cartpole(N=300)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD)
    xs = []
    for _ in range(3):
        s.factor(P.diag)
        x = P.rhs.copy()
        s.solve(x)
        xs.append(x)
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[1], xs[2])


def test_delayed_pivot_overflow_replans():
    """Minimal front slack forces the delayed-pivot re-planning path; the
    result still matches the oracle's inertia and pivot outcome."""
    P = W.poisson2d(N=32)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD,
                  delay_growth=1.0)
    s.factor(P.diag)
    o = OracleLDL(P.n, P.rowp, P.cols, order=AMD)
    o.factor(P.vals, P.diag)
    assert s.get_inertia() == o.inertia()
    assert s.node_info()["ndel"].sum() == o.node_info()["ndel"].sum() > 0
    x = P.rhs.copy()
    s.solve(x, refine=True)
    assert _relres(P.matrix(), x, P.rhs) <= 1e-10
    # refactor with different values reuses the enlarged plan
    s.factor(P.diag * 2.0)
    o.factor(P.vals, P.diag * 2.0)
    assert s.get_inertia() == o.inertia()


@pytest.mark.parametrize("name,kw", [("poisson2d", {}), ("cartpole", {}),
                                     ("elasticity3d", dict(N=12))])
def test_full_size_workloads_vs_oracle(name, kw):
    """Benchmark-size KKTs: inertia, delayed-pivot and 2x2 counts identical
    to the oracle; refined FP64 residual <= 1e-10."""
    print("workload", name, kw, flush=True)
    P = W.GENERATORS[name](**kw)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD)
    s.factor(P.diag)
    o = OracleLDL(P.n, P.rowp, P.cols, order=AMD)
    o.factor(P.vals, P.diag)
    assert s.get_inertia() == o.inertia()
    a, b = s.node_info(), o.node_info()
    assert a["ndel"].sum() == b["ndel"].sum()
    assert a["n2x2"][b["root"] == 0].sum() == b["n2x2"][b["root"] == 0].sum()
    x = P.rhs.copy()
    s.solve(x, refine=True)
    assert _relres(P.matrix(), x, P.rhs) <= 1e-10


# --- FP32 factor + FP64 refinement ----------------------------------------

LDLC = [c for c in CASES if int(c["solver_type"]) == SolverType.LDL]


@pytest.mark.parametrize("c", LDLC, ids=[str(c["name"]) for c in LDLC])
def test_fp32_factor_refined_vs_reference(c):
    """FP32 fronts, FP64 pivot decisions and D: identical inertia; FP64
    refinement reaches <= 1e-10 where the reference's FP64 solve does."""
    from amigo_amd.ldl import FP32_REFINE
    s = _engine(c, precision=FP32_REFINE, refine_tol=1e-10, max_refine=30)
    diag = c["diag"] if bool(c["has_diag"]) else None
    assert s.factor(diag) == 0
    assert s.get_inertia() == tuple(int(v) for v in c["inertia"])
    K = _K(c)
    r_ref = _relres(K, c["x"], c["rhs"])
    x = np.array(c["rhs"], dtype=np.float64, order="F", copy=True)
    s.solve(x, refine=True)
    if r_ref <= 1e-10:
        assert _relres(K, x, c["rhs"]) <= 1e-10
        assert np.abs(x - c["x"]).max() <= 1e-6 * np.abs(c["x"]).max()


@pytest.mark.parametrize("name,kw", [("poisson2d", {}), ("cartpole", {})])
def test_fp32_full_size_inertia_and_residual(name, kw):
    from amigo_amd.ldl import FP32_REFINE
    P = W.GENERATORS[name](**kw)
    s64 = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                    order=OrderingType.NESTED_DISSECTION)
    s64.factor(P.diag)
    s32 = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                    order=OrderingType.NESTED_DISSECTION, precision=FP32_REFINE,
                    max_refine=30)
    s32.factor(P.diag)
    assert s32.get_inertia() == s64.get_inertia()
    x = P.rhs.copy()
    s32.solve(x, refine=True)
    st = s32.stats()
    print("fp32 refine steps", st["refine_steps"], "residual", st["residual"])
    assert _relres(P.matrix(), x, P.rhs) <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_multi_rhs_columns_bitwise_equal_single_solves(prec):
    """The batched multi-right-hand-side kernels (several right-hand sides
    per workgroup in the large-node steps and gathers) sum every entry in
    the order of the one-right-hand-side kernels: each column of a 19-RHS
    solve is bitwise the single solve of that column (C2, ND)."""
    import torch
    from amigo_amd.ldl import FP32_REFINE
    P = W.poisson2d()
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                  order=OrderingType.NESTED_DISSECTION,
                  precision=FP32_REFINE if prec == "fp32" else 0)
    s.factor(P.diag)
    g = torch.Generator(device="cuda").manual_seed(3)
    B = torch.randn(19, P.n, dtype=torch.float64, device="cuda", generator=g).t()
    X = B.clone()
    s.solve(X)
    for j in (0, 3, 4, 7, 8, 18):
        xj = B[:, j].clone()
        s.solve(xj)
        assert torch.equal(xj, X[:, j]), j


def _solves_bitwise_equal_under_switch(name, prec, var):
    """Handles created with the environment switch `var` at 1 and at 0
    give the same solutions bit for bit, for 1 and 5 right-hand sides and the
    refined solve, under nested dissection."""
    import os

    import torch
    from amigo_amd.ldl import FP32_REFINE
    P = W.GENERATORS[name]()
    hs = []
    for flag in ("1", "0"):
        os.environ[var] = flag
        try:
            s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                          order=OrderingType.NESTED_DISSECTION,
                          precision=FP32_REFINE if prec == "fp32" else 0)
        finally:
            del os.environ[var]
        assert s.factor(P.diag) == 0
        hs.append(s)
    g = torch.Generator(device="cuda").manual_seed(9)
    B = torch.randn(5, P.n, dtype=torch.float64, device="cuda", generator=g).t()
    for nrhs in (1, 5):
        xa, xb = B[:, :nrhs].clone(), B[:, :nrhs].clone()
        hs[0].solve(xa)
        hs[1].solve(xb)
        assert torch.equal(xa, xb), nrhs
    xa = torch.as_tensor(P.rhs, device="cuda").clone()
    xb = xa.clone()
    assert hs[0].solve(xa, refine=True) == 0
    assert hs[1].solve(xb, refine=True) == 0
    assert torch.equal(xa, xb)
    assert _relres(P.matrix(), xa.cpu().numpy(), P.rhs) <= 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["poisson2d", "cartpole"])
def test_first_step_gather_bitwise_equal_gather_launch(name, prec):
    """Levels whose large nodes all hold pivots gather the front vector in
    the first forward step's workgroups (each the step's rows and its own
    tile's rows, with k_gather_big's sums) instead of a gather launch before
    the step (KKT_LDL_FUSE_GATHER=0 when the handle is created; solves of
    one right-hand side): the same solutions bit for bit."""
    _solves_bitwise_equal_under_switch(name, prec, "KKT_LDL_FUSE_GATHER")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("name", ["poisson2d", "cartpole"])
def test_merged_level_launches_bitwise_equal(name, prec):
    """A level's small fronts run in the launch of its large nodes' fused
    first forward step and of their backward L21^T products (k_fwd_level,
    k_bwd_level; one right-hand side) instead of launches of their own
    (KKT_LDL_MERGE_LEVELS=0): the same solutions bit for bit."""
    _solves_bitwise_equal_under_switch(name, prec, "KKT_LDL_MERGE_LEVELS")
