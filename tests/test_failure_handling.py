"""Failures are reported, never silent.

* NaN or infinite values in K or the diagonal: kkt_ldl_factor returns
  WARN_NONFINITE and counts the non-finite pivots; a refined solve returns
  WARN_NOT_CONVERGED with an infinite residual (a NaN residual never reads as
  converged); MI355XSolver.factor raises.
* A singular matrix factors to the end as in the reference, zero pivots
  counted positive in the inertia (include/sparse_ldl.h:1958), and the factor
  returns WARN_SINGULAR with the zero pivots in nzero.
* An FP32 refinement that cannot reach its target stops at the best iterate
  (the reported residual is that of the returned solution) and returns
  WARN_NOT_CONVERGED; MI355XSolver then redoes the solve with an FP64
  factorization of the same matrix and counts the fallback.
* MI355XSolver makes the host/device synchronisation calls of the
  reference's AmigoSolver (amigo/algorithm/solvers/amigo_solver.py:27-55).
* kkt_ldl_refactor (new diagonal, values kept on the device) equals a full
  factorization bit for bit; kkt_ldl_set_pivot_threshold equals creating
  the handle with that threshold.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from amigo_amd import workloads as W
from amigo_amd.ldl import (FP32_REFINE, FP64, WARN_NONFINITE, WARN_NOT_CONVERGED,
                           WARN_SINGULAR, CSRMat, OrderingType, SparseLDL, Vector)
from amigo_amd.solver import MI355XSolver
from oracle.pyoracle import AMD, OracleLDL

from .conftest import engine_order, golden_cases

pytestmark = pytest.mark.gpu


def _relres(P, x):
    return np.linalg.norm(P.matrix() @ x - P.rhs) / np.linalg.norm(P.rhs)


@pytest.mark.parametrize("where", ["values", "diag_inf"])
def test_nonfinite_input_is_reported(where):
    P = W.poisson2d(N=32)
    vals, diag = P.vals.copy(), P.diag.copy()
    if where == "values":
        vals[len(vals) // 3] = np.nan
    else:
        diag[P.n // 2] = np.inf
    for prec in (FP64, FP32_REFINE):
        s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, vals), order=OrderingType.AMD,
                      precision=prec)
        assert s.factor(diag) == WARN_NONFINITE
        assert s.stats()["num_nonfinite_pivots"] > 0
        x = P.rhs.copy()
        assert s.solve(x, refine=True) == WARN_NOT_CONVERGED
        assert s.stats()["residual"] == np.inf


def test_solver_raises_on_nonfinite_factor():
    P = W.poisson2d(N=16)
    vals = P.vals.copy()
    vals[7] = np.nan
    H = CSRMat(P.n, P.n, P.rowp, P.cols, vals)
    solver = MI355XSolver({}, _State(H))
    d = Vector(P.n)
    d.get_array()[:] = P.diag
    with pytest.raises(RuntimeError, match="NaN or infinite"):
        solver.factor(H, d)


def test_singular_matrix_warns_and_counts_zero_pivots():
    """Two rows and columns of a Poisson KKT zeroed out (diagonal kept as an
    explicit zero): two zero pivots, counted positive as the oracle does."""
    P = W.poisson2d(N=24)
    dead = np.array([5, P.n // 2])
    rowp, cols = P.rowp, P.cols
    rows = np.repeat(np.arange(P.n), np.diff(rowp))
    vals = P.vals.copy()
    vals[np.isin(rows, dead) | np.isin(cols, dead)] = 0.0   # pattern kept, values zero
    diag = P.diag.copy()
    diag[dead] = 0.0
    s = SparseLDL(CSRMat(P.n, P.n, rowp, cols, vals), order=OrderingType.AMD)
    assert s.factor(diag) == WARN_SINGULAR
    npos, nneg, nzero = s.inertia_and_zero()
    assert nzero == len(dead)
    o = OracleLDL(P.n, rowp, cols, order=AMD)
    o.factor(vals, diag)
    assert (npos, nneg) == o.inertia()
    assert s.stats()["num_zero_pivots"] == len(dead)


def test_refinement_stops_at_the_best_iterate():
    """An unreachable target (0): the refinement stops once a step no longer
    lowers the residual, returns the best iterate, and its reported residual
    is the residual of the solution it returns."""
    P = W.poisson2d(N=64)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                  order=OrderingType.NESTED_DISSECTION, precision=FP32_REFINE,
                  refine_tol=0.0, max_refine=25)
    s.factor(P.diag)
    x = P.rhs.copy()
    assert s.solve(x, refine=True) == WARN_NOT_CONVERGED
    st = s.stats()
    true = _relres(P, x)
    print("refine steps", st["refine_steps"], "reported", st["residual"], "true", true)
    assert st["refine_steps"] < 25
    assert st["residual"] <= 1e-12
    assert 0.5 * st["residual"] <= true <= 2.0 * st["residual"] + 1e-16


class _State:
    def __init__(self, hessian):
        self.hessian = hessian


class _RecordingVector(Vector):
    """A Vector that records the reference adapter's synchronisation calls."""

    def __init__(self, size, log, name):
        super().__init__(size)
        self.log, self.name = log, name

    def copy_device_to_host(self):
        self.log.append((self.name, "d2h"))

    def copy_host_to_device(self):
        self.log.append((self.name, "h2d"))


class _RecordingCSR(CSRMat):
    def __init__(self, P, log):
        super().__init__(P.n, P.n, P.rowp, P.cols, P.vals)
        self.log = log

    def copy_data_device_to_host(self):
        self.log.append(("hessian", "d2h"))


def test_solver_makes_the_reference_sync_calls():
    P = W.poisson2d(N=16)
    log = []
    H = _RecordingCSR(P, log)
    solver = MI355XSolver({}, _State(H))
    d = _RecordingVector(P.n, log, "diag")
    d.get_array()[:] = P.diag
    assert solver.factor(H, d) == 0
    assert log == [("diag", "d2h"), ("hessian", "d2h")]
    b = _RecordingVector(P.n, log, "b")
    b.get_array()[:] = P.rhs
    x = _RecordingVector(P.n, log, "x")
    log.clear()
    solver.solve(b, x)
    assert log == [("x", "d2h"), ("x", "h2d")]
    assert _relres(P, x.get_array()) <= 1e-10
    assert np.array_equal(b.get_array(), P.rhs)


def test_fp32_solver_falls_back_to_fp64():
    """A refinement target the FP32 factor cannot meet (1e-17): the adapter
    redoes the solve with an FP64 factorization and counts it."""
    P = W.poisson2d(N=32)
    H = CSRMat(P.n, P.n, P.rowp, P.cols, P.vals)
    solver = MI355XSolver({"ldl_precision": "fp32", "ldl_refine_tol": 1e-17}, _State(H))
    d = Vector(P.n)
    d.get_array()[:] = P.diag
    solver.factor(H, d)
    b, x = Vector(P.n), Vector(P.n)
    b.get_array()[:] = P.rhs
    solver.solve(b, x)
    assert solver.fallbacks == 1
    assert _relres(P, x.get_array()) <= 1e-12
    solver.solve(b, x)                      # the FP64 factor is reused
    assert solver.fallbacks == 2
    solver.factor(H, d)                     # a new factorization invalidates it
    solver.solve(b, x)
    assert solver.fallbacks == 3
    assert solver.stats()["fallbacks"] == 3


def test_wide_sigma_fp32_reaches_target_or_falls_back():
    """Late-IPM barrier diagonal spanning 1e-8 ... 1e8 on the control
    variables: the FP32 path (with its FP64 fallback) still returns a
    solution with relative residual <= 1e-10."""
    P = W.poisson2d(N=64)
    rng = np.random.default_rng(3)
    diag = P.diag.copy()
    u = np.arange(1, P.n, 3)
    diag[u] = 10.0 ** rng.uniform(-8, 8, size=len(u))
    H = CSRMat(P.n, P.n, P.rowp, P.cols, P.vals)
    solver = MI355XSolver({"ldl_precision": "fp32"}, _State(H))
    d = Vector(P.n)
    d.get_array()[:] = diag
    solver.factor(H, d)
    b, x = Vector(P.n), Vector(P.n)
    b.get_array()[:] = P.rhs
    solver.solve(b, x)
    K = sp.csr_matrix((P.vals, P.cols, P.rowp), shape=(P.n, P.n)) + sp.diags(diag)
    res = np.linalg.norm(K @ x.get_array() - P.rhs) / np.linalg.norm(P.rhs)
    print("fallbacks", solver.fallbacks, "residual", res, solver.ldl.stats()["refine_steps"])
    assert res <= 1e-10


@pytest.mark.parametrize("prec", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
def test_refactor_equals_factor(prec):
    P = W.cartpole(N=400)
    mat = CSRMat(P.n, P.n, P.rowp, P.cols, P.vals)
    a = SparseLDL(mat, order=OrderingType.AMD, precision=prec)
    b = SparseLDL(mat, order=OrderingType.AMD, precision=prec)
    a.factor(P.diag)
    d2 = P.diag.copy()
    d2[P.primal_idx] += 1e-3
    assert a.refactor(d2) == 0
    assert b.factor(d2) == 0
    xa, xb = P.rhs.copy(), P.rhs.copy()
    a.solve(xa)
    b.solve(xb)
    assert np.array_equal(xa, xb)
    assert a.get_inertia() == b.get_inertia()


def test_set_pivot_threshold_equals_construction():
    case = next(c for c in golden_cases() if str(c["name"]) == "rand3_o1_u0.4")
    n = int(case["n"])
    mat = CSRMat(n, n, case["rowp"], case["cols"], case["vals"])
    s = SparseLDL(mat, ustab=0.01, order=engine_order(case))
    s.set_pivot_threshold(ustab=0.4)
    s.factor(None)
    info = s.node_info()
    assert np.array_equal(info["npiv"], case["ni_npiv"])
    below = case["ni_root"] == 0
    assert np.array_equal(info["ndel"][below], case["ni_ndel"][below])
    assert s.get_inertia() == tuple(int(v) for v in case["inertia"])
