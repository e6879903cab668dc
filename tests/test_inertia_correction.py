"""Inertia correction (Algorithm IC) driving the solvers.

The reference's InertiaCorrector (amigo/algorithm/solvers/
inertia_correction.py:11-416) sees only the inertia returned after each
perturbed factorization.  Its restatement (oracle/inertia_correction.py)
drives the GPU solver through MI355XSolver and the oracle on the same KKT;
parity means the same sequence of (delta_w, delta_c, npos, nneg) at every
attempt.  The closed-form cases pin the restated loop itself: a 3 x 3 KKT
with one negative curvature direction of -1/2 inside the constraint null
space needs delta_w > 1/2, which the delta_w ladder reaches at 1.0
(1e-4, x100, x100), then at 8/3 on the next system (1/3 of the last value,
then x8); a KKT whose two constraint rows coincide is singular and is fixed
by IncreaseQuality followed by delta_c = 1e-8 mu^(1/4)."""

import numpy as np
import pytest

from amigo_amd import workloads as W
from oracle import pyoracle as po
from oracle.inertia_correction import InertiaCorrection


def _csr(K):
    """CSR with the full diagonal present (explicit zeros kept)."""
    K = np.asarray(K, float)
    n = K.shape[0]
    mask = (K != 0) | np.eye(n, dtype=bool)
    r, c = np.nonzero(mask)
    rowp = np.searchsorted(r, np.arange(n + 1)).astype(np.int32)
    return n, rowp, c.astype(np.int32), K[r, c]


CURV = np.array([[1.0, 0.0, 1.0], [0.0, -0.5, 0.0], [1.0, 0.0, 0.0]])   # x0, x1 | c
SING = np.array([[1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [1.0, 0.0, 0.0]])    # x0 | c0, c1


def _oracle_fn(n, rowp, cols, vals):
    o = po.OracleLDL(n, rowp, cols, order=po.AMD)

    def f(d):
        o.factor(vals, d)
        return o.inertia()
    return f


def test_closed_form_negative_curvature():
    n, rowp, cols, vals = _csr(CURV)
    ic = InertiaCorrection([0, 1], [2])
    f = _oracle_fn(n, rowp, cols, vals)
    assert ic.factor(f, np.zeros(3), mu=0.1)
    assert [(dw, npos, nneg) for dw, _, npos, nneg in ic.log] == [
        (0.0, 1, 2), (1e-4, 1, 2), (1e-2, 1, 2), (1.0, 2, 1)]
    assert ic.factor(f, np.zeros(3), mu=0.1)        # next system: starts from 1/3
    assert [dw for dw, _, _, _ in ic.log[4:]] == [0.0, 1.0 / 3.0, 8.0 / 3.0]
    assert ic.log[-1][2:] == (2, 1)


def test_closed_form_singular_jacobian():
    n, rowp, cols, vals = _csr(SING)
    ic = InertiaCorrection([0], [1, 2])
    pivtols = []
    assert ic.factor(_oracle_fn(n, rowp, cols, vals), np.zeros(3), mu=0.1,
                     set_pivtol=pivtols.append)
    dc = 1e-8 * 0.1 ** 0.25
    assert ic.log == [(0.0, 0.0, 2, 1), (0.0, 0.0, 2, 1), (0.0, dc, 1, 2)]
    assert pivtols == [1e-6, 1e-3]                  # IncreaseQuality: sqrt(pivtol)


def test_rosenbrock_needs_correction_oracle():
    """C0 snapshot (n = 5,998): wrong inertia at delta_w = 0, corrected by the
    ladder; the accepted inertia is the expected one."""
    P = W.rosenbrock()
    ic = InertiaCorrection(P.primal_idx, P.dual_idx)
    assert ic.factor(_oracle_fn(P.n, P.rowp, P.cols, P.vals), P.diag, mu=1e-2)
    assert len(ic.log) >= 3
    assert ic.log[0][2:] != P.expected_inertia()
    assert ic.log[-1][2:] == P.expected_inertia()


# ------------------------------------------------------------------- GPU --

class _State:
    def __init__(self, hessian):
        self.hessian = hessian


def _gpu_fn(n, rowp, cols, vals, **opts):
    from amigo_amd.ldl import CSRMat, Vector
    from amigo_amd.solver import MI355XSolver
    H = CSRMat(n, n, rowp, cols, vals)
    solver = MI355XSolver(opts, _State(H))
    diag = Vector(n)

    def f(d):
        diag.get_array()[:] = d
        solver.factor(H, diag)
        return solver.get_inertia()
    return f, solver


@pytest.mark.gpu
@pytest.mark.parametrize("K,primal,dual", [(CURV, [0, 1], [2]), (SING, [0], [1, 2])],
                         ids=["curvature", "singular"])
def test_gpu_closed_form_cases(K, primal, dual):
    n, rowp, cols, vals = _csr(K)
    logs = []
    for fn in (_oracle_fn(n, rowp, cols, vals), _gpu_fn(n, rowp, cols, vals)[0]):
        ic = InertiaCorrection(primal, dual)
        assert ic.factor(fn, np.zeros(n), mu=0.1)
        assert ic.factor(fn, np.zeros(n), mu=0.1)
        logs.append(ic.log)
    assert logs[0] == logs[1]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gpu_rosenbrock_same_attempts_as_oracle(prec):
    """Two consecutive IPM systems (the second with a rescaled barrier
    diagonal): every attempt's inertia on the GPU equals the oracle's."""
    P = W.rosenbrock()
    fg, solver = _gpu_fn(P.n, P.rowp, P.cols, P.vals, ldl_precision=prec)
    fo = _oracle_fn(P.n, P.rowp, P.cols, P.vals)
    logs = []
    for fn in (fo, fg):
        ic = InertiaCorrection(P.primal_idx, P.dual_idx)
        assert ic.factor(fn, P.diag, mu=1e-2)
        assert ic.factor(fn, P.diag * 0.5, mu=5e-3)
        logs.append(ic.log)
    assert logs[0] == logs[1]
    assert logs[1][-1][2:] == P.expected_inertia()
    assert len(logs[1]) >= 4


@pytest.mark.gpu
def test_gpu_solver_accepts_reference_vector_calls():
    """The calls the reference's corrector makes on its diagonal Vector
    (copy, add_scalar_at) before MI355XSolver.factor."""
    from amigo_amd.ldl import Vector
    n, rowp, cols, vals = _csr(CURV)
    _, solver = _gpu_fn(n, rowp, cols, vals)
    d = Vector(3)
    p = Vector(3)
    p.copy(d)
    p.add_scalar_at(np.array([0, 1], np.int32), 1e-12 + 1.0)
    solver.factor(solver.hessian, p)
    assert solver.get_inertia() == (2, 1)
    b = Vector(3)
    b.get_array()[:] = [1.0, 2.0, 3.0]
    x = Vector(3)
    solver.solve(b, x)
    Kp = CURV + np.diag([1.0 + 1e-12, 1.0 + 1e-12, 0.0])
    np.testing.assert_allclose(Kp @ x.get_array(), b.get_array(), atol=1e-12)
