"""Speculative large-front panels against the sequential panel.

Fronts above the LDS-resident size are factored by panels that choose their
pivots from the leading rows and then verify the choices against every row
(falling back to the per-pivot search at the first contradicted choice).
With the speculation switched off (KKT_LDL_PANEL_SPECULATION=0, read when
the handle is created) every panel runs the per-pivot search.  The two must
give the same factorization bit for bit: the same pivots, delays and 2x2
blocks per node, the same inertia and the same solution bits.

Cases: the Poisson control KKT (regular 1x1 / adjacent 2x2 pattern), 3D
elasticity, and random indefinite matrices with zero diagonals, whose fronts
need non-adjacent 2x2 pivots, 1x1 pivots away from the diagonal and delays
(the verification and fall-back paths), in FP64 and FP32.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from amigo_amd import workloads as W
from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SparseLDL

pytestmark = pytest.mark.gpu


def _random_indefinite(n, density, zero_frac, seed):
    rng = np.random.default_rng(seed)
    U = sp.triu(sp.random(n, n, density=density, random_state=rng), k=1).tocoo()
    d = rng.standard_normal(n)
    d[rng.random(n) < zero_frac] = 0.0
    r = np.concatenate([U.row, U.col, np.arange(n)])
    c = np.concatenate([U.col, U.row, np.arange(n)])
    v = np.concatenate([U.data, U.data, d])
    A = sp.csr_matrix((v, (r, c)), shape=(n, n))  # explicit zeros on the diagonal kept
    A.sort_indices()
    return (n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy(),
            np.zeros(n), rng.standard_normal(n))


def _forest():
    """Two Poisson control KKTs of different sizes as one block-diagonal
    matrix: two roots, the smaller one on a level that also holds large
    non-root fronts of the larger component."""
    parts = [W.poisson2d(N=96), W.poisson2d(N=40, seed=1)]
    K = sp.block_diag([sp.csr_matrix((P.vals, P.cols, P.rowp), shape=(P.n, P.n)) for P in parts],
                      format="csr")
    K.sort_indices()
    diag = np.concatenate([P.diag for P in parts])
    rhs = np.concatenate([P.rhs for P in parts])
    return (K.shape[0], K.indptr.astype(np.int32), K.indices.astype(np.int32), K.data.copy(),
            diag, rhs, OrderingType.NESTED_DISSECTION), parts


def _case(name):
    if name == "forest":
        return _forest()[0]
    if name == "poisson2d":
        P = W.poisson2d(N=96)
        return P.n, P.rowp, P.cols, P.vals, P.diag, P.rhs, OrderingType.NESTED_DISSECTION
    if name == "elasticity3d":
        P = W.elasticity3d(N=8)
        return P.n, P.rowp, P.cols, P.vals, P.diag, P.rhs, OrderingType.NESTED_DISSECTION
    seed = int(name.split("_")[1])
    return _random_indefinite(700, 0.02, 0.3, seed) + (OrderingType.AMD,)


def _run(monkeypatch, case, precision, spec):
    n, rowp, cols, vals, diag, rhs, order = case
    monkeypatch.setenv("KKT_LDL_PANEL_SPECULATION", spec)
    s = SparseLDL(CSRMat(n, n, rowp, cols, vals), order=order, precision=precision)
    s.factor(diag)
    x = rhs.copy()
    s.solve(x)
    return s.get_inertia(), s.node_info(), x


@pytest.mark.parametrize("precision", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", ["poisson2d", "elasticity3d", "forest", "random_1", "random_2",
                                  "random_3"])
def test_speculative_panel_equals_sequential(monkeypatch, name, precision):
    print("case", name, precision, flush=True)
    case = _case(name)
    ia, na, xa = _run(monkeypatch, case, precision, "1")
    ib, nb, xb = _run(monkeypatch, case, precision, "0")
    assert ia == ib
    for k in na:
        assert np.array_equal(na[k], nb[k]), k
    assert np.array_equal(xa, xb)
    assert sum(ia) == case[0]
    if name.startswith("random"):
        return
    # the solution is a solution
    n, rowp, cols, vals, diag, rhs, _ = case
    K = sp.csr_matrix((vals, cols, rowp), shape=(n, n)) + sp.diags(diag)
    r = np.linalg.norm(K @ xa - rhs) / np.linalg.norm(rhs)
    assert r < (1e-8 if precision == FP64 else 1e-2)


@pytest.mark.parametrize("precision", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
def test_forest_inertia_is_the_sum_of_its_components(monkeypatch, precision):
    """A level holding a root front (Bunch-Kaufman tests) next to large
    non-root fronts: the inertia of the block-diagonal matrix is the sum of
    its blocks' inertias, each factored on its own."""
    monkeypatch.setenv("KKT_LDL_PANEL_SPECULATION", "1")
    case, parts = _forest()
    ia, _, x = _run(monkeypatch, case, precision, "1")
    ib = [0, 0]
    for P in parts:
        s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                      order=OrderingType.NESTED_DISSECTION, precision=precision)
        s.factor(P.diag)
        pos, neg = s.get_inertia()
        ib[0] += pos
        ib[1] += neg
    assert tuple(ia) == tuple(ib)
    n, rowp, cols, vals, diag, rhs, _ = case
    K = sp.csr_matrix((vals, cols, rowp), shape=(n, n)) + sp.diags(diag)
    assert np.linalg.norm(K @ x - rhs) / np.linalg.norm(rhs) < (1e-8 if precision == FP64 else 1e-2)
