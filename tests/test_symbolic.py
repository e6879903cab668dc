"""The engine's host-side symbolic analysis (amigo_amd/csrc/symbolic.cpp,
reached through kkt_ldl_create, which needs no GPU) against the reference's
golden fixtures: permutation, supernode partition, tree and row patterns
bit-exact for NATURAL and AMD orderings."""

import numpy as np
import pytest

from amigo_amd import workloads as W
from amigo_amd.ldl import CSRMat, OrderingType, SolverType, SparseLDL
from oracle.pyoracle import GIVEN, OracleLDL

from .conftest import engine_order, golden_cases

CASES = golden_cases()


@pytest.mark.parametrize("c", CASES, ids=[str(c["name"]) for c in CASES])
def test_engine_symbolic_bit_exact(c):
    n = int(c["n"])
    s = SparseLDL(CSRMat(n, n, c["rowp"], c["cols"], c["vals"]),
                  solver_type=int(c["solver_type"]), ustab=float(c["ustab"]),
                  order=engine_order(c))
    perm, iperm = s.get_perm()
    assert np.array_equal(perm, c["perm"])
    assert np.array_equal(iperm, c["iperm"])
    nodes = s.get_supernodes()
    for k in ("node_ptr", "node_to_var", "children_ptr", "children", "contrib_ptr",
              "contrib_rows"):
        assert np.array_equal(nodes[k], c[k]), k


@pytest.mark.parametrize("name,kw", [("poisson2d", dict(N=48)), ("cartpole", dict(N=300)),
                                     ("elasticity3d", dict(N=5)), ("rosenbrock", {})])
def test_engine_symbolic_matches_oracle_on_workloads(name, kw):
    P = W.GENERATORS[name](**kw)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD)
    o = OracleLDL(P.n, P.rowp, P.cols)
    assert np.array_equal(s.get_perm()[0], o.perm()[0])
    a, b = s.get_supernodes(), o.nodes()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    st = s.stats()
    assert st["num_nodes"] == len(b["node_ptr"]) - 1
    par, post, cc = o.etree()
    c = cc.astype(float) - 1
    assert st["flops_factor"] == pytest.approx(float(np.sum(c * (c + 2))))
    assert st["nnz_l"] == int(cc.sum())


@pytest.mark.parametrize("name,kw", [("poisson2d", dict(N=40)), ("elasticity3d", dict(N=4)),
                                     ("cartpole", dict(N=150))])
def test_nested_dissection_is_a_valid_ordering(name, kw):
    """ND (this engine's own; METIS is absent) returns a permutation, and the
    analysis under it equals the oracle's analysis under the same
    permutation."""
    P = W.GENERATORS[name](**kw)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                  order=OrderingType.NESTED_DISSECTION)
    perm, iperm = s.get_perm()
    assert sorted(perm.tolist()) == list(range(P.n))
    assert np.array_equal(iperm[perm], np.arange(P.n))
    o = OracleLDL(P.n, P.rowp, P.cols, order=GIVEN, perm=perm)
    a, b = s.get_supernodes(), o.nodes()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_given_ordering_and_cholesky_flags():
    P = W.poisson2d(N=16)
    s_amd = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD)
    perm = s_amd.get_perm()[0]
    s_g = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), perm=perm)
    assert np.array_equal(s_g.get_perm()[0], perm)
    for k, v in s_amd.get_supernodes().items():
        assert np.array_equal(v, s_g.get_supernodes()[k])
    s_c = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), SolverType.CHOLESKY,
                    order=OrderingType.AMD)
    assert np.array_equal(s_c.get_perm()[0], perm)
