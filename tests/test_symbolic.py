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


def _grid(nx, nine):
    """Pattern of a 5- or 9-point grid operator (diagonal included)."""
    import scipy.sparse as sp
    idx = np.arange(nx * nx).reshape(nx, nx)
    offs = [(1, 0), (0, 1)] + ([(1, 1), (1, -1)] if nine else [])
    r, c = [], []
    for dx, dy in offs:
        a = idx[max(0, -dx):nx - max(0, dx), max(0, -dy):nx - max(0, dy)]
        b = idx[max(0, dx):nx + min(0, dx), max(0, dy):nx + min(0, dy)]
        r.append(a.ravel())
        c.append(b.ravel())
    r, c = np.concatenate(r), np.concatenate(c)
    n = nx * nx
    A = sp.coo_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    A = (A + A.T + sp.eye(n)).tocsr()
    A.sort_indices()
    return n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data


@pytest.mark.parametrize("nine", [False, True], ids=["5-point", "9-point"])
def test_nd_finds_grid_line_separators(nine):
    """On an N x N grid the best top separator is one grid line (N
    vertices); the root supernode of the nested dissection holds the top
    separator (and possibly a merged chain below it), and the factor's
    critical path follows the nested separators: N + 2 (N/2) + 2 (N/4) + ...
    = 3 N for straight lines, plus the leaf chains."""
    N = 96
    n, rowp, cols, vals = _grid(N, nine)
    s = SparseLDL(CSRMat(n, n, rowp, cols, vals), order=OrderingType.NESTED_DISSECTION)
    sn = s.get_supernodes()
    ns = np.diff(sn["node_ptr"])
    st = s.stats()
    root = ns[-1] if len(ns) else 0
    assert N <= root <= 2.2 * N
    assert st["crit_path_pivots"] <= 5.0 * N
    # a much better fill than AMD's on a 2D grid is not required, but the
    # nested dissection must not be worse than the natural band order
    band = SparseLDL(CSRMat(n, n, rowp, cols, vals), order=OrderingType.NATURAL).stats()
    assert st["flops_factor"] < 0.5 * band["flops_factor"]


def test_nd_is_deterministic():
    P = W.poisson2d(N=48)
    perms = [SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                       order=OrderingType.NESTED_DISSECTION).get_perm()[0] for _ in range(3)]
    assert all(np.array_equal(perms[0], q) for q in perms[1:])


def test_pairing_makes_2x2_partners_adjacent():
    """With the matrix values (vals_hint), predicted 2x2 partners are placed
    next to each other inside fundamental supernodes: the oracle, factoring
    in that order, meets most 2x2 pivots of large fronts as a pivot and its
    next candidate; without the values (index order inside separators) most
    of them are distant."""
    import ctypes

    from oracle import pyoracle as po
    lib = po._load(po.ORACLE_SO, "orc_")
    lib.orc_pivot_kinds.argtypes = [ctypes.POINTER(ctypes.c_longlong)]
    P = W.poisson2d(N=96)

    def kinds(perm):
        o = OracleLDL(P.n, P.rowp, P.cols, order=GIVEN, perm=perm)
        o.factor(P.vals, P.diag)
        out = (ctypes.c_longlong * 5)()
        lib.orc_pivot_kinds(out)
        return list(out)   # 1x1 at k, 2x2 with k+1, 1x1 later, 2x2 later, delay

    paired = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                       order=OrderingType.NESTED_DISSECTION)
    mat = CSRMat(P.n, P.n, P.rowp, P.cols, P.vals)
    plain = SparseLDL(mat, order=OrderingType.NESTED_DISSECTION)
    # the same ordering without the values: build it through the C ABI
    # with vals_hint NULL (the Python mirror always passes the values)
    import amigo_amd.ldl as L
    o = L.Options()
    plain.lib.kkt_ldl_default_options(ctypes.byref(o))
    o.order = OrderingType.NESTED_DISSECTION
    h = ctypes.c_void_p()
    assert plain.lib.kkt_ldl_create(P.n, L._ip(P.rowp), L._ip(P.cols), ctypes.byref(o),
                                    ctypes.byref(h)) == 0
    p0, ip0 = np.zeros(P.n, np.int32), np.zeros(P.n, np.int32)
    plain.lib.kkt_ldl_get_perm(h, L._ip(p0), L._ip(ip0))
    plain.lib.kkt_ldl_destroy(h)
    kp, k0 = kinds(paired.get_perm()[0]), kinds(p0)
    assert kp[1] > 4 * kp[3]          # adjacent 2x2 dominate with the values
    assert k0[3] > k0[1]              # distant 2x2 dominate without them
