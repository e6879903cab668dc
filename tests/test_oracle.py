"""The oracle (CPU restatement of the reference's SparseLDL) pinned against
the reference's own outputs, known answers and brute force.

* golden fixtures written by the compiled reference (tests/golden): ordering,
  supernodes and row patterns bit-exact, inertia exact, per-node pivot
  outcome exact below the root, solutions to rounding;
* the reference's tests/unit/test_ldl.py assertions (inertia from eigh,
  multi-RHS error < 1e-11 against x = ones);
* elimination tree and column counts against dense symbolic elimination;
* dense factorizations against numpy eigh / solve and scipy.linalg.ldl.
"""

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp

from oracle.pyoracle import AMD, CHOLESKY, GIVEN, LDL, NATURAL, OracleLDL

from .conftest import golden_cases

CASES = golden_cases()
IDS = [str(c["name"]) for c in CASES]


def _oracle_for(c):
    return OracleLDL(int(c["n"]), c["rowp"], c["cols"], order=int(c["order"]),
                     ustab=float(c["ustab"]), pivot_tol=float(c["pivot_tol"]),
                     solver_type=int(c["solver_type"]))


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_symbolic_bit_exact_vs_reference(c):
    o = _oracle_for(c)
    perm, iperm = o.perm()
    assert np.array_equal(perm, c["perm"])
    assert np.array_equal(iperm, c["iperm"])
    nodes = o.nodes()
    for k in ("node_ptr", "node_to_var", "children_ptr", "children", "contrib_ptr",
              "contrib_rows"):
        assert np.array_equal(nodes[k], c[k]), k


@pytest.mark.parametrize("c", CASES, ids=IDS)
def test_numeric_vs_reference(c):
    o = _oracle_for(c)
    diag = c["diag"] if bool(c["has_diag"]) else None
    o.factor(c["vals"], diag)
    if int(c["solver_type"]) == LDL:
        assert o.inertia() == tuple(int(v) for v in c["inertia"])
    info = o.node_info()
    nonroot = c["ni_root"] == 0
    # Below the root the pivot sequence is the reference's; at the root it is
    # LAPACK dsytrf's, whose blocked variant is vendor specific.
    for k in ("npiv", "ndel", "n2x2"):
        assert np.array_equal(info[k][nonroot], c["ni_" + k][nonroot]), k
    assert np.array_equal(info["npiv"], c["ni_npiv"])
    x = o.solve(c["rhs"])
    n = int(c["n"])
    K = sp.csr_matrix((c["vals"], c["cols"], c["rowp"]), shape=(n, n))
    if diag is not None:
        K = K + sp.diags(diag)
    b = c["rhs"].reshape(n, -1)
    r = np.linalg.norm(K @ x.reshape(n, -1) - b) / np.linalg.norm(b)
    rr = np.linalg.norm(K @ c["x"].reshape(n, -1) - b) / np.linalg.norm(b)
    assert r <= max(1e-10, 10 * rr)
    scale = np.abs(c["x"]).max()
    assert np.abs(x - c["x"]).max() <= 1e-6 * scale


# --- the reference's own test_ldl.py assertions, on the oracle ------------

HAND = [c for c in CASES if str(c["name"]).startswith("hand")]


@pytest.mark.parametrize("c", HAND, ids=[str(c["name"]) for c in HAND])
def test_hand_matrix_inertia_matches_eigh(c):
    n = int(c["n"])
    A = sp.csr_matrix((c["vals"], c["cols"], c["rowp"]), shape=(n, n)).toarray()
    ev = np.linalg.eigvalsh(A)
    o = _oracle_for(c)
    o.factor(c["vals"])
    assert o.inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
    x = o.solve(c["rhs"])
    assert np.linalg.norm(x - c["exact"]) < 1e-12


POIS = [c for c in CASES if str(c["name"]).startswith("poisson")]


@pytest.mark.parametrize("c", POIS, ids=[str(c["name"]) for c in POIS])
def test_poisson_multi_rhs(c):
    o = _oracle_for(c)
    o.factor(c["vals"])
    if int(c["solver_type"]) == LDL:
        assert o.inertia() == (int(c["n"]), 0)
    X = o.solve(c["rhs"])
    assert X.shape[1] == 7
    err = max(np.linalg.norm(X[:, k] - c["exact"]) for k in range(7))
    assert err < 1e-11


# --- textbook quantities against brute force -------------------------------

def _random_pattern(rng, n, density):
    A = sp.random(n, n, density=density, random_state=int(rng.integers(1 << 30)))
    A = (A + A.T + sp.eye(n)).tocsr()
    A.sort_indices()
    return A


def _dense_symbolic(P):
    """Boolean dense elimination: parent and column counts of L."""
    n = P.shape[0]
    S = P.copy().astype(bool)
    parent = -np.ones(n, int)
    cc = np.zeros(n, int)
    for j in range(n):
        below = np.nonzero(S[j + 1:, j])[0] + j + 1
        cc[j] = 1 + len(below)
        if len(below):
            parent[j] = below[0]
            S[np.ix_(below, below)] = True
    return parent, cc


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("order", [NATURAL, AMD])
def test_etree_and_colcounts_bruteforce(seed, order):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 14))
    A = _random_pattern(rng, n, 0.25)
    o = OracleLDL(n, A.indptr, A.indices, order=order)
    perm, _ = o.perm()
    par, post, cc = o.etree()
    D = A.toarray() != 0
    Dp = D[np.ix_(perm, perm)]
    bpar, bcc = _dense_symbolic(Dp)
    assert np.array_equal(par, bpar)
    assert np.array_equal(cc, bcc)
    # post-order: every node after its descendants, a permutation
    pos = np.empty(n, int)
    pos[post] = np.arange(n)
    assert sorted(post) == list(range(n))
    for v in range(n):
        if par[v] >= 0:
            assert pos[v] < pos[par[v]]


@pytest.mark.parametrize("seed", range(10))
def test_dense_indefinite_against_numpy(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(2, 40))
    M = rng.standard_normal((n, n))
    M = M + M.T
    M[rng.random((n, n)) < 0.3] = 0.0
    M = (M + M.T) / 2 + np.diag(rng.standard_normal(n) * (rng.random(n) < 0.5))
    A = sp.csr_matrix(M + 0 * np.eye(n))
    A = (A + sp.eye(n) * 0.0).tocsr()
    A.setdiag(M.diagonal())
    A.sort_indices()
    ev = np.linalg.eigvalsh(A.toarray())
    if np.min(np.abs(ev)) < 1e-8:
        pytest.skip("near-singular draw")
    for order in (NATURAL, AMD):
        for ustab in (0.01, 0.4):
            o = OracleLDL(n, A.indptr, A.indices, order=order, ustab=ustab)
            o.factor(A.data)
            assert o.inertia() == (int(np.sum(ev >= 0)), int(np.sum(ev < 0)))
            b = rng.standard_normal(n)
            x = o.solve(b)
            assert np.allclose(x, np.linalg.solve(A.toarray(), b), rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("seed", range(6))
def test_root_bunch_kaufman_against_scipy_ldl(seed):
    """A fully dense pattern is one root front: the restated dsytf2/dsytrs."""
    rng = np.random.default_rng(200 + seed)
    n = int(rng.integers(2, 60))
    M = rng.standard_normal((n, n))
    M = M + M.T
    A = sp.csr_matrix(np.ones((n, n)))
    A.data = M[A.nonzero()].copy()
    A.sort_indices()
    o = OracleLDL(n, A.indptr, A.indices, order=NATURAL)
    o.factor(A.data)
    lu, d, _ = scipy.linalg.ldl(M, lower=True)
    evd = np.linalg.eigvalsh(d)
    assert o.inertia() == (int(np.sum(evd >= 0)), int(np.sum(evd < 0)))
    b = rng.standard_normal(n)
    assert np.allclose(o.solve(b), np.linalg.solve(M, b), rtol=1e-7, atol=1e-9)


def test_given_perm_reproduces_amd():
    """ORDER_GIVEN with the AMD permutation gives the AMD analysis."""
    c = [c for c in CASES if str(c["name"]) == "kkt_rosenbrock"][0]
    n = int(c["n"])
    a = OracleLDL(n, c["rowp"], c["cols"], order=AMD)
    g = OracleLDL(n, c["rowp"], c["cols"], order=GIVEN, perm=c["perm"])
    for k, v in a.nodes().items():
        assert np.array_equal(v, g.nodes()[k])


def test_planted_ordering_mistake_is_caught():
    """A transposed pair in the permutation changes the supernode tree that
    the bit-exact check compares, so a plausible ordering slip fails it."""
    c = [c for c in CASES if str(c["name"]) == "kkt_rosenbrock"][0]
    n = int(c["n"])
    perm = c["perm"].copy()
    perm[[10, 11]] = perm[[11, 10]]
    g = OracleLDL(n, c["rowp"], c["cols"], order=GIVEN, perm=perm)
    nodes = g.nodes()
    same = all(np.array_equal(nodes[k], c[k]) for k in nodes)
    assert not same
