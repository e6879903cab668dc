"""The seeded KKT generators: sizes of the benchmark configurations,
symmetric patterns with explicit diagonals, and the KKT block layout."""

import numpy as np
import pytest
import scipy.sparse as sp

from amigo_amd import workloads as W


@pytest.mark.parametrize("name,kw,n", [
    ("rosenbrock", {}, 5998),
    ("poisson2d", dict(N=16), 3 * 17 ** 2),
    ("cartpole", dict(N=50), 9 * 50 - 4),
    ("elasticity3d", dict(N=3), 9 * 4 ** 3),
])
def test_shapes_and_pattern(name, kw, n):
    P = W.GENERATORS[name](**kw)
    assert P.n == n
    A = sp.csr_matrix((P.vals, P.cols, P.rowp), shape=(n, n))
    assert abs(A - A.T).max() < 1e-12
    Pat = sp.csr_matrix((np.ones_like(P.vals), P.cols, P.rowp), shape=(n, n))
    assert (Pat != Pat.T).nnz == 0
    for i in range(n):
        row = P.cols[P.rowp[i]:P.rowp[i + 1]]
        assert np.all(np.diff(row) > 0)
        assert i in row
    assert len(P.primal_idx) + len(P.dual_idx) <= n
    assert np.all(P.diag[P.dual_idx] <= 0)


def test_rosenbrock_hessian_is_exact():
    """KKT (x, x) block equals the Hessian of the Lagrangian by central
    differences of its analytic gradient."""
    n = 12
    P = W.rosenbrock(n=n)
    rng = np.random.default_rng(0)
    x = rng.uniform(-2, 2, size=n)
    lam = rng.uniform(0, 1, size=n - 1)
    # regenerate with the same draws as the generator
    Q = W.rosenbrock(n=n)
    A = sp.csr_matrix((Q.vals, Q.cols, Q.rowp), shape=(Q.n, Q.n)).toarray()
    ix = 3 * np.arange(n)
    il = 3 * np.arange(n - 1) + 2

    def grad(x, lam):
        g = np.zeros(n)
        g[:-1] += -2 * (1 - x[:-1]) - 400 * x[:-1] * (x[1:] - x[:-1] ** 2)
        g[1:] += 200 * (x[1:] - x[:-1] ** 2)
        g[:-1] += 2 * lam * x[:-1]
        g[1:] += 2 * lam * x[1:]
        return g

    rng = np.random.default_rng(0)
    x0 = rng.uniform(-2.0, 2.0, size=n)
    l0 = rng.uniform(0.0, 1.0, size=n - 1)
    h = 1e-6
    H = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = h
        H[:, j] = (grad(x0 + e, l0) - grad(x0 - e, l0)) / (2 * h)
    assert np.allclose(A[np.ix_(ix, ix)], H, atol=1e-5)
    J = A[np.ix_(il, ix)]
    Jref = np.zeros((n - 1, n))
    Jref[np.arange(n - 1), np.arange(n - 1)] = 2 * x0[:-1]
    Jref[np.arange(n - 1), np.arange(1, n)] = 2 * x0[1:]
    assert np.allclose(J, Jref)


def test_poisson_blocks():
    P = W.poisson2d(N=8, alpha=1e-4)
    A = sp.csr_matrix((P.vals, P.cols, P.rowp), shape=(P.n, P.n)).toarray()
    nn = 81
    Y, U, Lm = 3 * np.arange(nn), 3 * np.arange(nn) + 1, 3 * np.arange(nn) + 2
    free = np.setdiff1d(np.arange(nn), P.info["n_fixed"] and
                        np.nonzero((np.arange(nn) // 9 == 0) | (np.arange(nn) % 9 == 0) |
                                   (np.arange(nn) // 9 == 8) | (np.arange(nn) % 9 == 8))[0])
    Muu = A[np.ix_(U, U)]
    assert np.allclose(Muu, Muu.T)
    assert np.all(np.linalg.eigvalsh(Muu) > 0)
    # mass matrix integrates 1 over the unit square: alpha * sum = alpha
    assert np.isclose(Muu.sum(), 1e-4)
    KLy = A[np.ix_(Lm[free], Y[free])]
    assert np.allclose(KLy, KLy.T)
