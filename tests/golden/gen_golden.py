"""Write the golden fixtures of the reference's SparseLDL path.

Runs the reference's own SparseLDL, compiled from its headers by
`make -C oracle ref` (oracle/_ref/libamigo_ref.so, built only where the
reference tree is present), on:

* the 7 hand matrices and the seeded, permuted 2D Poisson grids of the
  reference's tests/unit/test_ldl.py:20-118 (ustab = 0.4, NATURAL and AMD;
  CHOLESKY + AMD for the Poisson grids), x = ones as exact solution;
* seeded random symmetric indefinite matrices (NATURAL, AMD; ustab 0.4, 0.01);
* the KKT snapshots of amigo_amd.workloads at test sizes (AMD, ustab 0.01,
  pivot_tol 1e-14, delay_growth 2.0 as AmigoSolver uses them).

For each case it records the inputs, the permutation, the supernode
partition and tree, the row pattern below each node, the inertia, the per
node pivot outcome (pivots, delayed, 2x2 blocks, root flag) and the solution
for a seeded right-hand side.  Output: tests/golden/*.npz.

    python tests/golden/gen_golden.py
"""

import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from amigo_amd import workloads as W  # noqa: E402
from oracle.pyoracle import AMD, CHOLESKY, LDL, NATURAL, RefLDL  # noqa: E402

HAND = [
    [[2, 1, 0, 0, 0], [1, 2, 1, 0, 0], [0, 1, 2, 1, 0], [0, 0, 1, 2, 1], [0, 0, 0, 1, 2]],
    [[0, 1, 0, 0, 0], [1, 0, 1, 0, 0], [0, 1, 2, 1, 0], [0, 0, 1, 2, 1], [0, 0, 0, 1, 2]],
    [[0, 0, 0, 1, 1], [0, 0, 1, 0, 1], [0, 1, 2, 0, 1], [1, 0, 0, 0, 1], [1, 1, 1, 1, 0]],
    [[0, 1, 0, 0, 1], [1, 0, 2, 0, 1], [0, 2, 0, 0, 1], [0, 0, 0, 2, 1], [1, 1, 1, 1, 0]],
    [[0, 1], [1, 0]],
    [[1, 1], [1, 0]],
    [[1, 2, 3], [2, 1, 3], [3, 3, 0]],
]


def poisson_2d(n):
    """tests/unit/test_ldl.py:85-118 under np.random.seed(0)."""
    m = n - 1
    T = sp.diags([-np.ones(m - 1), 2.0 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1],
                 format="csr")
    I = sp.eye(m, format="csr")
    A = ((sp.kron(I, T) + sp.kron(T, I)) * float(n * n)).tocsr()
    perm = np.random.permutation(m * m)
    return A[perm, :][:, perm].tocsr()


def random_indefinite(seed, n):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=min(1.0, 4.0 / n), random_state=int(rng.integers(1 << 30)))
    A = (A + A.T).tocsr()
    d = rng.normal(size=n) * (rng.random(n) < 0.5)
    A = (A + sp.diags(d) + sp.eye(n) * 0.0).tocsr()
    A.setdiag(A.diagonal())  # explicit diagonal in every row
    A.sort_indices()
    return A


def record(name, A, order, ustab, solver_type=LDL, diag=None, rhs=None, nrhs=1,
           pivot_tol=1e-14, exact=None):
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    r = RefLDL(n, A.indptr, A.indices, order=order, ustab=ustab, pivot_tol=pivot_tol,
               solver_type=solver_type)
    perm, iperm = r.perm()
    nodes = r.nodes()
    flag = r.factor(A.data, diag)
    if rhs is None:
        rhs = np.random.default_rng(7).standard_normal((n, nrhs))
        if nrhs == 1:
            rhs = rhs[:, 0]
    x = r.solve(rhs)
    inertia = r.inertia() if solver_type == LDL else (0, 0)
    info = r.node_info()
    out = dict(
        name=name, n=n, order=order, ustab=ustab, pivot_tol=pivot_tol,
        solver_type=solver_type, rowp=A.indptr.astype(np.int32),
        cols=A.indices.astype(np.int32), vals=A.data.astype(np.float64),
        diag=np.zeros(0) if diag is None else np.asarray(diag, np.float64),
        has_diag=diag is not None, rhs=np.asarray(rhs, np.float64),
        x=np.asarray(x, np.float64), perm=perm, iperm=iperm,
        inertia=np.array(inertia, np.int32), flag=flag,
        exact=np.zeros(0) if exact is None else exact, **nodes,
        **{"ni_" + k: v for k, v in info.items()},
    )
    return out


def main():
    cases = []
    for i, e in enumerate(HAND):
        M = np.array(e, float)
        for order in (NATURAL, AMD):
            xs = np.ones(M.shape[0])
            cases.append(record(f"hand{i}_o{order}", M, order, 0.4, rhs=M @ xs, exact=xs))
    for nn in (10, 20, 50):
        np.random.seed(0)
        A = poisson_2d(nn)
        xs = np.ones(A.shape[0])
        b = np.asfortranarray(np.repeat((A @ xs)[:, None], 7, axis=1))
        for order in (NATURAL, AMD):
            cases.append(record(f"poisson{nn}_o{order}", A, order, 0.4, rhs=b, exact=xs))
        cases.append(record(f"poisson{nn}_chol", A, AMD, 0.4, solver_type=CHOLESKY,
                            rhs=b, exact=xs))
    for seed, n in ((1, 40), (2, 120), (3, 300)):
        A = random_indefinite(seed, n)
        for order in (NATURAL, AMD):
            for us in (0.4, 0.01):
                cases.append(record(f"rand{seed}_o{order}_u{us}", A, order, us))
    for name, kw in (("rosenbrock", {}), ("poisson2d", dict(N=32)),
                     ("cartpole", dict(N=200)), ("elasticity3d", dict(N=4))):
        P = W.GENERATORS[name](**kw)
        K = sp.csr_matrix((P.vals, P.cols, P.rowp), shape=(P.n, P.n))
        cases.append(record(f"kkt_{name}", K, AMD, 0.01, diag=P.diag, rhs=P.rhs))
    for c in cases:
        np.savez_compressed(os.path.join(HERE, c["name"] + ".npz"), **c)
    print(f"wrote {len(cases)} fixtures to {HERE}")


if __name__ == "__main__":
    main()
