"""Seeded KKT snapshots for the interior-point linear-algebra workloads.

Each generator returns a :class:`KKTSnapshot`: the KKT Hessian of the
Lagrangian over all variables (primal x, slacks s and multipliers lambda) as a
full symmetric CSR with sorted columns and an explicit diagonal in every row,
the barrier diagonal ``diag`` that the interior-point method hands to
``factor()`` (Sigma + delta_w on primal rows, -delta_c on multiplier rows) and a
seeded right-hand side.  This is the layout the reference's IPM produces:
Sigma is not stored in the CSR but added at factorization time
(amigo/algorithm/solvers/inertia_correction.py:280-294,
include/interior_point_backend.h:183-200), fixed variables keep identity rows
(include/fixed_variables.h:66-69) and inequalities get a slack with
``c(x) - s = 0`` (amigo/model.py:1768-1853).

Workloads (sizes as in the benchmark configurations):

* ``rosenbrock``  -- chained Rosenbrock, n = 2,000, n-1 constraints
  x_i^2 + x_{i+1}^2 <= 1.5, bounds [-2, 2]; KKT n = 5,998.
* ``cartpole``    -- trapezoidal direct collocation of the cart-pole
  (examples/trajectory/cart/cart_pole.py:40-82 dynamics), N = 20,000 nodes,
  4 states + 1 control + 4 defect multipliers per node; KKT n = 9N - 4.
* ``poisson2d``   -- Q1 Poisson distributed control on a 256 x 256 mesh,
  0.5||y - y_d||_M^2 + alpha/2 ||u||_M^2, alpha = 1e-4, -1 <= u <= 1,
  y = 0 on the boundary; KKT n = 3 * 257^2 = 198,147.
* ``elasticity3d`` -- Q1 linear elasticity distributed control on a 40^3 hex
  mesh, E ~ lognormal(0, 0.5) per element; KKT n = 3 * 3 * 41^3 = 620,289.

The per-element Hessians are derived by hand (Rosenbrock, Q1 mass/stiffness)
or with torch autograd on the CPU (cart-pole dynamics).
"""

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp


@dataclass
class KKTSnapshot:
    name: str
    n: int
    rowp: np.ndarray      # int32[n+1]
    cols: np.ndarray      # int32[nnz]
    vals: np.ndarray      # float64[nnz]
    diag: np.ndarray      # float64[n], added to the diagonal at factor time
    rhs: np.ndarray       # float64[n]
    primal_idx: np.ndarray
    dual_idx: np.ndarray
    info: dict = field(default_factory=dict)

    @property
    def nnz(self):
        return int(self.rowp[-1])

    def matrix(self, with_diag=True):
        """The assembled KKT operator as scipy CSR (Sigma included)."""
        A = sp.csr_matrix((self.vals, self.cols, self.rowp), shape=(self.n, self.n))
        if with_diag:
            A = A + sp.diags(self.diag)
        return A.tocsr()

    def expected_inertia(self):
        """(n_primal, n_dual): what the inertia correction accepts
        (amigo/algorithm/solvers/inertia_correction.py:269-278)."""
        return len(self.primal_idx), len(self.dual_idx)


def _csr_from_coo(n, r, c, v):
    """Symmetric CSR with sorted columns and an explicit diagonal per row."""
    d = np.arange(n)
    r = np.concatenate([r, d])
    c = np.concatenate([c, d])
    v = np.concatenate([v, np.zeros(n)])
    A = sp.coo_matrix((v, (r, c)), shape=(n, n)).tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32),
            A.data.astype(np.float64))


def _fix_rows(n, rowp, cols, vals, fixed):
    """Zero rows/columns of fixed variables and put 1 on their diagonal."""
    isfix = np.zeros(n, bool)
    isfix[fixed] = True
    rows = np.repeat(np.arange(n), np.diff(rowp))
    kill = isfix[rows] | isfix[cols]
    vals = vals.copy()
    vals[kill] = 0.0
    vals[kill & (rows == cols)] = 1.0
    return vals


def _barrier_sigma(rng, m, lb, ub, mu, active_frac):
    """Sigma = zl/(x-lb) + zu/(ub-x) at a late interior-point iterate with
    zl, zu on the central path (z = mu/gap) and a fraction of the bounds
    nearly active."""
    lo = lb if np.isfinite(lb) else ub - 4.0
    hi = ub if np.isfinite(ub) else lb + 4.0
    x = lo + (hi - lo) * (0.05 + 0.9 * rng.random(m))
    act = rng.random(m) < active_frac
    side = rng.random(m) < 0.5
    gap = 10.0 ** rng.uniform(-6, -2, size=m)
    x = np.where(act & side, lb + gap, x)
    x = np.where(act & ~side, ub - gap, x)
    sig = np.zeros(m)
    if np.isfinite(lb):
        g = x - lb
        sig += (mu / g) / g
    if np.isfinite(ub):
        g = ub - x
        sig += (mu / g) / g
    return x, sig


def rosenbrock(n=2000, seed=0, mu=1e-2, delta_c=0.0):
    """Chained Rosenbrock KKT at x0 ~ U[-2, 2].

    Layout per index i: x_i at 3i, then (for i < n-1) slack s_i at 3i+1 and
    multiplier lambda_i at 3i+2.
    """
    rng = np.random.default_rng(seed)
    m = n - 1
    x = rng.uniform(-2.0, 2.0, size=n)
    lam = rng.uniform(0.0, 1.0, size=m)
    ix = 3 * np.arange(n)
    isl = 3 * np.arange(m) + 1
    il = 3 * np.arange(m) + 2
    N = n + 2 * m
    r, c, v = [], [], []

    def add(a, b, val):
        r.append(a)
        c.append(b)
        v.append(val)

    # Hessian of sum (1-x_i)^2 + 100 (x_{i+1} - x_i^2)^2 + sum lam_i c_i
    h_d = np.zeros(n)
    h_d[:m] += 2.0 + 1200.0 * x[:m] ** 2 - 400.0 * x[1:] + 2.0 * lam
    h_d[1:] += 200.0 + 2.0 * lam
    add(ix, ix, h_d)
    off = -400.0 * x[:m]
    add(ix[:m], ix[1:], off)
    add(ix[1:], ix[:m], off)
    # Jacobian of c_i - s_i
    for a, val in ((ix[:m], 2.0 * x[:m]), (ix[1:], 2.0 * x[1:]), (isl, -np.ones(m))):
        add(il, a, val)
        add(a, il, val)
    rowp, cols, vals = _csr_from_coo(N, np.concatenate(r), np.concatenate(c),
                                     np.concatenate(v))
    diag = np.zeros(N)
    # bounds [-2, 2] on x; slack upper bound 1.5 (c(x) <= 1.5)
    _, sx = _barrier_sigma(rng, n, -2.0, 2.0, mu, 0.1)
    _, ss = _barrier_sigma(rng, m, -np.inf, 1.5, mu, 0.1)
    diag[ix] = sx + 1e-12
    diag[isl] = ss + 1e-12
    diag[il] = -delta_c
    rhs = rng.standard_normal(N)
    primal = np.sort(np.concatenate([ix, isl]))
    return KKTSnapshot("rosenbrock", N, rowp, cols, vals, diag, rhs, primal, il,
                       dict(n_x=n, n_con=m))


def _q1_2d():
    Me = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]], float) / 36.0
    Ke = np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]],
                  float) / 6.0
    return Me, Ke


def poisson_targets(N, seed=0, ngauss=8):
    """y_d at the mesh nodes: sum of seeded Gaussians on [0,1]^2."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0, 1, size=(ngauss, 2))
    wid = rng.uniform(0.05, 0.2, size=ngauss)
    amp = rng.uniform(-1, 1, size=ngauss)
    t = np.linspace(0, 1, N + 1)
    X, Y = np.meshgrid(t, t, indexing="ij")
    yd = np.zeros_like(X)
    for k in range(ngauss):
        yd += amp[k] * np.exp(-((X - cen[k, 0]) ** 2 + (Y - cen[k, 1]) ** 2)
                              / (2 * wid[k] ** 2))
    return yd.ravel(), rng


def poisson2d(N=256, seed=0, alpha=1e-4, mu=1e-6, delta_c=0.0):
    """2D Poisson distributed-control QP KKT on an N x N Q1 mesh.

    Unknowns per mesh node k (row-major over (N+1)^2 nodes): state y at 3k,
    control u at 3k+1, adjoint lambda at 3k+2.  KKT blocks:
    [[M, 0, K^T], [0, alpha M, -M^T], [K, -M, 0]]; boundary y and lambda
    are fixed (identity rows).
    """
    yd, rng = poisson_targets(N, seed)
    nn = (N + 1) ** 2
    h = 1.0 / N
    Me, Ke = _q1_2d()
    Me = Me * h * h
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    b = (ii * (N + 1) + jj).ravel()
    conn = np.stack([b, b + (N + 1), b + (N + 1) + 1, b + 1], axis=1)  # ccw
    er = np.repeat(conn, 4, axis=1).ravel()
    ec = np.tile(conn, (1, 4)).ravel()
    ne = conn.shape[0]
    mv = np.tile(Me.ravel(), ne)
    kv = np.tile(Ke.ravel(), ne)
    Y, U, L = 3 * er, 3 * er + 1, 3 * er + 2
    Yc, Uc, Lc = 3 * ec, 3 * ec + 1, 3 * ec + 2
    r = np.concatenate([Y, U, L, Yc, L, U])
    c = np.concatenate([Yc, Uc, Yc, L, Uc, Lc])
    v = np.concatenate([mv, alpha * mv, kv, kv, -mv, -mv])
    n = 3 * nn
    rowp, cols, vals = _csr_from_coo(n, r, c, v)
    gi, gj = np.divmod(np.arange(nn), N + 1)
    bnd = np.nonzero((gi == 0) | (gj == 0) | (gi == N) | (gj == N))[0]
    fixed = np.sort(np.concatenate([3 * bnd, 3 * bnd + 2]))
    vals = _fix_rows(n, rowp, cols, vals, fixed)
    diag = np.zeros(n)
    _, su = _barrier_sigma(rng, nn, -1.0, 1.0, mu, 0.2)
    diag[3 * np.arange(nn) + 1] = su + 1e-12
    isfix = np.zeros(n, bool)
    isfix[fixed] = True
    yidx = 3 * np.arange(nn)
    diag[yidx[~isfix[yidx]]] += 1e-12
    lidx = 3 * np.arange(nn) + 2
    diag[lidx[~isfix[lidx]]] = -delta_c
    rhs = rng.standard_normal(n)
    rhs[fixed] = 0.0
    primal = np.sort(np.concatenate([yidx, 3 * np.arange(nn) + 1]))
    dual = lidx[~isfix[lidx]]
    return KKTSnapshot("poisson2d", n, rowp, cols, vals, diag, rhs, primal, dual,
                       dict(N=N, alpha=alpha, n_fixed=len(fixed), yd=yd))


def poisson2d_elements(N=256, alpha=1e-4):
    """The element view of :func:`poisson2d`'s KKT values: one component
    group of N^2 Q1 elements, each with the 12 variables
    (y_0..y_3, u_0..u_3, lambda_0..lambda_3) of its 4 nodes and the 12 x 12
    Hessian of its Lagrangian term
    [[M_e, 0, K_e^T], [0, alpha M_e, -M_e], [K_e, -M_e, 0]].

    Returns dict(n, rowp, cols, idx (N^2 x 12), H (N^2 x 12 x 12), fixed):
    assembling H into (rowp, cols) with the boundary y and lambda fixed gives
    poisson2d(N).vals (up to the order of floating-point additions)."""
    P = poisson2d(N=N, alpha=alpha)
    h = 1.0 / N
    Me, Ke = _q1_2d()
    Me = Me * h * h
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    b = (ii * (N + 1) + jj).ravel()
    conn = np.stack([b, b + (N + 1), b + (N + 1) + 1, b + 1], axis=1)
    idx = np.concatenate([3 * conn, 3 * conn + 1, 3 * conn + 2], axis=1).astype(np.int32)
    He = np.zeros((12, 12))
    He[0:4, 0:4] = Me
    He[4:8, 4:8] = alpha * Me
    He[8:12, 0:4] = Ke
    He[0:4, 8:12] = Ke.T
    He[8:12, 4:8] = -Me
    He[4:8, 8:12] = -Me.T
    H = np.broadcast_to(He, (conn.shape[0], 12, 12)).copy()
    nn = (N + 1) ** 2
    gi, gj = np.divmod(np.arange(nn), N + 1)
    bnd = np.nonzero((gi == 0) | (gj == 0) | (gi == N) | (gj == N))[0]
    fixed = np.sort(np.concatenate([3 * bnd, 3 * bnd + 2])).astype(np.int32)
    return dict(n=P.n, rowp=P.rowp, cols=P.cols, idx=idx, H=H, fixed=fixed, vals=P.vals)


def poisson2d_ipm_state(N=256, seed=0, mu=1e-6, active_frac=0.2):
    """An interior-point iterate of :func:`poisson2d`'s problem in the
    reference's OptProblemInfo layout (include/interior_point_backend.h:54-62):
    the primals are the free states y (no bounds) and the controls u in
    [-1, 1]; the constraints are the free adjoints lambda (target 0); the
    fixed boundary y and lambda are neither.  u sits strictly inside its
    bounds with a fraction nearly active, the bound duals on the central path
    (z = mu / gap), and grad, the primal step and the dual steps are seeded
    normals.

    Returns dict(n, pidx, lbx, ubx, cidx, lbh, mu, x, zl, zu, grad, dx, dzl,
    dzu)."""
    rng = np.random.default_rng(seed + 17)
    nn = (N + 1) ** 2
    n = 3 * nn
    gi, gj = np.divmod(np.arange(nn), N + 1)
    bnd = (gi == 0) | (gj == 0) | (gi == N) | (gj == N)
    y = 3 * np.arange(nn)[~bnd]
    u = 3 * np.arange(nn) + 1
    lam = 3 * np.arange(nn)[~bnd] + 2
    pidx = np.sort(np.concatenate([y, u])).astype(np.int32)
    isu = (pidx % 3) == 1
    lbx = np.where(isu, -1.0, -np.inf)
    ubx = np.where(isu, 1.0, np.inf)
    x = rng.standard_normal(n)
    xu = -1.0 + 2.0 * (0.05 + 0.9 * rng.random(len(u)))
    act = rng.random(len(u)) < active_frac
    side = rng.random(len(u)) < 0.5
    gap = 10.0 ** rng.uniform(-6, -2, size=len(u))
    xu = np.where(act & side, -1.0 + gap, xu)
    xu = np.where(act & ~side, 1.0 - gap, xu)
    x[u] = xu
    xp = x[pidx]
    zl = np.where(isu, mu / (xp - lbx), 0.0)
    zu = np.where(isu, mu / (ubx - xp), 0.0)
    grad = rng.standard_normal(n)
    return dict(n=n, pidx=pidx, lbx=lbx, ubx=ubx, cidx=lam.astype(np.int32),
                lbh=np.zeros(len(lam)), mu=mu, x=x, zl=zl, zu=zu, grad=grad,
                dx=rng.standard_normal(n), dzl=rng.standard_normal(len(pidx)),
                dzu=rng.standard_normal(len(pidx)))


def _cartpole_derivs(q, u, w, g=9.81, L0=0.5, m1=1.0, m2=0.3):
    """f(q,u), df/d(q,u) and the Hessian of w^T f for every node (torch)."""
    import torch
    from torch.func import hessian, jacrev, vmap

    def f(z):
        q1, q3, uu = z[1], z[3], z[4]
        s, c = torch.sin(q1), torch.cos(q1)
        den = m1 + m2 * s * s
        a = (uu + m2 * L0 * s * q3 * q3 + m2 * g * c * s) / den
        b = (-uu * c - m2 * L0 * c * s * q3 * q3 - (m1 + m2) * g * s) / (L0 * den)
        return torch.stack([z[2], z[3], a, b])

    Z = torch.tensor(np.concatenate([q, u[:, None]], axis=1), dtype=torch.float64)
    W = torch.tensor(w, dtype=torch.float64)
    F = vmap(f)(Z)
    J = vmap(jacrev(f))(Z)
    H = vmap(hessian(lambda z, ww: (ww * f(z)).sum()))(Z, W)
    return F.numpy(), J.numpy(), H.numpy()


def cartpole(N=20000, seed=0, tf=None, mu=1e-4, delta_c=0.0):
    """Trapezoidal collocation of the cart-pole with tracking cost.

    Node k holds z_k = (q_k[0:4], u_k) at 9k..9k+4 and, for k < N-1, the
    defect multipliers lambda_k[0:4] at 9k+5..9k+8.
    Defect_k = q_{k+1} - q_k - h/2 (f(z_k) + f(z_{k+1})).
    Cost 0.5||q - q_ref||^2 + 0.5 u^2 with q_ref ~ U[-1, 1].
    """
    rng = np.random.default_rng(seed)
    tf = tf if tf is not None else 2.0 * N / 100.0
    h = tf / (N - 1)
    q = rng.uniform(-1, 1, size=(N, 4))
    u = rng.uniform(-1, 1, size=N)
    lam = rng.standard_normal((N - 1, 4))
    # w_k = lambda weight on f(z_k): -(h/2)(lambda_{k-1} + lambda_k)
    wsum = np.zeros((N, 4))
    wsum[:-1] += lam
    wsum[1:] += lam
    _, J, H = _cartpole_derivs(q, u, -0.5 * h * wsum)
    n = 9 * N - 4
    zi = 9 * np.arange(N)[:, None] + np.arange(5)[None, :]       # N x 5
    li = 9 * np.arange(N - 1)[:, None] + 5 + np.arange(4)[None, :]  # (N-1) x 4
    r, c, v = [], [], []
    Hn = H + np.eye(5)[None]  # tracking cost Hessian
    r.append(np.repeat(zi, 5, axis=1).ravel())
    c.append(np.tile(zi, (1, 5)).ravel())
    v.append(Hn.reshape(N, 25).ravel())
    # d defect_k / d z_k = -[I 0] - h/2 J_k ; d defect_k / d z_{k+1} = [I 0] - h/2 J_{k+1}
    E = np.zeros((4, 5))
    E[:, :4] = np.eye(4)
    A0 = -E[None] - 0.5 * h * J[:-1]
    A1 = E[None] - 0.5 * h * J[1:]
    for A, zz in ((A0, zi[:-1]), (A1, zi[1:])):
        rr = np.repeat(li, 5, axis=1).ravel()
        cc = np.tile(zz, (1, 4)).ravel()
        vv = A.reshape(N - 1, 20).ravel()
        r += [rr, cc]
        c += [cc, rr]
        v += [vv, vv]
    rowp, cols, vals = _csr_from_coo(n, np.concatenate(r), np.concatenate(c),
                                     np.concatenate(v))
    diag = np.zeros(n)
    _, su = _barrier_sigma(rng, N, -10.0, 10.0, mu, 0.05)
    _, sx = _barrier_sigma(rng, N, -5.0, 5.0, mu, 0.05)
    diag[zi[:, 4]] = su
    diag[zi[:, 0]] = sx
    diag[zi.ravel()] += 1e-12
    diag[li.ravel()] = -delta_c
    rhs = rng.standard_normal(n)
    return KKTSnapshot("cartpole", n, rowp, cols, vals, diag, rhs,
                       np.sort(zi.ravel()), np.sort(li.ravel()), dict(N=N, h=h))


def _hex8_elasticity(nu=0.3):
    """Unit-modulus stiffness (24x24) and scalar mass (8x8) of a unit cube
    trilinear element, 2x2x2 Gauss quadrature."""
    g = np.array([-1, 1]) / np.sqrt(3.0)
    nodes = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                      [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float) * 2 - 1
    lam = nu / ((1 + nu) * (1 - 2 * nu))
    mu_ = 1.0 / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu_
    D[np.arange(3, 6), np.arange(3, 6)] = mu_
    Ke = np.zeros((24, 24))
    Me = np.zeros((8, 8))
    for a in g:
        for b in g:
            for cq in g:
                p = np.array([a, b, cq])
                Nn = np.prod(1 + nodes * p, axis=1) / 8
                dN = np.zeros((8, 3))
                for d in range(3):
                    t = nodes[:, d].copy()
                    o = [e for e in range(3) if e != d]
                    dN[:, d] = t * (1 + nodes[:, o[0]] * p[o[0]]) * \
                        (1 + nodes[:, o[1]] * p[o[1]]) / 8
                dN = dN * 2.0  # d/dx on [0,1] = 2 d/dxi
                B = np.zeros((6, 24))
                for k in range(8):
                    B[0, 3 * k] = dN[k, 0]
                    B[1, 3 * k + 1] = dN[k, 1]
                    B[2, 3 * k + 2] = dN[k, 2]
                    B[3, 3 * k] = dN[k, 1]
                    B[3, 3 * k + 1] = dN[k, 0]
                    B[4, 3 * k + 1] = dN[k, 2]
                    B[4, 3 * k + 2] = dN[k, 1]
                    B[5, 3 * k] = dN[k, 2]
                    B[5, 3 * k + 2] = dN[k, 0]
                w = 1.0 / 8.0  # det J of [0,1]^3 map times unit weights
                Ke += B.T @ D @ B * w
                Me += np.outer(Nn, Nn) * w
    return Ke, Me


def elasticity3d(N=40, seed=0, alpha=1e-4, mu=1e-6, delta_c=0.0):
    """3D linear-elasticity distributed-control QP KKT on an N^3 hex mesh.

    Per mesh node k: displacement y (3 dofs) at 9k..9k+2, body-force control
    u (3) at 9k+3..9k+5, adjoint lambda (3) at 9k+6..9k+8.  Blocks as in
    :func:`poisson2d` with the elasticity stiffness K(E_e) and the vector
    mass matrix; nodes on the x = 0 face are clamped (y, lambda fixed);
    -1 <= u <= 1.
    """
    rng = np.random.default_rng(seed)
    Ke1, Me1 = _hex8_elasticity()
    h = 1.0 / N
    nn = (N + 1) ** 3
    E = np.exp(rng.normal(0.0, 0.5, size=N ** 3))
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    base = ((i * (N + 1) + j) * (N + 1) + k).ravel()
    sx, sy, sz = (N + 1) ** 2, N + 1, 1
    off = np.array([0, sx, sx + sy, sy, sz, sx + sz, sx + sy + sz, sy + sz])
    conn = base[:, None] + off[None, :]                     # ne x 8
    dof = (9 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)
    ne = conn.shape[0]
    Kv = (E[:, None] * (Ke1 * h).ravel()[None, :]).ravel()   # stiffness ~ h
    Mv3 = np.kron(Me1 * h ** 3, np.eye(3))                   # vector mass
    Mv = np.tile(Mv3.ravel(), ne)
    er = np.repeat(dof, 24, axis=1).ravel()
    ec = np.tile(dof, (1, 24)).ravel()
    Y, U, L = er, er + 3, er + 6
    Yc, Uc, Lc = ec, ec + 3, ec + 6
    r = np.concatenate([Y, U, L, Yc, L, U])
    c = np.concatenate([Yc, Uc, Yc, L, Uc, Lc])
    v = np.concatenate([Mv, alpha * Mv, Kv, Kv, -Mv, -Mv])
    n = 9 * nn
    rowp, cols, vals = _csr_from_coo(n, r, c, v)
    gi = np.arange(nn) // ((N + 1) ** 2)
    clamp = np.nonzero(gi == 0)[0]
    fixed = np.sort(np.concatenate([9 * clamp[:, None] + np.arange(3),
                                    9 * clamp[:, None] + 6 + np.arange(3)], axis=1).ravel())
    vals = _fix_rows(n, rowp, cols, vals, fixed)
    diag = np.zeros(n)
    uidx = (9 * np.arange(nn)[:, None] + 3 + np.arange(3)).ravel()
    _, su = _barrier_sigma(rng, len(uidx), -1.0, 1.0, mu, 0.2)
    diag[uidx] = su + 1e-12
    isfix = np.zeros(n, bool)
    isfix[fixed] = True
    yidx = (9 * np.arange(nn)[:, None] + np.arange(3)).ravel()
    diag[yidx[~isfix[yidx]]] += 1e-12
    lidx = (9 * np.arange(nn)[:, None] + 6 + np.arange(3)).ravel()
    diag[lidx[~isfix[lidx]]] = -delta_c
    rhs = rng.standard_normal(n)
    rhs[fixed] = 0.0
    primal = np.sort(np.concatenate([yidx, uidx]))
    return KKTSnapshot("elasticity3d", n, rowp, cols, vals, diag, rhs, primal,
                       lidx[~isfix[lidx]], dict(N=N, alpha=alpha))


def _bilinear(N, nc):
    """(N+1)^2 x nc^2 bilinear interpolation from an nc x nc node grid on
    [0,1]^2 to the (N+1)^2 mesh nodes (row-major in both)."""
    t = np.linspace(0.0, 1.0, N + 1)
    s = t * (nc - 1)
    i0 = np.minimum(np.floor(s).astype(int), nc - 2)
    w = s - i0
    rows, cols, vals = [], [], []
    for a in range(N + 1):
        for da, wa in ((0, 1.0 - w[a]), (1, w[a])):
            for db in (0, 1):
                b = np.arange(N + 1)
                wb = np.where(db == 0, 1.0 - w, w)
                rows.append(a * (N + 1) + b)
                cols.append((i0[a] + da) * nc + i0 + db)
                vals.append(wa * wb)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=((N + 1) ** 2, nc * nc))


def poisson2d_sensitivity(N=256, ncoarse=32, nout=64, seed=0, **kw):
    """Post-optimality sensitivity workload on :func:`poisson2d`: the target
    field is y_d = B p with p the ncoarse^2 coefficients of a bilinear field
    on a coarse grid; outputs J_k(z) = c_k^T z with c_k ~ N(0, 1).

    Returns (snapshot, dict) where the dict holds B, the y-row map, the
    gradient Jacobian dg/dp (n x ncoarse^2, sparse), the KKT right-hand side
    map r(p) = -(dg/dp) p, the outputs C (nout x n) and p0 (the Gaussian
    target sampled on the coarse grid)."""
    P = poisson2d(N=N, seed=seed, **kw)
    nn = (N + 1) ** 2
    h = 1.0 / N
    Me, _ = _q1_2d()
    Me = Me * h * h
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    b = (ii * (N + 1) + jj).ravel()
    conn = np.stack([b, b + (N + 1), b + (N + 1) + 1, b + 1], axis=1)
    M = sp.coo_matrix((np.tile(Me.ravel(), conn.shape[0]),
                       (np.repeat(conn, 4, axis=1).ravel(), np.tile(conn, (1, 4)).ravel())),
                      shape=(nn, nn)).tocsr()
    B = _bilinear(N, ncoarse)
    gi, gj = np.divmod(np.arange(nn), N + 1)
    free = ~((gi == 0) | (gj == 0) | (gi == N) | (gj == N))
    MB = (sp.diags(free.astype(float)) @ M @ B).tocsr()        # free y rows only
    Py = sp.csr_matrix((np.ones(nn), (3 * np.arange(nn), np.arange(nn))), shape=(P.n, nn))
    dgdp = -(Py @ MB).tocsr()                                   # g_y = M (y - B p) + ...
    rng = np.random.default_rng(seed + 1)
    C = rng.standard_normal((nout, P.n))
    tc = np.linspace(0, 1, ncoarse)
    Xc, Yc = np.meshgrid(tc, tc, indexing="ij")
    yd_full, _ = poisson_targets(N, seed)
    p0 = (B.T @ yd_full) / np.maximum(B.T @ np.ones(nn), 1e-300)
    return P, dict(B=B, dgdp=dgdp, C=C, p0=p0, ncoarse=ncoarse)


GENERATORS = {
    "rosenbrock": rosenbrock,
    "cartpole": cartpole,
    "poisson2d": poisson2d,
    "elasticity3d": elasticity3d,
}
