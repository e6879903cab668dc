"""Post-optimality derivatives by adjoint solves on the stored factors.

Pinned by the reference's closed form (examples/post_opt_quadratic/
quadratic.py:1-51: f* = (5a^2 - 8ab + 5b^2)/9, so df*/da = (10a - 8b)/9 and
df*/db = (10b - 8a)/9), by central finite differences (h = 1e-6), and by the
direct method (dz/dp = -K^{-1} dg/dp)."""

import numpy as np
import pytest
import scipy.sparse as sp

from amigo_amd import workloads as W
from oracle.pyoracle import NATURAL, OracleLDL


def _quadratic_kkt(a, b):
    # objective x1^2 + x1 x2 + x2^2 - a x1 - b x2: K = Hessian, g = K x - [a, b]
    K = sp.csr_matrix(np.array([[2.0, 1.0], [1.0, 2.0]]))
    K.sort_indices()
    xs = np.linalg.solve(K.toarray(), [a, b])
    out_wrt_x = 2.0 * xs[None, :]          # output f = x1^2 + x2^2
    grad_wrt_p = -np.eye(2)                 # dg/d(a, b)
    exact = np.array([[(10 * a - 8 * b) / 9, (10 * b - 8 * a) / 9]])
    return K, out_wrt_x, grad_wrt_p, exact


@pytest.mark.parametrize("a,b", [(1.0, 3.0), (-2.0, 0.5)])
def test_adjoint_formula_closed_form_on_oracle(a, b):
    """The adjoint formula of ipm_driver.py:332-343 with the oracle's solve."""
    K, out_wrt_x, grad_wrt_p, exact = _quadratic_kkt(a, b)
    o = OracleLDL(2, K.indptr, K.indices, order=NATURAL)
    o.factor(K.data)
    psi = o.solve(-out_wrt_x.T)
    dfdp = (grad_wrt_p.T @ psi).T
    assert np.allclose(dfdp, exact, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("a,b", [(1.0, 3.0), (-2.0, 0.5)])
def test_post_opt_closed_form_gpu(a, b):
    from amigo_amd.adjoint import post_optimality_derivatives
    from amigo_amd.ldl import CSRMat, OrderingType, SparseLDL
    K, out_wrt_x, grad_wrt_p, exact = _quadratic_kkt(a, b)
    s = SparseLDL(CSRMat(2, 2, K.indptr, K.indices, K.data), order=OrderingType.AMD)
    s.factor()
    dfdp = post_optimality_derivatives(s, out_wrt_x, grad_wrt_p)
    assert np.allclose(dfdp, exact, atol=1e-12)


@pytest.mark.gpu
def test_poisson_sensitivity_adjoint_vs_fd_and_direct():
    import torch
    from amigo_amd.adjoint import kkt_solve, post_optimality_derivatives
    from amigo_amd.ldl import CSRMat, OrderingType, SparseLDL
    P, S = W.poisson2d_sensitivity(N=32, ncoarse=8, nout=6)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                  order=OrderingType.NESTED_DISSECTION)
    s.factor(P.diag)
    G = S["dgdp"]
    C = S["C"]
    dfdp = post_optimality_derivatives(s, C, G)           # 6 x 64, one batched solve
    # direct method: dz/dp = -K^{-1} G, all 64 columns in one solve
    Z = -torch.as_tensor(G.toarray(), device="cuda")
    Z = Z.t().contiguous().t()
    s.solve(Z, refine=True)
    direct = C @ Z.cpu().numpy()
    assert np.allclose(dfdp, direct, rtol=1e-9, atol=1e-12 * np.abs(direct).max())
    # central differences of J(p) = C K^{-1} r(p), r(p) = -G p.  J is affine
    # in p, so the difference is exact for any step; a unit step keeps the
    # solve's 1e-10 relative residual from being amplified by 1/h.
    p0 = S["p0"]
    h = 1.0
    rng = np.random.default_rng(3)
    for j in rng.choice(G.shape[1], 8, replace=False):
        e = np.zeros(G.shape[1])
        e[j] = h
        Js = []
        for sgn in (1.0, -1.0):
            z = torch.as_tensor(-(G @ (p0 + sgn * e)), device="cuda")
            s.solve(z, refine=True)
            Js.append(C @ z.cpu().numpy())
        fd = (Js[0] - Js[1]) / (2 * h)
        assert np.allclose(dfdp[:, j], fd, rtol=1e-6, atol=1e-9 * np.abs(dfdp).max())
    # autograd through the factors gives the same Jacobian rows
    p = torch.tensor(p0, device="cuda", requires_grad=True)
    Gt = torch.sparse_csr_tensor(torch.as_tensor(G.indptr, dtype=torch.int64),
                                 torch.as_tensor(G.indices, dtype=torch.int64),
                                 torch.as_tensor(G.data), size=G.shape).to("cuda")
    z = kkt_solve(s, -(Gt @ p))
    Ct = torch.as_tensor(C, device="cuda")
    J = Ct @ z
    for i in range(C.shape[0]):
        (gp,) = torch.autograd.grad(J[i], p, retain_graph=True)
        assert np.allclose(gp.cpu().numpy(), dfdp[i], rtol=1e-9,
                           atol=1e-12 * np.abs(dfdp).max())


@pytest.mark.gpu
def test_c4_full_size_adjoint_vs_oracle_and_central_differences():
    """The post-optimality workload at its stated size: the 256 x 256
    Poisson control KKT (n = 198,147), p = 1,024 bilinear target
    coefficients on a 32 x 32 grid, 64 seeded output functionals.  The 64
    adjoints go through one batched solve on the stored FP64 factors.  Checked
    against (a) the oracle's FP64 LDL^T (the reference's SparseLDL restated)
    in the same elimination order, solving the same 64 adjoint systems as the
    reference's compute_post_opt_derivatives does one at a time
    (amigo/algorithm/ipm_driver.py:332-343), and (b) central differences with
    h = 1e-6 on 8 seeded parameters."""
    import time

    import torch
    from amigo_amd.adjoint import post_optimality_derivatives
    from amigo_amd.ldl import CSRMat, OrderingType, SparseLDL
    from oracle.pyoracle import GIVEN

    P, S = W.poisson2d_sensitivity(N=256, ncoarse=32, nout=64)
    G, C, p0 = S["dgdp"], S["C"], S["p0"]
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                  order=OrderingType.NESTED_DISSECTION)
    assert s.factor(P.diag) == 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dfdp = post_optimality_derivatives(s, C, G)          # 64 x 1024
    torch.cuda.synchronize()
    print(f"64 adjoints + dfdp: {1e3 * (time.perf_counter() - t0):.1f} ms")
    assert dfdp.shape == (64, 1024) and np.all(np.isfinite(dfdp))
    # (a) the oracle: K psi = -C^T for all 64 outputs, dfdp = (G^T psi)^T
    o = OracleLDL(P.n, P.rowp, P.cols, order=GIVEN, perm=s.get_perm()[0])
    o.factor(P.vals, P.diag)
    assert o.inertia() == s.get_inertia()
    psi = o.solve(-C.T)
    ref = (G.T @ psi).T
    err = np.abs(dfdp - ref).max() / np.abs(ref).max()
    print(f"dfdp vs oracle: max relative difference {err:.2e}")
    assert err <= 1e-9
    # the device-resident path: inputs in HBM, (dg/dp)^T psi on the GPU
    from amigo_amd.adjoint import DeviceJacobian
    dd = post_optimality_derivatives(s, torch.as_tensor(C, device="cuda"), DeviceJacobian(G),
                                     as_numpy=False)
    assert isinstance(dd, torch.Tensor) and dd.is_cuda and tuple(dd.shape) == (64, 1024)
    err_d = np.abs(dd.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"device path vs oracle: max relative difference {err_d:.2e}")
    assert err_d <= 1e-9
    # (b) central differences: J(p) = C K^{-1} r(p), r(p) = -G p.  The
    # tolerance comes from the solves' own residuals: with psi_k = K^{-1} c_k,
    # the error of c_k^T z~ is psi_k^T (K z~ - r), at most
    # ||psi_k|| ||K z~ - r||, so the difference quotient is within
    # ||psi_k|| (||res+|| + ||res-||) / (2h) of the exact derivative.
    K = P.matrix()
    Psi = torch.as_tensor(np.asfortranarray(C.T), device="cuda")
    assert s.solve(Psi, refine=True) == 0
    psi_norm = np.linalg.norm(Psi.cpu().numpy(), axis=0)           # ||K^{-1} c_k||
    h = 1e-6
    rng = np.random.default_rng(4)
    worst = worst_ratio = 0.0
    for j in rng.choice(G.shape[1], 8, replace=False):
        Js, rn = [], 0.0
        for sgn in (1.0, -1.0):
            pj = p0.copy()
            pj[j] += sgn * h
            r = -(G @ pj)
            z = torch.as_tensor(r, device="cuda")
            assert s.solve(z, refine=True) == 0
            zh = z.cpu().numpy()
            rn += np.linalg.norm(K @ zh - r)
            Js.append(C @ zh)
        fd = (Js[0] - Js[1]) / (2 * h)
        bound = psi_norm * rn / (2 * h) + 1e-9 * np.abs(dfdp).max()
        diff = np.abs(fd - dfdp[:, j])
        worst = max(worst, diff.max() / np.abs(dfdp[:, j]).max())
        worst_ratio = max(worst_ratio, float((diff / bound).max()))
        assert np.all(diff <= bound), (j, float((diff / bound).max()))
    print(f"central differences (h = 1e-6) vs adjoint: max relative difference {worst:.2e}, "
          f"at most {worst_ratio:.2f} of the residual bound")
