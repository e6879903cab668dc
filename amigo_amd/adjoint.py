"""Adjoint solves on the stored LDL^T factors: post-optimality derivatives.

``KKTSolve`` is a ``torch.autograd.Function`` whose forward solves K x = b
with the factors held by a ``SparseLDL`` (FP64 refinement to the solver's
tolerance) and whose backward solves with the same factors again: K is
symmetric, so the vector-Jacobian product of b -> K^{-1} b is K^{-1} g.
Both directions accept one right-hand side (n,) or a batch (n, k), which the
engine solves in a single multi-RHS call.

``post_optimality_derivatives`` is the adjoint method of the reference's
``Optimizer.compute_post_opt_derivatives`` (amigo/algorithm/ipm_driver.py:
297-353): with the KKT matrix factored at the optimum,

    K psi_i = -d out_i / dx,    dfdp[i, :] = d out_i / dp + (d g / dp)^T psi_i,

where g is the KKT residual (the gradient of the Lagrangian).  The reference
solves one output at a time with nrhs = 1; here all outputs go through one
batched solve.
"""

import numpy as np
import torch


def _as_cuda(x):
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device="cuda")


def _colmajor(x):
    """(n,) stays; (n, k) becomes a column-major (Fortran-order) copy."""
    if x.dim() == 1:
        return x.contiguous().clone()
    return x.t().contiguous().t()


class KKTSolve(torch.autograd.Function):
    """x = K^{-1} b on the stored factors; backward x_bar -> K^{-1} x_bar."""

    @staticmethod
    def forward(ctx, b, solver, refine):
        x = _colmajor(b.detach())
        solver.solve(x, refine=refine)
        ctx.solver = solver
        ctx.refine = refine
        return x

    @staticmethod
    def backward(ctx, gx):
        g = _colmajor(gx.detach())
        ctx.solver.solve(g, refine=ctx.refine)
        return g, None, None


def kkt_solve(solver, b, refine=True):
    """Differentiable K^{-1} b (b: CUDA float64 tensor, (n,) or (n, k))."""
    return KKTSolve.apply(b, solver, refine)


def post_optimality_derivatives(solver, out_wrt_x, grad_wrt_p, out_wrt_p=None,
                                refine=True, as_numpy=True):
    """dfdp (nout x np) by the adjoint method, all outputs in one solve.

    out_wrt_x:  (nout, n) d out / dx at the optimum (dense array or tensor)
    grad_wrt_p: (n, np) d g / dp: scipy sparse (multiplied on the host), a
                dense array or tensor, or a ``DeviceJacobian`` (sparse, kept
                in HBM and multiplied on the GPU)
    out_wrt_p:  (nout, np) explicit d out / dp, or None
    as_numpy:   return a numpy array (host copy); False returns the CUDA
                tensor, so that a caller whose inputs already live in HBM
                moves nothing over PCIe
    """
    A = _as_cuda(out_wrt_x)
    if A.dim() == 1:
        A = A[None, :]
    psi = _colmajor(-A.t())            # n x nout, column major
    solver.solve(psi, refine=refine)   # K psi = -d out / dx
    if isinstance(grad_wrt_p, DeviceJacobian):
        dfdp = grad_wrt_p.tmatmul(psi).t()
    elif hasattr(grad_wrt_p, "tocsr"):   # scipy sparse: (d g/dp)^T psi on the host
        G = grad_wrt_p.tocsr()
        dfdp = torch.as_tensor((G.T @ psi.cpu().numpy()).T)
    else:
        dfdp = psi.t() @ _as_cuda(grad_wrt_p)
    if out_wrt_p is not None:
        dfdp = dfdp + torch.as_tensor(np.asarray(out_wrt_p, dtype=np.float64),
                                      device=dfdp.device)
    return dfdp.cpu().numpy() if as_numpy else dfdp


class DeviceJacobian:
    """A sparse Jacobian d g / dp (n x np) held in HBM as the CSR matrix of
    its transpose, for (d g / dp)^T psi on the GPU (one sparse x dense
    product for every output at once)."""

    def __init__(self, G):
        Gt = G.T.tocsr()
        Gt.sort_indices()
        self.shape = G.shape
        self.t = torch.sparse_csr_tensor(
            torch.as_tensor(Gt.indptr, dtype=torch.int64, device="cuda"),
            torch.as_tensor(Gt.indices, dtype=torch.int64, device="cuda"),
            torch.as_tensor(Gt.data, dtype=torch.float64, device="cuda"),
            size=(G.shape[1], G.shape[0]))

    def tmatmul(self, psi):
        """(d g / dp)^T psi for psi (n x k) on the device: (np x k)."""
        return torch.sparse.mm(self.t, psi)
