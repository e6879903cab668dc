"""kkt_ldl_solve_refined_noalloc: the refined solve with its buffers in a
static per-GPU scratch.  Through the C ABI it must give what
kkt_ldl_solve_refined gives (bit for bit: the same kernels in the same
order), on two calls in a row, the second with a smaller system, without
any device allocation; sizes beyond its scratch are rejected with a clear
error."""

import ctypes

import numpy as np
import pytest

from amigo_amd import workloads as W
from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SparseLDL

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.tensor(np.asfortranarray(a), device="cuda", dtype=torch.float64).t().contiguous().t()


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _solve(s, which, X):
    """Solve the device tensor X (n x k, column major) in place: "plain"
    (kkt_ldl_solve), "refined" (kkt_ldl_solve_refined) or "noalloc"."""
    import torch
    k = 1 if X.dim() == 1 else X.shape[1]
    p = ctypes.c_void_p(X.data_ptr())
    if which == "noalloc":
        st = s.lib.kkt_ldl_solve_refined_noalloc(s.h, k, p, s.n, _stream())
    else:
        fn = s.lib.kkt_ldl_solve if which == "plain" else s.lib.kkt_ldl_solve_refined
        st = fn(s.h, k, p, s.n, 1, _stream())
    torch.cuda.synchronize()
    return st


def _case(N, prec, nrhs, seed):
    P = W.poisson2d(N=N)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.NESTED_DISSECTION,
                  precision=prec)
    assert s.factor(P.diag) == 0
    rng = np.random.default_rng(seed)
    B = np.column_stack([P.rhs] + [rng.standard_normal(P.n) for _ in range(nrhs - 1)])
    return P, s, B


@pytest.mark.parametrize("prec", [FP32_REFINE, FP64], ids=["fp32", "fp64"])
def test_noalloc_matches_refined_solve_on_two_calls(prec):
    import torch
    # first call: a larger system with 4 right-hand sides
    P, s, B = _case(64, prec, 4, 0)
    warm = _dev(B)
    assert _solve(s, "plain", warm) == 0     # sizes the solve buffers for 4
    ref = _dev(B)
    assert _solve(s, "refined", ref) == 0
    ref_steps = s.stats()["refine_steps"]
    got = _dev(B)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    assert _solve(s, "noalloc", got) == 0
    assert torch.cuda.mem_get_info()[0] == free0
    assert s.stats()["refine_steps"] == ref_steps
    assert torch.equal(got, ref)
    K = P.matrix()
    X = got.cpu().numpy()
    assert np.linalg.norm(K @ X - B) / np.linalg.norm(B) <= 1e-10
    # second call: a smaller system, one right-hand side (the scratch still
    # holds the first call's contents)
    P2, s2, B2 = _case(24, prec, 1, 1)
    ref2 = _dev(B2[:, 0])
    assert _solve(s2, "refined", ref2) == 0
    got2 = _dev(B2[:, 0])
    free1 = torch.cuda.mem_get_info()[0]
    assert _solve(s2, "noalloc", got2) == 0
    assert torch.cuda.mem_get_info()[0] == free1
    assert torch.equal(got2, ref2)
    x2 = got2.cpu().numpy()
    assert np.linalg.norm(P2.matrix() @ x2 - B2[:, 0]) / np.linalg.norm(B2[:, 0]) <= 1e-10


def test_noalloc_rejects_sizes_beyond_its_scratch():
    import torch
    P, s, B = _case(16, FP64, 2, 2)
    x = _dev(B)
    # solve buffers sized for one right-hand side only
    st = _solve(s, "noalloc", x)
    assert st == -7 and b"solve buffers" in s.lib.kkt_ldl_last_error()
    big = torch.zeros((P.n, 65), dtype=torch.float64, device="cuda").t().contiguous().t()
    assert _solve(s, "noalloc", big) == -7
    assert b"exceeds its scratch" in s.lib.kkt_ldl_last_error()
    # ldx * nrhs above 2^20 entries
    ldx = (1 << 20) // 2 + 1
    huge = torch.zeros(ldx * 2, dtype=torch.float64, device="cuda")
    st = s.lib.kkt_ldl_solve_refined_noalloc(s.h, 2, ctypes.c_void_p(huge.data_ptr()), ldx,
                                             _stream())
    assert st == -7 and b"exceeds its scratch" in s.lib.kkt_ldl_last_error()
