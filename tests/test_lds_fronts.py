"""Fronts held whole in LDS: the right-looking kernels (k_front_rl, one
workgroup per front of a tick; k_front_chain, one workgroup walking a run
of single-front ticks) against the oracle node by node, and against the
blocked LDS kernel (k_front_lds, selected with KKT_LDL_RL_FRONTS=0 when the
handle is created) on the same plan.

The cart-pole KKT under AMD is one chain of 2,858 fronts of 68 rows (every
tick holds one LDS front: a single k_front_chain launch); under nested
dissection its fronts are LDS fronts many per tick.  Bar: per-node pivot,
delay and 2x2 counts below the root identical to the oracle's, inertia
identical, refined residual <= 1e-10.
"""

import os

import numpy as np
import pytest

from amigo_amd import workloads as W
from amigo_amd.ldl import FP32_REFINE, FP64, CSRMat, OrderingType, SparseLDL
from oracle.pyoracle import AMD, GIVEN, OracleLDL

pytestmark = pytest.mark.gpu


def _relres(K, x, b):
    return np.linalg.norm(K @ x - b) / np.linalg.norm(b)


def _node_parity(s, o):
    a, b = s.node_info(), o.node_info()
    nonroot = b["root"] == 0
    assert np.array_equal(a["npiv"][nonroot], b["npiv"][nonroot])
    assert np.array_equal(a["ndel"][nonroot], b["ndel"][nonroot])
    assert np.array_equal(a["n2x2"][nonroot], b["n2x2"][nonroot])


@pytest.mark.parametrize("order", ["amd", "nd"])
@pytest.mark.parametrize("prec", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
def test_cartpole_full_size_node_parity(order, prec):
    print("cartpole", order, prec, flush=True)
    P = W.cartpole()
    ot = OrderingType.AMD if order == "amd" else OrderingType.NESTED_DISSECTION
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=ot, precision=prec,
                  max_refine=30)
    assert s.factor(P.diag) == 0
    st = s.stats()
    if order == "amd":
        assert st["num_levels"] == st["num_nodes"]  # one front per level: a chain
    if order == "amd":
        o = OracleLDL(P.n, P.rowp, P.cols, order=AMD)
        assert np.array_equal(s.get_perm()[0], o.perm()[0])
    else:
        o = OracleLDL(P.n, P.rowp, P.cols, order=GIVEN, perm=s.get_perm()[0])
    o.factor(P.vals, P.diag)
    assert s.get_inertia() == o.inertia()
    _node_parity(s, o)
    x = P.rhs.copy()
    s.solve(x, refine=True)
    assert _relres(P.matrix(), x, P.rhs) <= 1e-10


@pytest.mark.parametrize("name,kw,order", [("cartpole", dict(N=2000), OrderingType.AMD),
                                           ("poisson2d", dict(N=64),
                                            OrderingType.NESTED_DISSECTION),
                                           ("rosenbrock", {}, OrderingType.AMD)])
@pytest.mark.parametrize("prec", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
def test_right_looking_matches_blocked_lds_kernel(name, kw, order, prec):
    """Same plan, same decisions: the right-looking kernels and the blocked
    k_front_lds give the same per-node pivot outcome and inertia, and
    solutions that agree to rounding."""
    print("rl vs blocked", name, prec, flush=True)
    P = W.GENERATORS[name](**kw)
    out = []
    for flag in ("1", "0"):
        os.environ["KKT_LDL_RL_FRONTS"] = flag
        try:
            s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=order,
                          precision=prec, max_refine=30)
        finally:
            del os.environ["KKT_LDL_RL_FRONTS"]
        assert s.factor(P.diag) == 0
        x = P.rhs.copy()
        s.solve(x, refine=True)
        out.append((s.get_inertia(), s.node_info(), x))
    (i1, n1, x1), (i0, n0, x0) = out
    assert i1 == i0
    for k in ("npiv", "ndel", "n2x2"):
        assert np.array_equal(n1[k], n0[k]), k
    assert np.abs(x1 - x0).max() <= 1e-8 * np.abs(x0).max()
    assert _relres(P.matrix(), x1, P.rhs) <= 1e-10


@pytest.mark.parametrize("prec", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
def test_large_root_threshold_then_bunch_kaufman(prec):
    """A root front too large for LDS is factored with the threshold tests
    first and Bunch-Kaufman on the columns they delay (mode 3); with
    KKT_LDL_ROOT_BK=1 it keeps Bunch-Kaufman throughout.  Both give the
    oracle's inertia (the reference's root is LAPACK's Bunch-Kaufman; the
    inertia does not depend on the pivot sequence), the same pivot outcome
    below the root, all of the root's columns pivoted, and refined
    residuals <= 1e-10 on the full C2 KKT under nested dissection."""
    print("root mode", prec, flush=True)
    P = W.poisson2d()
    res = {}
    for flag in ("0", "1"):
        os.environ["KKT_LDL_ROOT_BK"] = flag
        try:
            s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                          order=OrderingType.NESTED_DISSECTION, precision=prec, max_refine=30)
        finally:
            del os.environ["KKT_LDL_ROOT_BK"]
        assert s.factor(P.diag) == 0
        x = P.rhs.copy()
        s.solve(x, refine=True)
        assert _relres(P.matrix(), x, P.rhs) <= 1e-10
        res[flag] = (s.get_inertia(), s.node_info(), x, s.get_perm()[0])
    o = OracleLDL(P.n, P.rowp, P.cols, order=GIVEN, perm=res["0"][3])
    o.factor(P.vals, P.diag)
    b = o.node_info()
    for flag in ("0", "1"):
        inertia, info, _, _ = res[flag]
        assert inertia == o.inertia()
        root = b["root"] != 0
        assert np.array_equal(info["npiv"], b["npiv"])
        assert info["ndel"][root].sum() == 0
        for k in ("ndel", "n2x2"):
            assert np.array_equal(info[k][~root], b[k][~root]), k
    assert np.abs(res["0"][2] - res["1"][2]).max() <= 1e-8 * np.abs(res["1"][2]).max()


def test_chain_overflow_replans():
    """Minimal front slack on the AMD cart-pole chain: fronts overflow inside
    the single k_front_chain launch (each failed front marks its ancestors
    failed within the same launch), the host enlarges the slack and
    refactors; the result matches the oracle node by node."""
    P = W.cartpole(N=3000)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD,
                  delay_growth=1.0)
    assert s.factor(P.diag) == 0
    assert s.stats()["replans"] > 0
    o = OracleLDL(P.n, P.rowp, P.cols, order=AMD)
    o.factor(P.vals, P.diag)
    assert s.get_inertia() == o.inertia()
    _node_parity(s, o)
    x = P.rhs.copy()
    s.solve(x, refine=True)
    assert _relres(P.matrix(), x, P.rhs) <= 1e-10


def test_elasticity_fronts_beyond_lds_bookkeeping():
    """30^3 elasticity KKT (n = 268,119) in FP32 + refinement: its largest
    front (13,392 rows) takes the global-memory panel with the front's
    vars/pre kept in HBM (they do not fit LDS beside the panel), the others
    the LDS panels; inertia equal to the structural one (every row positive
    except the free multipliers) and a refined residual <= 1e-10."""
    from amigo_amd.ldl import FP32_REFINE
    P = W.elasticity3d(N=30)
    s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                  order=OrderingType.NESTED_DISSECTION, precision=FP32_REFINE, max_refine=30)
    assert s.stats()["max_front"] > 12288
    assert s.factor(P.diag) == 0
    nd = len(P.dual_idx)
    assert s.get_inertia() == (P.n - nd, nd)
    x = P.rhs.copy()
    assert s.solve(x, refine=True) == 0
    assert _relres(P.matrix(), x, P.rhs) <= 1e-10


@pytest.mark.parametrize("prec", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
def test_chain_solve_bitwise_equal_level_launches(prec):
    """The AMD cart-pole KKT's solve levels hold one front each: the forward
    and backward sweeps walk the chain in one launch per right-hand side
    (k_forward_chain / k_backward_chain) instead of one launch per level
    (KKT_LDL_SOLVE_CHAINS=0).  Same per-node code: bitwise equal for 1 and
    3 right-hand sides and for the refined solve, refined residual <= 1e-10."""
    import torch
    P = W.cartpole()
    hs = []
    for flag in ("1", "0"):
        os.environ["KKT_LDL_SOLVE_CHAINS"] = flag
        try:
            s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals), order=OrderingType.AMD,
                          precision=prec, max_refine=30)
        finally:
            del os.environ["KKT_LDL_SOLVE_CHAINS"]
        assert s.factor(P.diag) == 0
        hs.append(s)
    g = torch.Generator(device="cuda").manual_seed(7)
    B = torch.randn(3, P.n, dtype=torch.float64, device="cuda", generator=g).t()
    for nrhs in (1, 3):
        xa, xb = B[:, :nrhs].clone(), B[:, :nrhs].clone()
        hs[0].solve(xa)
        hs[1].solve(xb)
        assert torch.equal(xa, xb), nrhs
    xa = torch.as_tensor(P.rhs, device="cuda").clone()
    xb = xa.clone()
    assert hs[0].solve(xa, refine=True) == 0
    assert hs[1].solve(xb, refine=True) == 0
    assert torch.equal(xa, xb)
    assert _relres(P.matrix(), xa.cpu().numpy(), P.rhs) <= 1e-10


@pytest.mark.parametrize("prec", [FP64, FP32_REFINE], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", ["poisson2d", "cartpole"])
def test_lds_front_instances_bitwise_equal(name, prec):
    """LDS front launches take the kernel instance for their front bound (32,
    48, 64, 80 or 96 rows: fewer registers for smaller bounds); with
    KKT_LDL_RL_INSTANCES=0 every launch takes the 96-row instance.  Products
    are rounded before they are added in every instance, so the factors are
    the same bit for bit: equal solutions for 1 right-hand side and for the
    refined solve, equal inertia, under nested dissection."""
    import torch
    P = W.GENERATORS[name]()
    hs = []
    for flag in ("1", "0"):
        os.environ["KKT_LDL_RL_INSTANCES"] = flag
        try:
            s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, P.vals),
                          order=OrderingType.NESTED_DISSECTION, precision=prec)
        finally:
            del os.environ["KKT_LDL_RL_INSTANCES"]
        assert s.factor(P.diag) == 0
        hs.append(s)
    assert hs[0].get_inertia() == hs[1].get_inertia()
    g = torch.Generator(device="cuda").manual_seed(5)
    b = torch.randn(P.n, dtype=torch.float64, device="cuda", generator=g)
    xa, xb = b.clone(), b.clone()
    hs[0].solve(xa)
    hs[1].solve(xb)
    assert torch.equal(xa, xb)
    xa = torch.as_tensor(P.rhs, device="cuda").clone()
    xb = xa.clone()
    assert hs[0].solve(xa, refine=True) == 0
    assert hs[1].solve(xb, refine=True) == 0
    assert torch.equal(xa, xb)
    assert _relres(P.matrix(), xa.cpu().numpy(), P.rhs) <= 1e-10
