"""KKT value assembly: oracle pinned to the reference's own host assembly
(golden fixtures and, where built, the compiled reference), and the GPU
gather kernel bitwise equal to the oracle.

The reference sums the contributions of one CSR entry in contribution order
starting from 0.0 (CSRMat::add_row after CSRMat::zero, include/csr_matrix.h:
215-220, 555-567); the kernel sums them in the same order, so every
comparison here is exact (np.array_equal), not a tolerance."""

import glob
import os

import numpy as np
import pytest

from oracle import kkt_assembly as KA
from oracle import pyoracle as po

from .conftest import ROOT, has_gpu

GOLD = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "assembly", "*.npz")))


def _load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def _case_id(path):
    return os.path.splitext(os.path.basename(path))[0]


# ---------------------------------------------------------------- oracle --

def test_fixtures_present():
    assert {_case_id(p) for p in GOLD} >= {"hand6", "hand6_nofix", "rand300", "poisson8"}


@pytest.mark.parametrize("path", GOLD, ids=_case_id)
def test_oracle_matches_reference_fixture(path):
    g = _load(path)
    v = KA.assemble(int(g["n"]), g["rowp"], g["cols"], g["crow"], g["ccol"], g["cvals"],
                    g["fixed"])
    assert np.array_equal(v, g["vals"])
    slots = KA.contribution_slots(int(g["n"]), g["rowp"], g["cols"], g["crow"], g["ccol"])
    assert int(np.sum(slots < 0)) == int(g["ndropped"])


def test_hand_case_by_hand():
    """The 6 x 6 fixture worked out by hand: duplicates add, entries outside
    the pattern vanish, fixed row 3 (no diagonal in the pattern) and its
    column become 0, and out-of-range fixed indices are ignored."""
    g = _load(os.path.join(ROOT, "tests", "golden", "assembly", "hand6.npz"))
    want = {(0, 0): 3.75, (0, 5): 3.0, (1, 0): 0.5, (1, 2): -1.0, (2, 2): 6.0,
            (4, 4): -2.0, (5, 0): 4.0, (5, 5): 0.125}
    rowp, cols = g["rowp"], g["cols"]
    for r in range(6):
        for p in range(rowp[r], rowp[r + 1]):
            assert g["vals"][p] == want.get((r, int(cols[p])), 0.0), (r, cols[p])
    assert int(g["ndropped"]) == 2


def test_element_contribution_order_brute_force():
    """q = ncomp*(ncomp*e + k) + i addresses (idx[e][k], idx[e][i]), the loop
    order of the reference's add_hessian_kernel."""
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 50, size=(7, 4))
    r, c = KA.element_contributions(idx)
    q = 0
    for e in range(7):
        for k in range(4):
            for i in range(4):
                assert (r[q], c[q]) == (idx[e][k], idx[e][i])
                q += 1


def test_element_values_are_hessian_columns():
    from amigo_amd.assembly import element_values
    rng = np.random.default_rng(4)
    H = rng.standard_normal((3, 5, 5))
    v = element_values(H)
    for e in range(3):
        for k in range(5):
            assert np.array_equal(v[25 * e + 5 * k: 25 * e + 5 * k + 5], H[e] @ np.eye(5)[k])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_against_dense_brute_force(seed):
    """Dense accumulation with explicit masks on small random cases."""
    rng = np.random.default_rng(seed)
    n = 25
    M = rng.random((n, n)) < 0.2
    M = M | M.T | np.eye(n, dtype=bool)
    M[5, 5] = False
    rows, cls = np.nonzero(M)
    rowp = np.searchsorted(rows, np.arange(n + 1)).astype(np.int32)
    cols = cls.astype(np.int32)
    crow = rng.integers(0, n, 400)
    ccol = rng.integers(0, n, 400)
    cv = rng.standard_normal(400)
    fixed = [5, 7, -3, n + 2]
    D = np.zeros((n, n))
    for a, b, x in zip(crow, ccol, cv):
        if M[a, b]:
            D[a, b] += x
    for f in (5, 7):
        D[f, :] = 0.0
        D[:, f] = 0.0
        if M[f, f]:
            D[f, f] = 1.0
    v = KA.assemble(n, rowp, cols, crow, ccol, cv, fixed)
    assert np.array_equal(v, D[rows, cls])


def test_poisson_elements_reproduce_the_kkt_values():
    from amigo_amd import workloads as W
    E = W.poisson2d_elements(N=12)
    r, c = KA.element_contributions(E["idx"])
    cv = np.ascontiguousarray(E["H"].transpose(0, 2, 1)).ravel()
    v = KA.assemble(E["n"], E["rowp"], E["cols"], r, c, cv, E["fixed"])
    np.testing.assert_allclose(v, E["vals"], rtol=0, atol=1e-15)


@pytest.mark.skipif(not po.ref_available(), reason="compiled reference not built")
@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_matches_compiled_reference_random(seed):
    rng = np.random.default_rng(seed)
    n = 120
    M = rng.random((n, n)) < 0.05
    M = M | M.T | np.eye(n, dtype=bool)
    rows, cls = np.nonzero(M)
    rowp = np.searchsorted(rows, np.arange(n + 1)).astype(np.int32)
    cols = cls.astype(np.int32)
    idx = rng.integers(0, n, size=(200, 6))
    r, c = KA.element_contributions(idx)
    cv = rng.standard_normal(len(r)) * 10.0 ** rng.integers(-8, 8, len(r))
    fixed = rng.choice(n, 9, replace=False)
    want = po.ref_assemble(n, rowp, cols, r, c, cv, fixed, ncomp=6)
    assert np.array_equal(KA.assemble(n, rowp, cols, r, c, cv, fixed), want)


# --------------------------------------------------------- C ABI (host) --

def _c_create(crow, ccol, fixed):
    import ctypes
    from amigo_amd import assembly as A
    lib = A._lib()
    rowp = np.array([0, 2, 4], np.int32)
    cols = np.array([0, 1, 0, 1], np.int32)
    crow = np.asarray(crow, np.int32)
    ccol = np.asarray(ccol, np.int32)
    fx = np.asarray(fixed, np.int32)
    h = ctypes.c_void_p()
    st = lib.kkt_asm_create(2, rowp.ctypes.data_as(A._i32p), cols.ctypes.data_as(A._i32p),
                            len(crow), crow.ctypes.data_as(A._i32p),
                            ccol.ctypes.data_as(A._i32p), len(fx), fx.ctypes.data_as(A._i32p),
                            ctypes.byref(h))
    return lib, st, h


def test_c_abi_rejects_contributions_out_of_range():
    lib, st, _ = _c_create([0, 1], [0, 5], [])
    assert st < 0 and b"out of range" in lib.kkt_asm_last_error()


def test_c_abi_ignores_fixed_indices_out_of_range():
    """As FixedVariables ignores them: pattern set-up passes validation (and,
    without a device, fails only at the first HIP allocation)."""
    lib, st, h = _c_create([0, 1], [0, 1], [-4, 9, 1])
    if has_gpu():
        assert st == 0
        lib.kkt_asm_destroy(h)
    else:
        assert st < 0 and b"HIP" in lib.kkt_asm_last_error()


def test_create_checks_arguments_without_a_gpu():
    from amigo_amd.assembly import KKTAssembler
    rowp = np.array([0, 2, 4], np.int32)
    cols = np.array([0, 1, 0, 1], np.int32)
    with pytest.raises(RuntimeError, match="out of range"):
        KKTAssembler(2, rowp, cols, np.array([0, 2], np.int32), np.array([0, 0], np.int32))
    with pytest.raises(ValueError):
        KKTAssembler(2, rowp, cols, np.array([0], np.int32), np.array([0, 1], np.int32))


# ------------------------------------------------------------------- GPU --

def _gpu_assembler(g):
    from amigo_amd.assembly import KKTAssembler
    return KKTAssembler(int(g["n"]), g["rowp"], g["cols"], g["crow"], g["ccol"], g["fixed"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLD, ids=_case_id)
def test_gpu_matches_reference_fixture(path):
    import torch
    g = _load(path)
    A = _gpu_assembler(g)
    assert A.ndropped == int(g["ndropped"])
    host = A.values(g["cvals"])
    assert np.array_equal(host, g["vals"])
    dev = A.values(torch.tensor(g["cvals"], device="cuda"))
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy(), g["vals"])


@pytest.mark.gpu
def test_gpu_poisson_full_size_bitwise_and_repeatable():
    """Q1 Poisson KKT at 256 x 256 (65,536 elements x 144 contributions):
    bitwise equal to the oracle on two value sets, the same output buffer
    reused, and unchanged on a repeat."""
    import torch

    from amigo_amd import workloads as W
    from amigo_amd.assembly import KKTAssembler, element_values
    E = W.poisson2d_elements(N=256)
    A = KKTAssembler.from_elements(E["n"], E["rowp"], E["cols"], [E["idx"]], E["fixed"])
    r, c = KA.element_contributions(E["idx"])
    rng = np.random.default_rng(5)
    out = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
    for H in (E["H"], E["H"] * (1.0 + rng.standard_normal(E["H"].shape))):
        cv = element_values(H)
        want = KA.assemble(E["n"], E["rowp"], E["cols"], r, c, cv, E["fixed"])
        got = A.values(torch.tensor(cv, device="cuda"), out=out)
        torch.cuda.synchronize()
        assert np.array_equal(got.cpu().numpy(), want)
        again = A.values(torch.tensor(cv, device="cuda"))
        torch.cuda.synchronize()
        assert np.array_equal(again.cpu().numpy(), want)
    np.testing.assert_allclose(A.values(element_values(E["H"])), E["vals"], rtol=0, atol=1e-15)


@pytest.mark.gpu
def test_gpu_assembled_kkt_factors_like_the_snapshot():
    """Assembled on the GPU, handed to the factorization on the GPU without
    a host round trip: same inertia and solution as the snapshot's values."""
    import torch

    from amigo_amd import workloads as W
    from amigo_amd.assembly import KKTAssembler, element_values
    from amigo_amd.ldl import CSRMat, OrderingType, SparseLDL
    N = 64
    E = W.poisson2d_elements(N=N)
    P = W.poisson2d(N=N)
    A = KKTAssembler.from_elements(E["n"], E["rowp"], E["cols"], [E["idx"]], E["fixed"])
    vals = A.values(torch.tensor(element_values(E["H"]), device="cuda"))
    diag = torch.tensor(P.diag, device="cuda")
    xs = []
    inertias = []
    for v in (vals, torch.tensor(P.vals, device="cuda")):
        s = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, v), order=OrderingType.AMD)
        s.factor(diag)
        x = torch.tensor(P.rhs, device="cuda")
        s.solve(x, refine=True)
        inertias.append(tuple(s.get_inertia()))
        xs.append(x.cpu().numpy())
    assert inertias[0] == inertias[1]
    assert inertias[0][1] == len(P.dual_idx)
    assert np.linalg.norm(P.matrix() @ xs[0] - P.rhs) <= 1e-10 * np.linalg.norm(P.rhs)
    np.testing.assert_allclose(xs[0], xs[1], rtol=0, atol=1e-9 * np.abs(xs[1]).max())


@pytest.mark.gpu
def test_gpu_empty_contributions():
    import torch

    from amigo_amd.assembly import KKTAssembler
    rowp = np.array([0, 2, 4], np.int32)
    cols = np.array([0, 1, 0, 1], np.int32)
    A = KKTAssembler(2, rowp, cols, np.zeros(0, np.int32), np.zeros(0, np.int32), [1])
    v = A.values(torch.zeros(0, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    assert v.cpu().tolist() == [0.0, 0.0, 0.0, 1.0]


def test_gpu_flag_consistent():
    assert isinstance(has_gpu(), bool)
