"""Interior-point condensation: the oracle pinned by hand values and by the
algebra of the Newton system, then the HIP kernels (include/kkt_ipm.h)
against it -- bitwise for the vectors, exactly for maxima and the
fraction-to-the-boundary arg-min, to rounding for sums."""

import ctypes
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ipm_condense as O

from .conftest import ROOT


def random_state(n, seed, frac_p=0.7, frac_c=0.2, dup=False):
    """A random iterate: bounds of every kind (none, lower, upper, both),
    primals strictly inside them, positive bound duals where a bound
    exists."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    npr, nc = int(frac_p * n), int(frac_c * n)
    pidx = np.sort(perm[:npr]).astype(np.int32)
    cidx = np.sort(perm[npr:npr + nc]).astype(np.int32)
    if dup and npr > 3:
        pidx[1] = pidx[npr - 1]            # a primal position listed twice
        cidx = np.append(cidx, pidx[2]).astype(np.int32)  # primal and constraint
    kind = rng.integers(0, 4, npr)         # 0 none, 1 lower, 2 upper, 3 both
    lo = rng.uniform(-3, 0, npr)
    hi = lo + rng.uniform(0.5, 3, npr)
    lbx = np.where((kind == 1) | (kind == 3), lo, -np.inf)
    ubx = np.where((kind == 2) | (kind == 3), hi, np.inf)
    x = rng.standard_normal(n)
    t = 0.02 + 0.96 * rng.random(npr)
    xp = np.where(kind == 3, lo + t * (hi - lo),
                  np.where(kind == 1, lo + rng.uniform(1e-3, 2, npr),
                           np.where(kind == 2, hi - rng.uniform(1e-3, 2, npr), x[pidx])))
    x[pidx] = xp  # a position listed twice keeps its last value
    zl = np.where(np.isinf(lbx), 0.0, rng.uniform(1e-3, 5, npr))
    zu = np.where(np.isinf(ubx), 0.0, rng.uniform(1e-3, 5, npr))
    return dict(n=n, pidx=pidx, lbx=lbx, ubx=ubx, cidx=cidx,
                lbh=rng.standard_normal(len(cidx)), mu=0.1, x=x, zl=zl, zu=zu,
                grad=rng.standard_normal(n), dx=rng.standard_normal(n),
                dzl=rng.standard_normal(npr), dzu=rng.standard_normal(npr))


# ---------------------------------------------------------------------------
# The oracle, pinned
# ---------------------------------------------------------------------------
def test_oracle_hand_values():
    """A 5-entry vector worked by hand: primals at 0 (both bounds), 2 (lower),
    3 (none); the constraint at 4; entry 1 untouched."""
    pidx = np.array([0, 2, 3])
    lbx = np.array([0.0, 1.0, -np.inf])
    ubx = np.array([2.0, np.inf, np.inf])
    x = np.array([0.5, 7.0, 3.0, -1.0, 9.0])
    zl = np.array([1.0, 2.0, 0.0])
    zu = np.array([3.0, 0.0, 0.0])
    grad = np.array([1.0, 5.0, 2.0, 4.0, 6.0])
    mu = 0.5
    diag = O.compute_diagonal(5, pidx, lbx, ubx, x, zl, zu)
    # 1/0.5 + 3/1.5 = 4;  2/2 = 1
    assert diag.tolist() == [4.0, 0.0, 1.0, 0.0, 0.0]
    res, nrm = O.compute_residual(5, pidx, lbx, ubx, np.array([4]), np.array([1.0]), mu,
                                  x, zl, zu, grad)
    # entry 0: r = 1 - 1 + 3 + (0.5*1 - 0.5)/0.5 - (1.5*3 - 0.5)/1.5 = 3 - 8/3
    # entry 2: r = 2 - 2 + 0 + (2*2 - 0.5)/2 = 1.75 ; entry 3: r = 4 ; entry 4: -(6 - 1)
    assert res[0] == pytest.approx(-(3.0 - 8.0 / 3.0), abs=1e-15)
    assert res[1:].tolist() == [0.0, -1.75, -4.0, -5.0]
    assert nrm == pytest.approx(np.sqrt((1 / 3) ** 2 + 1.75 ** 2 + 16 + 25), rel=1e-15)
    dzl, dzu = O.compute_bound_dual_step(pidx, lbx, ubx, mu, x, zl, zu,
                                         np.array([0.25, 0, -1.0, 3.0, 0]))
    # dzl0 = -(0 + 1*0.25)/0.5 = -0.5 ; dzu0 = -(4 - 3*0.25)/1.5 ; dzl1 = -(3.5 - 2)/2
    assert dzl.tolist() == [-0.5, -0.75, 0.0]
    assert dzu[0] == pytest.approx(-3.25 / 1.5, abs=1e-15) and dzu[1:].tolist() == [0.0, 0.0]


def test_oracle_residual_closed_form():
    """The z terms cancel: the condensed residual at a primal is
    -grad + mu/gap_l - mu/gap_u (the reference's header comment,
    interior_point_backend.h:27-29); at a constraint -(grad - lbh)."""
    S = random_state(4000, 1)
    res, _ = O.compute_residual(S["n"], S["pidx"], S["lbx"], S["ubx"], S["cidx"], S["lbh"],
                                S["mu"], S["x"], S["zl"], S["zu"], S["grad"])
    xp = S["x"][S["pidx"]]
    with np.errstate(divide="ignore"):
        want = (-S["grad"][S["pidx"]] + np.where(np.isinf(S["lbx"]), 0.0, S["mu"] / (xp - S["lbx"]))
                - np.where(np.isinf(S["ubx"]), 0.0, S["mu"] / (S["ubx"] - xp)))
    with np.errstate(divide="ignore"):
        bar = (np.where(np.isinf(S["lbx"]), 0.0, S["mu"] / (xp - S["lbx"]))
               + np.where(np.isinf(S["ubx"]), 0.0, S["mu"] / (S["ubx"] - xp)))
    scale = np.abs(S["grad"][S["pidx"]]) + np.abs(S["zl"]) + np.abs(S["zu"]) + bar + 1
    assert np.all(np.abs(res[S["pidx"]] - want) <= 1e-14 * scale)
    assert np.array_equal(res[S["cidx"]], -(S["grad"][S["cidx"]] - S["lbh"]))
    other = np.setdiff1d(np.arange(S["n"]), np.concatenate([S["pidx"], S["cidx"]]))
    assert np.all(res[other] == 0.0)


def test_oracle_newton_consistency():
    """Sigma, the residual and the bound-dual back-substitution are three
    faces of one elimination: for any primal step dx,
    -dzl + dzu - Sigma dx = -res - (grad - zl + zu) at the primals, and
    each dz solves its linearized complementarity gap dz + z (+-dx) = -(gap z - mu)."""
    S = random_state(3000, 2)
    n, pidx = S["n"], S["pidx"]
    diag = O.compute_diagonal(n, pidx, S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"])
    res, _ = O.compute_residual(n, pidx, S["lbx"], S["ubx"], S["cidx"], S["lbh"], S["mu"],
                                S["x"], S["zl"], S["zu"], S["grad"])
    dzl, dzu = O.compute_bound_dual_step(pidx, S["lbx"], S["ubx"], S["mu"], S["x"], S["zl"],
                                         S["zu"], S["dx"])
    dx = S["dx"][pidx]
    lhs = -dzl + dzu - diag[pidx] * dx
    rhs = -res[pidx] - (S["grad"][pidx] - S["zl"] + S["zu"])
    scale = np.abs(lhs) + np.abs(rhs) + np.abs(diag[pidx] * dx) + 1
    assert np.all(np.abs(lhs - rhs) <= 1e-12 * scale)
    xp = S["x"][pidx]
    hl, hu = ~np.isinf(S["lbx"]), ~np.isinf(S["ubx"])
    gl, gu = xp - S["lbx"], S["ubx"] - xp
    el = gl * dzl + S["zl"] * dx + (gl * S["zl"] - S["mu"])
    eu = gu * dzu - S["zu"] * dx + (gu * S["zu"] - S["mu"])
    assert np.all(np.abs(el[hl]) <= 1e-12 * (np.abs(gl * S["zl"]) + np.abs(S["zl"] * dx) + 1)[hl])
    assert np.all(np.abs(eu[hu]) <= 1e-12 * (np.abs(gu * S["zu"]) + np.abs(S["zu"] * dx) + 1)[hu])
    assert np.all(dzl[~hl] == 0.0) and np.all(dzu[~hu] == 0.0)


def test_oracle_max_step_is_the_largest_feasible_step():
    """The returned steps keep every gap at >= (1 - tau) of its value, the
    binding term is tight, and a slightly longer step violates it; a serial
    loop over the primals (the reference's) gives the same indices."""
    tau = 0.99
    for seed in range(4):
        S = random_state(500, 10 + seed)
        ax, xi, az, zi = O.compute_max_step(tau, S["pidx"], S["lbx"], S["ubx"], S["x"],
                                            S["zl"], S["zu"], S["dx"], S["dzl"], S["dzu"])
        # serial loop as written in the reference
        bx, bxi, bz, bzi = 1.0, -1, 1.0, -1
        for i, idx in enumerate(S["pidx"]):
            xv, dx = S["x"][idx], S["dx"][idx]
            if not np.isinf(S["lbx"][i]):
                if dx < 0.0 and -tau * (xv - S["lbx"][i]) / dx < bx:
                    bx, bxi = -tau * (xv - S["lbx"][i]) / dx, idx
                if S["dzl"][i] < 0.0 and -tau * S["zl"][i] / S["dzl"][i] < bz:
                    bz, bzi = -tau * S["zl"][i] / S["dzl"][i], idx
            if not np.isinf(S["ubx"][i]):
                if dx > 0.0 and tau * (S["ubx"][i] - xv) / dx < bx:
                    bx, bxi = tau * (S["ubx"][i] - xv) / dx, idx
                if S["dzu"][i] < 0.0 and -tau * S["zu"][i] / S["dzu"][i] < bz:
                    bz, bzi = -tau * S["zu"][i] / S["dzu"][i], idx
        assert (ax, xi, az, zi) == (bx, bxi, bz, bzi)
        xp, dx = S["x"][S["pidx"]], S["dx"][S["pidx"]]
        hl, hu = ~np.isinf(S["lbx"]), ~np.isinf(S["ubx"])
        gl = xp + ax * dx - S["lbx"]
        gu = S["ubx"] - (xp + ax * dx)
        slack = 1e-12 * (1 + np.abs(xp))
        assert np.all(gl[hl] >= (1 - tau) * (xp - S["lbx"])[hl] - slack[hl])
        assert np.all(gu[hu] >= (1 - tau) * (S["ubx"] - xp)[hu] - slack[hu])
        assert np.all((S["zl"] + az * S["dzl"] >= (1 - tau) * S["zl"] - 1e-12)[hl])
        assert np.all((S["zu"] + az * S["dzu"] >= (1 - tau) * S["zu"] - 1e-12)[hu])
        if xi >= 0:
            i = int(np.flatnonzero(S["pidx"] == xi)[0])
            a2 = ax * (1 + 1e-9)
            lo = xp[i] + a2 * dx[i] - S["lbx"][i] < (1 - tau) * (xp[i] - S["lbx"][i])
            hi_ = S["ubx"][i] - (xp[i] + a2 * dx[i]) < (1 - tau) * (S["ubx"][i] - xp[i])
            assert lo or hi_


def test_oracle_complementarity_and_errors():
    S = random_state(2000, 5)
    xp = S["x"][S["pidx"]]
    hl, hu = ~np.isinf(S["lbx"]), ~np.isinf(S["ubx"])
    prods = np.concatenate([((xp - S["lbx"]) * S["zl"])[hl], ((S["ubx"] - xp) * S["zu"])[hu]])
    avg, xi = O.compute_complementarity(S["pidx"], S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"])
    assert avg == pytest.approx(prods.mean(), rel=1e-13)
    assert xi == pytest.approx(min(1.0, max(0.0, prods.min() / prods.mean())), rel=1e-13)
    d, p, c = O.compute_kkt_error(S["pidx"], S["lbx"], S["ubx"], S["cidx"], S["lbh"], S["mu"],
                                  S["x"], S["zl"], S["zu"], S["grad"])
    assert d == np.abs(S["grad"][S["pidx"]] - S["zl"] + S["zu"]).max()
    assert p == np.abs(S["grad"][S["cidx"]] - S["lbh"]).max()
    assert c == np.abs(prods - S["mu"]).max()
    # no bounds at all: average 0, uniformity 1
    assert O.compute_complementarity(np.array([0]), np.array([-np.inf]), np.array([np.inf]),
                                     np.zeros(1), np.zeros(1), np.zeros(1)) == (0.0, 1.0)


def test_oracle_duplicate_positions_take_the_last_value():
    """A position listed twice takes its last primal; a position that is both
    a primal and a constraint takes the constraint (written after)."""
    S = random_state(200, 6, dup=True)
    n, pidx = S["n"], S["pidx"]
    diag = O.compute_diagonal(n, pidx, S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"])
    res, _ = O.compute_residual(n, pidx, S["lbx"], S["ubx"], S["cidx"], S["lbh"], S["mu"],
                                S["x"], S["zl"], S["zu"], S["grad"])
    want_d, want_r = np.zeros(n), np.zeros(n)
    for i, idx in enumerate(pidx):  # the reference's loops, entry by entry
        xv = S["x"][idx]
        s, r = 0.0, S["grad"][idx] - S["zl"][i] + S["zu"][i]
        if not np.isinf(S["lbx"][i]):
            g = xv - S["lbx"][i]
            s += S["zl"][i] / g
            r += (g * S["zl"][i] - S["mu"]) / g
        if not np.isinf(S["ubx"][i]):
            g = S["ubx"][i] - xv
            s += S["zu"][i] / g
            r -= (g * S["zu"][i] - S["mu"]) / g
        want_d[idx], want_r[idx] = s, -r
    for j, idx in enumerate(S["cidx"]):
        want_r[idx] = -(S["grad"][idx] - S["lbh"][j])
    assert np.array_equal(diag, want_d) and np.array_equal(res, want_r)


# ---------------------------------------------------------------------------
# The C ABI without a GPU
# ---------------------------------------------------------------------------
def test_every_ipm_symbol_is_exported():
    from amigo_amd import ipm as I
    from amigo_amd import ldl as L
    lib = L.load_library()
    text = open(os.path.join(ROOT, "include", "kkt_ipm.h")).read()
    decl = sorted(set(re.findall(r"\b(kkt_ipm_[a-z_]+)\s*\(", text)))
    assert len(decl) == 9
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(I.EXPORTS) == decl


def test_create_rejects_bad_indices_before_touching_the_device():
    from amigo_amd import ipm as I
    lib = I._lib()
    h = ctypes.c_void_p()
    pidx = np.array([0, 5], np.int32)
    b = np.zeros(2)
    f64 = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.POINTER(ctypes.c_int)
    st = lib.kkt_ipm_create(4, 2, pidx.ctypes.data_as(i32), b.ctypes.data_as(f64),
                            b.ctypes.data_as(f64), 0, None, None, ctypes.byref(h))
    assert st < 0 and b"out of range" in lib.kkt_ipm_last_error()
    st = lib.kkt_ipm_create(4, 2, None, None, None, 0, None, None, ctypes.byref(h))
    assert st < 0 and b"null" in lib.kkt_ipm_last_error()
    assert lib.kkt_ipm_diagonal(None, None, None, None, None, None) < 0


# ---------------------------------------------------------------------------
# GPU parity through the C ABI
# ---------------------------------------------------------------------------
def _gpu(S):
    import torch
    from amigo_amd.ipm import InteriorPointKernels
    K = InteriorPointKernels(S["n"], S["pidx"], S["lbx"], S["ubx"], S["cidx"], S["lbh"])
    d = {k: torch.tensor(S[k], device="cuda") for k in ("x", "zl", "zu", "grad", "dx", "dzl", "dzu")}
    return K, d


def _check_state(S):
    import torch
    K, d = _gpu(S)
    n, pidx = S["n"], S["pidx"]
    st = (d["x"], d["zl"], d["zu"])
    diag = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    K.compute_diagonal(st, diag)
    want = O.compute_diagonal(n, pidx, S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"])
    assert np.array_equal(diag.cpu().numpy(), want)
    res = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    nrm = K.compute_residual(S["mu"], st, d["grad"], res)
    want_r, want_n = O.compute_residual(n, pidx, S["lbx"], S["ubx"], S["cidx"], S["lbh"],
                                        S["mu"], S["x"], S["zl"], S["zu"], S["grad"])
    assert np.array_equal(res.cpu().numpy(), want_r)
    assert nrm == pytest.approx(want_n, rel=1e-13, abs=0.0)
    px = d["dx"]
    upd = (torch.empty_like(px), torch.empty_like(d["zl"]), torch.empty_like(d["zu"]))
    K.compute_update(S["mu"], st, px, upd)
    wl, wu = O.compute_bound_dual_step(pidx, S["lbx"], S["ubx"], S["mu"], S["x"], S["zl"],
                                       S["zu"], S["dx"])
    assert torch.equal(upd[0], px)
    assert np.array_equal(upd[1].cpu().numpy(), wl) and np.array_equal(upd[2].cpu().numpy(), wu)
    got = K.compute_max_step(0.99, st, (d["dx"], d["dzl"], d["dzu"]))
    assert got == O.compute_max_step(0.99, pidx, S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"],
                                     S["dx"], S["dzl"], S["dzu"])
    avg, xi = K.compute_complementarity(st)
    wa, wx = O.compute_complementarity(pidx, S["lbx"], S["ubx"], S["x"], S["zl"], S["zu"])
    assert avg == pytest.approx(wa, rel=1e-13) and xi == pytest.approx(wx, rel=1e-12)
    assert K.compute_kkt_error(S["mu"], st, d["grad"]) == O.compute_kkt_error(
        pidx, S["lbx"], S["ubx"], S["cidx"], S["lbh"], S["mu"], S["x"], S["zl"], S["zu"],
        S["grad"])


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed,dup", [(1, 0, False), (37, 1, False), (5000, 2, False),
                                        (5000, 3, True), (300001, 4, False)])
def test_gpu_parity_random(n, seed, dup):
    _check_state(random_state(n, seed, dup=dup))


@pytest.mark.gpu
def test_gpu_parity_poisson2d_full_size():
    """The 256 x 256 Poisson control iterate (n = 198,147): Sigma equals the
    workload's barrier diagonal formula, everything matches the oracle."""
    from amigo_amd import workloads as W
    S = W.poisson2d_ipm_state(N=256)
    assert S["n"] == 198147
    _check_state(S)


@pytest.mark.gpu
def test_gpu_max_step_all_steps_feasible():
    """Steps that leave every bound inactive: ax = az = 1, indices -1."""
    import torch
    S = random_state(2000, 7)
    S["dx"][:] = 0.0
    S["dzl"][:] = np.abs(S["dzl"])
    S["dzu"][:] = np.abs(S["dzu"])
    K, d = _gpu(S)
    assert K.compute_max_step(0.99, (d["x"], d["zl"], d["zu"]),
                              (d["dx"], d["dzl"], d["dzu"])) == (1.0, -1, 1.0, -1)


@pytest.mark.gpu
def test_gpu_newton_step_on_the_device():
    """One condensed Newton step of the Poisson control problem without
    leaving the device: Sigma -> factor -> condensed residual -> solve ->
    bound-dual back-substitution -> fraction to the boundary.  The solve
    meets the 1e-10 residual bar on K + Sigma, and the recovered dz satisfy
    their linearized complementarity."""
    import torch
    from amigo_amd import workloads as W
    from amigo_amd.ldl import CSRMat, OrderingType, SparseLDL
    N = 64
    P = W.poisson2d(N=N)
    S = W.poisson2d_ipm_state(N=N)
    K, d = _gpu(S)
    st = (d["x"], d["zl"], d["zu"])
    diag = torch.empty(S["n"], dtype=torch.float64, device="cuda")
    K.compute_diagonal(st, diag)
    diag[torch.tensor(S["pidx"], device="cuda", dtype=torch.long)] += 1e-12
    ldl = SparseLDL(CSRMat(P.n, P.n, P.rowp, P.cols, torch.tensor(P.vals, device="cuda")),
                    order=OrderingType.NESTED_DISSECTION)
    ldl.factor(diag)
    res = torch.empty(S["n"], dtype=torch.float64, device="cuda")
    K.compute_residual(S["mu"], st, d["grad"], res)
    px = res.clone()
    ldl.solve(px)
    A = P.matrix(with_diag=False) + sp.diags(diag.cpu().numpy())
    r = res.cpu().numpy()
    assert np.linalg.norm(A @ px.cpu().numpy() - r) <= 1e-10 * np.linalg.norm(r)
    upd = (torch.empty_like(px), torch.empty_like(d["zl"]), torch.empty_like(d["zu"]))
    K.compute_update(S["mu"], st, px, upd)
    ax, xi, az, zi = K.compute_max_step(0.995, st, upd)
    assert 0.0 < ax <= 1.0 and 0.0 < az <= 1.0
    xp = S["x"][S["pidx"]]
    dx = px.cpu().numpy()[S["pidx"]]
    isu = ~np.isinf(S["lbx"])
    gl = xp - S["lbx"]
    e = gl * upd[1].cpu().numpy() + S["zl"] * dx + (gl * S["zl"] - S["mu"])
    assert np.all(np.abs(e[isu]) <= 1e-12 * (1 + np.abs(S["zl"] * dx) + np.abs(gl * S["zl"]))[isu])
