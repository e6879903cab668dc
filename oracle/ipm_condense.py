"""CPU restatement of the reference's interior-point condensation -- test
infrastructure, not product code.  Only tests/ and bench.py's cpu_baseline
leg import it.

Follows include/interior_point_backend.h and the InteriorPointOptimizer
wrappers that zero the output vectors first
(include/interior_point_optimizer.h:233-285, 399-433):

  compute_diagonal         interior_point_backend.h:179-199
  compute_residual         interior_point_backend.h:146-176
  compute_bound_dual_step  interior_point_backend.h:213-236
  compute_max_step         interior_point_backend.h:242-287, called with
                           ax = az = 1, xi = zi = -1 (amigo/amigo.cpp:788-794)
  compute_complementarity  interior_point_backend.h:313-339 and
                           interior_point_optimizer.h:399-433
  compute_kkt_error        interior_point_backend.h:345-369

Every elementwise value is the reference's expression evaluated in the same
order with IEEE double operations (numpy does not contract a*b + c), so the
vectors are bitwise those of the serial loops.  A position listed twice takes
its last value, and constraint positions are written after primal ones, as
the loops leave them.  Sums run in the serial order (np.cumsum is a
sequential scan).
"""

import numpy as np


def _scatter(out, idx, vals):
    idx = np.asarray(idx)
    if len(np.unique(idx)) == len(idx):
        out[idx] = vals
    else:
        for k in range(len(idx)):  # serial loop: later positions overwrite
            out[idx[k]] = vals[k]


def _has(b):
    return ~np.isinf(b)


def compute_diagonal(n, pidx, lbx, ubx, x, zl, zu):
    """Sigma at the primal positions, 0 elsewhere."""
    diag = np.zeros(n)
    xv = x[pidx]
    with np.errstate(divide="ignore", invalid="ignore"):
        sigma = np.zeros(len(pidx))
        sigma = np.where(_has(lbx), sigma + zl / (xv - lbx), sigma)
        sigma = np.where(_has(ubx), sigma + zu / (ubx - xv), sigma)
    _scatter(diag, pidx, sigma)
    return diag


def compute_residual(n, pidx, lbx, ubx, cidx, lbh, mu, x, zl, zu, grad):
    """(res, ||res||_2): the condensed right-hand side."""
    res = np.zeros(n)
    xv = x[pidx]
    with np.errstate(divide="ignore", invalid="ignore"):
        r = grad[pidx] - zl + zu
        gl = xv - lbx
        r = np.where(_has(lbx), r + (gl * zl - mu) / gl, r)
        gu = ubx - xv
        r = np.where(_has(ubx), r - (gu * zu - mu) / gu, r)
    _scatter(res, pidx, -r)
    cidx = np.asarray(cidx, np.int64)
    if len(cidx):
        _scatter(res, cidx, -(grad[cidx] - lbh))
    return res, float(np.sqrt(np.cumsum(res * res)[-1])) if n else 0.0


def compute_bound_dual_step(pidx, lbx, ubx, mu, x, zl, zu, px):
    """(dzl, dzu) from the primal step px."""
    xv, dx = x[pidx], px[pidx]
    with np.errstate(divide="ignore", invalid="ignore"):
        gl = xv - lbx
        dzl = np.where(_has(lbx), -((gl * zl - mu) + zl * dx) / gl, 0.0)
        gu = ubx - xv
        dzu = np.where(_has(ubx), -((gu * zu - mu) - zu * dx) / gu, 0.0)
    return dzl, dzu


def compute_max_step(tau, pidx, lbx, ubx, x, zl, zu, dxv, dzl, dzu):
    """(ax, xi, az, zi): the serial loop keeps the first term strictly below
    the running minimum, starting from 1."""
    pidx = np.asarray(pidx)
    xv, dx = x[pidx], dxv[pidx]
    inf = np.inf
    hl, hu = _has(lbx), _has(ubx)
    with np.errstate(divide="ignore", invalid="ignore"):
        ax_l = np.where(hl & (dx < 0.0), -tau * (xv - lbx) / dx, inf)
        ax_u = np.where(hu & (dx > 0.0), tau * (ubx - xv) / dx, inf)
        az_l = np.where(hl & (dzl < 0.0), -tau * zl / dzl, inf)
        az_u = np.where(hu & (dzu < 0.0), -tau * zu / dzu, inf)
    out = []
    for a, b in ((ax_l, ax_u), (az_l, az_u)):
        per = np.minimum(a, b)
        m = per.min() if len(per) else inf
        if m < 1.0:
            out += [float(m), int(pidx[np.flatnonzero(per == m)[0]])]
        else:
            out += [1.0, -1]
    return tuple(out)


def compute_complementarity(pidx, lbx, ubx, x, zl, zu):
    """(avg, xi): mean complementarity product over the finite bounds and
    the uniformity min/avg clamped to [0, 1]."""
    xv = x[pidx]
    hl, hu = _has(lbx), _has(ubx)
    with np.errstate(invalid="ignore"):
        cl = (xv - lbx) * zl
        cu = (ubx - xv) * zu
    # serial order: for each primal, its lower then its upper product
    prods = np.stack([np.where(hl, cl, np.nan), np.where(hu, cu, np.nan)], 1).ravel()
    prods = prods[~np.isnan(prods)] if len(prods) else prods
    s = float(np.cumsum(prods)[-1]) if len(prods) else 0.0
    cnt = float(len(prods))
    mn = float(prods.min()) if len(prods) else np.finfo(np.float64).max
    avg = s / cnt if cnt > 0.0 else 0.0
    xi = 1.0
    if avg > 0.0:
        xi = max(0.0, min(1.0, mn / avg))
    return avg, xi


def compute_kkt_error(pidx, lbx, ubx, cidx, lbh, mu, x, zl, zu, grad):
    """(dual, primal, comp) infinity-norm errors."""
    xv = x[pidx]
    dual = float(np.max(np.abs(grad[pidx] - zl + zu), initial=0.0))
    with np.errstate(invalid="ignore"):
        cl = np.abs((xv - lbx) * zl - mu)
        cu = np.abs((ubx - xv) * zu - mu)
    comp = float(max(np.max(cl[_has(lbx)], initial=0.0), np.max(cu[_has(ubx)], initial=0.0)))
    cidx = np.asarray(cidx, np.int64)
    primal = float(np.max(np.abs(grad[cidx] - lbh), initial=0.0)) if len(cidx) else 0.0
    return dual, primal, comp
