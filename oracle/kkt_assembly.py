"""CPU restatement of the reference's KKT value assembly -- test
infrastructure, not product code.  Only tests/ and bench.py's cpu_baseline
leg import it.

Follows OptimizationProblem::hessian (include/optimization_problem.h:
1063-1087):
  1. matrix->zero()                         (include/csr_matrix.h:215-220)
  2. per element e and column k, one CSRMat::add_row of the Hessian-vector
     product H_e[:, k] into row idx[e][k], columns idx[e][:]
     (include/component_group.h:491-528); add_row binary-searches each column
     in the row and silently skips columns outside the pattern
     (include/csr_matrix.h:555-567)
  3. FixedVariables::zero_rows_and_columns: every entry in the row or column
     of a fixed variable becomes 0 and the diagonal entry of a fixed row 1;
     fixed indices outside [0, n) are ignored (include/fixed_variables.h:
     14-48, 66-69, 88-130).
The additions of one entry happen in contribution order starting from 0.0,
so the result is bitwise defined.
"""

import numpy as np


def element_contributions(idx):
    """(row, col) of each contribution of an element group with indices
    idx (nelem x ncomp), in the reference's order q = ncomp*(ncomp*e + k) + i
    for the entry (idx[e][k], idx[e][i])."""
    idx = np.asarray(idx, np.int64)
    ne, nc = idx.shape
    row = np.repeat(idx, nc, axis=1).ravel()     # idx[e][k] repeated over i
    col = np.tile(idx, (1, nc)).ravel()          # idx[e][i]
    return row.astype(np.int32), col.astype(np.int32)


def contribution_slots(n, rowp, cols, crow, ccol):
    """CSR position of each contribution, -1 where add_row would skip it."""
    rowp = np.asarray(rowp, np.int64)
    cols = np.asarray(cols, np.int64)
    crow = np.asarray(crow, np.int64)
    ccol = np.asarray(ccol, np.int64)
    # sorted columns inside each row: search the key row*n + col in the
    # row-major key array of the pattern
    key = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowp)) * n + cols
    q = crow * n + ccol
    pos = np.searchsorted(key, q)
    ok = pos < len(key)
    ok[ok] = key[pos[ok]] == q[ok]
    return np.where(ok, pos, -1)


def assemble(n, rowp, cols, crow, ccol, cvals, fixed=()):
    """CSR values from contributions cvals at (crow, ccol); fixed rows and
    columns replaced by identity rows."""
    rowp = np.asarray(rowp)
    cols = np.asarray(cols)
    slot = contribution_slots(n, rowp, cols, crow, ccol)
    vals = np.zeros(int(rowp[n]))
    keep = slot >= 0
    np.add.at(vals, slot[keep], np.asarray(cvals, np.float64)[keep])   # in order
    fx = np.asarray(fixed, np.int64)
    fx = fx[(fx >= 0) & (fx < n)]
    if len(fx):
        isfix = np.zeros(n, bool)
        isfix[fx] = True
        r = np.repeat(np.arange(n), np.diff(rowp))
        hit = isfix[r] | isfix[cols]
        vals[hit] = 0.0
        vals[isfix[r] & (r == cols)] = 1.0
    return vals
