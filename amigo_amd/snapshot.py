"""KKT snapshot wire format: one interior-point linear system on disk.

A snapshot holds what the IPM hands to the linear solver at one iteration,
so that KKT systems from real amigo runs can be replayed against the engine
(and against the oracle) without the model that produced them:

* the KKT matrix K in full symmetric CSR (sorted columns, diagonal present,
  int32 indices, float64 values) -- the reference's CSRMat layout
  (include/csr_matrix.h:101-115, optimization_problem.h:804-917);
* the barrier diagonal Sigma (+ delta_w / -delta_c) that factor() adds to
  the diagonal (amigo/algorithm/solvers/amigo_solver.py:27-40);
* right-hand sides (n x nrhs, column major);
* the variable type of every row, with the reference's OptVarType bit values
  (include/optimization_problem.h:17-23: FIXED 2, PRIMAL 4, SLACK 8,
  DUAL_INEQUALITY 16, DUAL_EQUALITY 32), from which the primal and
  constraint index lists follow as in
  OptimizationProblem::initialize_primal_and_constraints
  (optimization_problem.h:1370-1388);
* the fixed rows (identity rows of FixedVariables, fixed_variables.h:66-69);
* a JSON metadata block (generator, iteration, mu, ...).

Layout (little endian), version 1::

    8  bytes  magic b"AKKTSNAP"
    4  u32    version (1)
    4  u32    flags: 1 diag, 2 rhs, 4 var_types, 8 fixed rows
    8  i64    n
    8  i64    nnz
    8  i64    nrhs
    8  i64    nfixed
    4  u32    metadata length in bytes (m)
    4  u32    0 (reserved)
    (n+1) i32 rowp, nnz i32 cols, nnz f64 vals
    [n f64 diag] [n*nrhs f64 rhs] [n i32 var_types] [nfixed i32 fixed]
    m bytes   UTF-8 JSON metadata
    8  u64    BLAKE2b-64 digest of every byte before it

Reading executes nothing from the file (fixed-layout arrays and JSON); a
wrong magic, an unknown version or flag, a truncated file or a checksum
mismatch raises ValueError.  The reference serializes its models, not its
linear systems (amigo/model.py:1348-1456 serialize / deserialize), so this
format has no counterpart there; ``from_problem`` builds a snapshot from the
reference's own objects (an OptimizationProblem's Hessian CSRMat, the
diagonal and right-hand-side Vectors, the var-type array) without changing
them.

    python -m amigo_amd.snapshot --emit DIR    # C0-C3 snapshots
"""

import hashlib
import json
import struct

import numpy as np

MAGIC = b"AKKTSNAP"
VERSION = 1
F_DIAG, F_RHS, F_TYPES, F_FIXED = 1, 2, 4, 8
_HEADER = struct.Struct("<8sIIqqqqII")

# include/optimization_problem.h:17-23
FIXED, PRIMAL, SLACK, DUAL_INEQUALITY, DUAL_EQUALITY = 2, 4, 8, 16, 32
PRIMAL_MASK = PRIMAL | SLACK | FIXED
DUAL_MASK = DUAL_INEQUALITY | DUAL_EQUALITY


def _digest(chunks):
    """64-bit BLAKE2b digest of a sequence of byte strings."""
    h = hashlib.blake2b(digest_size=8)
    for c in chunks:
        h.update(c)
    return struct.unpack("<Q", h.digest())[0]


class Snapshot:
    """One KKT linear system.  Arrays are numpy; ``rhs`` is (n,) or
    (n, nrhs); ``var_types`` and ``fixed`` may be None."""

    def __init__(self, rowp, cols, vals, diag=None, rhs=None, var_types=None, fixed=None,
                 meta=None):
        self.rowp = np.ascontiguousarray(rowp, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.n = len(self.rowp) - 1
        if self.n < 0 or len(self.cols) != self.rowp[-1] or len(self.vals) != len(self.cols):
            raise ValueError("inconsistent CSR arrays")
        self.diag = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        self.rhs = None if rhs is None else np.asarray(rhs, dtype=np.float64)
        self.var_types = (None if var_types is None
                          else np.ascontiguousarray(var_types, dtype=np.int32))
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.int32)
        self.meta = dict(meta or {})
        for name, a in (("diag", self.diag), ("var_types", self.var_types)):
            if a is not None and a.shape != (self.n,):
                raise ValueError(f"{name} must have length n = {self.n}")
        if self.rhs is not None and self.rhs.shape[0] != self.n:
            raise ValueError("rhs must have n rows")

    @property
    def nnz(self):
        return int(self.rowp[-1])

    def primal_indices(self):
        """Rows of type PRIMAL, SLACK or FIXED (optimization_problem.h:1378-1382)."""
        return np.nonzero(self.var_types & PRIMAL_MASK)[0].astype(np.int32)

    def constraint_indices(self):
        """Rows of type DUAL_INEQUALITY or DUAL_EQUALITY (:1385-1387)."""
        return np.nonzero(self.var_types & DUAL_MASK)[0].astype(np.int32)

    def matrix(self, with_diag=True):
        import scipy.sparse as sp
        K = sp.csr_matrix((self.vals, self.cols, self.rowp), shape=(self.n, self.n))
        if with_diag and self.diag is not None:
            K = K + sp.diags(self.diag)
        return K.tocsr()


def save(path, snap):
    """Write a Snapshot (see the module docstring for the layout)."""
    flags = ((F_DIAG if snap.diag is not None else 0) | (F_RHS if snap.rhs is not None else 0) |
             (F_TYPES if snap.var_types is not None else 0) |
             (F_FIXED if snap.fixed is not None else 0))
    nrhs = 0
    rhs = b""
    if snap.rhs is not None:
        r = snap.rhs.reshape(snap.n, -1)
        nrhs = r.shape[1]
        rhs = np.asfortranarray(r).astype("<f8").tobytes(order="F")
    meta = json.dumps(snap.meta, sort_keys=True).encode("utf-8")
    nfixed = 0 if snap.fixed is None else len(snap.fixed)
    chunks = [
        _HEADER.pack(MAGIC, VERSION, flags, snap.n, snap.nnz, nrhs, nfixed, len(meta), 0),
        snap.rowp.astype("<i4").tobytes(), snap.cols.astype("<i4").tobytes(),
        snap.vals.astype("<f8").tobytes(),
        b"" if snap.diag is None else snap.diag.astype("<f8").tobytes(),
        rhs,
        b"" if snap.var_types is None else snap.var_types.astype("<i4").tobytes(),
        b"" if snap.fixed is None else snap.fixed.astype("<i4").tobytes(),
        meta,
    ]
    with open(path, "wb") as f:
        for c in chunks:
            f.write(c)
        f.write(struct.pack("<Q", _digest(chunks)))


def load(path):
    """Read a Snapshot; raises ValueError on anything that is not a complete
    version-1 snapshot."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < _HEADER.size + 8:
        raise ValueError("truncated snapshot")
    magic, version, flags, n, nnz, nrhs, nfixed, mlen, _ = _HEADER.unpack_from(data, 0)
    if magic != MAGIC:
        raise ValueError("not a KKT snapshot (bad magic)")
    if version != VERSION:
        raise ValueError(f"unsupported snapshot version {version}")
    if flags & ~(F_DIAG | F_RHS | F_TYPES | F_FIXED):
        raise ValueError(f"unknown snapshot flags {flags:#x}")
    if min(n, nnz, nrhs, nfixed) < 0:
        raise ValueError("negative size in snapshot header")
    sizes = [("rowp", "<i4", n + 1), ("cols", "<i4", nnz), ("vals", "<f8", nnz)]
    if flags & F_DIAG:
        sizes.append(("diag", "<f8", n))
    if flags & F_RHS:
        sizes.append(("rhs", "<f8", n * nrhs))
    if flags & F_TYPES:
        sizes.append(("var_types", "<i4", n))
    if flags & F_FIXED:
        sizes.append(("fixed", "<i4", nfixed))
    need = _HEADER.size + sum(np.dtype(t).itemsize * c for _, t, c in sizes) + mlen + 8
    if len(data) != need:
        raise ValueError(f"truncated or oversized snapshot ({len(data)} bytes, {need} expected)")
    (stored,) = struct.unpack_from("<Q", data, need - 8)
    if _digest([data[:need - 8]]) != stored:
        raise ValueError("snapshot checksum mismatch")
    out, off = {}, _HEADER.size
    for name, t, c in sizes:
        a = np.frombuffer(data, dtype=t, count=c, offset=off)
        off += a.nbytes
        out[name] = a.astype(t[1:], copy=True)
    meta = json.loads(data[off:off + mlen].decode("utf-8")) if mlen else {}
    rhs = out.get("rhs")
    if rhs is not None:
        rhs = rhs.reshape(nrhs, n).T.copy() if nrhs > 1 else rhs
    return Snapshot(out["rowp"], out["cols"], out["vals"], out.get("diag"), rhs,
                    out.get("var_types"), out.get("fixed"), meta)


def _array(v):
    if v is None:
        return None
    if hasattr(v, "get_array"):
        return np.array(v.get_array(), copy=True)
    if hasattr(v, "detach"):
        return v.detach().cpu().numpy().copy()
    return np.array(v, copy=True)


def from_problem(hessian, diagonal=None, rhs=None, var_types=None, fixed=None, meta=None):
    """A Snapshot from the reference's objects, as the IPM holds them at a
    factor/solve call: ``hessian`` a CSRMat (get_nonzero_structure,
    get_data; copy_data_device_to_host is called first when present, as
    AmigoSolver.factor does), ``diagonal`` and ``rhs`` Vectors (or arrays),
    ``var_types`` the OptVarType array of the problem (Vector or array),
    ``fixed`` the fixed row indices."""
    fn = getattr(hessian, "copy_data_device_to_host", None)
    if fn is not None:
        fn()
    for v in (diagonal, rhs):
        fn = getattr(v, "copy_device_to_host", None)
        if fn is not None:
            fn()
    nrows, ncols, nnz, rowp, cols = hessian.get_nonzero_structure()
    if nrows != ncols:
        raise ValueError("KKT matrix must be square")
    return Snapshot(np.array(rowp, copy=True), np.array(cols, copy=True),
                    _array(hessian.get_data()), _array(diagonal), _array(rhs),
                    _array(var_types), _array(fixed), meta)


def from_kkt(P, meta=None):
    """A Snapshot of a workload KKTSnapshot (amigo_amd.workloads), typed as
    amigo/model.py:1733-1760 types a model's variables: every row PRIMAL,
    the constraint rows DUAL_EQUALITY (the generators write inequalities
    with a slack, c(x) - s = 0), and rows kept as identity rows (diagonal 1,
    every other entry of the row 0: FixedVariables, fixed_variables.h:66-69)
    FIXED."""
    types = np.full(P.n, PRIMAL, np.int32)
    types[np.asarray(P.dual_idx, dtype=np.int64)] = DUAL_EQUALITY
    rows = np.repeat(np.arange(P.n), np.diff(P.rowp))
    on_diag = P.cols == rows
    offdiag_nz = np.zeros(P.n, bool)
    np.logical_or.at(offdiag_nz, rows[~on_diag], P.vals[~on_diag] != 0.0)
    diag_one = np.zeros(P.n, bool)
    diag_one[rows[on_diag & (P.vals == 1.0)]] = True
    fixed = np.nonzero(diag_one & ~offdiag_nz)[0].astype(np.int32)
    types[fixed] = FIXED
    m = {"name": P.name}
    m.update(meta or {})
    return Snapshot(P.rowp, P.cols, P.vals, P.diag, P.rhs, types, fixed, m)


def _emit(outdir):
    import os

    from . import workloads as W
    os.makedirs(outdir, exist_ok=True)
    for tag, gen in (("c0_rosenbrock", W.rosenbrock), ("c1_cartpole", W.cartpole),
                     ("c2_poisson2d_256", W.poisson2d), ("c3_elasticity3d_40", W.elasticity3d)):
        P = gen()
        path = os.path.join(outdir, tag + ".kkt")
        save(path, from_kkt(P, {"generator": gen.__name__, "config": tag}))
        print(path, P.n, P.nnz, flush=True)


if __name__ == "__main__":
    import sys
    if len(sys.argv) == 3 and sys.argv[1] == "--emit":
        _emit(sys.argv[2])
    else:
        print(__doc__)
