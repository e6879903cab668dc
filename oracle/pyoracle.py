"""ctypes front end of the oracle -- test infrastructure, not product code.

Loads oracle/_build/libldl_oracle.so (the CPU restatement of the reference's
SparseLDL path) and, when present, oracle/_ref/libamigo_ref.so (the
reference's own SparseLDL compiled from its headers).  Both expose the same
operations; the classes below give them one Python interface so that tests
can compare them field by field.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "_build", "libldl_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libamigo_ref.so")

NATURAL, AMD, GIVEN = 0, 1, 2
LDL, CHOLESKY = 0, 1

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def _ip(a):
    return a.ctypes.data_as(_i32p)


def _dp(a):
    return a.ctypes.data_as(_f64p) if a is not None else None


_libs = {}


def _load(path, prefix):
    if path in _libs:
        return _libs[path]
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    getattr(lib, prefix + "destroy").argtypes = [vp]
    for name in ("num_nodes", "contrib_total"):
        f = getattr(lib, prefix + name)
        f.argtypes = [vp]
        f.restype = ctypes.c_int
    getattr(lib, prefix + "get_perm").argtypes = [vp, _i32p, _i32p]
    getattr(lib, prefix + "get_nodes").argtypes = [vp] + [_i32p] * 6
    f = getattr(lib, prefix + "factor")
    f.argtypes = [vp, _f64p, _f64p]
    f.restype = ctypes.c_int
    getattr(lib, prefix + "solve").argtypes = [vp, ctypes.c_int, _f64p, ctypes.c_int]
    getattr(lib, prefix + "inertia").argtypes = [vp, _i32p, _i32p]
    getattr(lib, prefix + "node_info").argtypes = [vp] + [_i32p] * 4
    if prefix == "orc_":
        lib.orc_analyze.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int]
        lib.orc_analyze.restype = vp
        lib.orc_get_etree.argtypes = [vp, _i32p, _i32p, _i32p]
        lib.orc_factor_nnz.argtypes = [vp]
        lib.orc_factor_nnz.restype = ctypes.c_longlong
        lib.orc_num_snodes.argtypes = [vp]
        lib.orc_num_snodes.restype = ctypes.c_int
    else:
        lib.ref_analyze.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int]
        lib.ref_analyze.restype = vp
        lib.ref_assemble.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p, _i32p,
                                     _i32p, _f64p, ctypes.c_int, _i32p, _f64p]
    _libs[path] = lib
    return lib


def ref_available():
    return os.path.exists(REF_SO)


def ref_assemble(n, rowp, cols, crow, ccol, cvals, fixed=(), ncomp=None):
    """The reference's own host assembly (CSRMat::add_row per call, then
    FixedVariables::zero_rows_and_columns).  Contributions are grouped into
    add_row calls of ncomp consecutive entries sharing a row (one call per
    element column, as the reference makes them); ncomp=None makes one call
    per contribution."""
    lib = _load(REF_SO, "ref_")
    crow = np.ascontiguousarray(crow, np.int32)
    ccol = np.ascontiguousarray(ccol, np.int32)
    cv = np.ascontiguousarray(cvals, np.float64)
    nc = 1 if ncomp is None else int(ncomp)
    assert len(crow) % nc == 0
    ncalls = len(crow) // nc
    call_row = np.ascontiguousarray(crow[::nc])
    assert np.all(crow.reshape(ncalls, nc) == call_row[:, None])
    call_ptr = np.arange(0, len(crow) + 1, nc, dtype=np.int32)
    fx = np.ascontiguousarray(fixed, np.int32)
    out = np.zeros(int(rowp[n]))
    lib.ref_assemble(n, _ip(np.ascontiguousarray(rowp, np.int32)),
                     _ip(np.ascontiguousarray(cols, np.int32)), ncalls, _ip(call_row),
                     _ip(call_ptr), _ip(ccol), _dp(cv), len(fx), _ip(fx), _dp(out))
    return out


class _Base:
    prefix = None

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def __del__(self):
        if getattr(self, "h", None):
            self._f("destroy")(self.h)
            self.h = None

    def perm(self):
        p = np.zeros(self.n, np.int32)
        ip = np.zeros(self.n, np.int32)
        self._f("get_perm")(self.h, _ip(p), _ip(ip))
        return p, ip

    def nodes(self):
        N = self._f("num_nodes")(self.h)
        nc = self._f("contrib_total")(self.h)
        out = dict(
            node_ptr=np.zeros(N + 1, np.int32),
            node_to_var=np.zeros(self.n, np.int32),
            children_ptr=np.zeros(N + 2, np.int32),
            children=np.zeros(N, np.int32),
            contrib_ptr=np.zeros(N + 1, np.int32),
            contrib_rows=np.zeros(max(nc, 1), np.int32),
        )
        self._f("get_nodes")(self.h, *[_ip(out[k]) for k in
                                       ("node_ptr", "node_to_var", "children_ptr",
                                        "children", "contrib_ptr", "contrib_rows")])
        out["contrib_rows"] = out["contrib_rows"][:nc]
        return out

    def factor(self, vals, diag=None):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        d = None if diag is None else np.ascontiguousarray(diag, dtype=np.float64)
        return self._f("factor")(self.h, _dp(vals), _dp(d))

    def solve(self, b):
        """Solve in place semantics on a copy; b is (n,) or (n, nrhs)."""
        x = np.array(b, dtype=np.float64, order="F", copy=True)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        self._f("solve")(self.h, nrhs, _dp(x), self.n)
        return x

    def inertia(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        self._f("inertia")(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def node_info(self):
        N = self._f("num_nodes")(self.h)
        arrs = [np.zeros(N, np.int32) for _ in range(4)]
        self._f("node_info")(self.h, *[_ip(a) for a in arrs])
        return dict(zip(("npiv", "ndel", "n2x2", "root"), arrs))


class OracleLDL(_Base):
    """CPU restatement of SparseLDL (order: NATURAL, AMD or GIVEN perm)."""

    prefix = "orc_"

    def __init__(self, n, rowp, cols, order=AMD, perm=None, ustab=0.01,
                 pivot_tol=1e-14, solver_type=LDL):
        self.lib = _load(ORACLE_SO, "orc_")
        self.n = int(n)
        self.rowp = np.ascontiguousarray(rowp, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        p = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
        self.h = self.lib.orc_analyze(self.n, _ip(self.rowp), _ip(self.cols), order,
                                      _ip(p) if p is not None else None,
                                      ustab, pivot_tol, solver_type)

    def etree(self):
        par = np.zeros(self.n, np.int32)
        post = np.zeros(self.n, np.int32)
        cc = np.zeros(self.n, np.int32)
        self.lib.orc_get_etree(self.h, _ip(par), _ip(post), _ip(cc))
        return par, post, cc

    def factor_nnz(self):
        return self.lib.orc_factor_nnz(self.h)

    def num_snodes(self):
        return self.lib.orc_num_snodes(self.h)


class RefLDL(_Base):
    """The reference's own SparseLDL (compiled by `make -C oracle ref`)."""

    prefix = "ref_"

    def __init__(self, n, rowp, cols, order=AMD, ustab=0.01, pivot_tol=1e-14,
                 delay_growth=2.0, solver_type=LDL):
        self.lib = _load(REF_SO, "ref_")
        self.n = int(n)
        self.rowp = np.ascontiguousarray(rowp, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.h = self.lib.ref_analyze(self.n, _ip(self.rowp), _ip(self.cols), order,
                                      ustab, pivot_tol, delay_growth, solver_type)
