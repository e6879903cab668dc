"""Python interface of the MI355X KKT LDL^T engine (libkktldl.so, C ABI).

Mirrors the reference's pybind surface for this path (amigo/amigo.cpp:316-429,
648-732): ``OrderingType``, ``SolverType``, ``CSRMat(nrows, ncols, rowp, cols,
vals)``, ``Vector``, and ``SparseLDL(mat, solver_type, ustab, pivot_tol,
delay_growth, order)`` with ``factor(diagonal)``, ``solve(x)`` and
``get_inertia()``.  Buffers may be numpy arrays / ``Vector`` objects (host)
or torch CUDA tensors (device-resident; the engine then reads and writes
HBM directly on torch's current stream).

There is no CPU fallback: if the HIP library is missing or no GPU is
present, construction of ``SparseLDL`` raises.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkktldl.so")


class OrderingType:
    """Orderings with the reference's numbering (include/ordering_utils.h:21,
    bound at amigo/amigo.cpp:324-334), so that the reference's own enum values
    (``int(amigo.OrderingType.X)``) select the same ordering here.  DEFAULT is
    AMD, as in a reference build without METIS; NESTED_DISSECTION is this
    engine's own nested dissection (METIS is not available).  MULTI_COLOR is
    rejected (the reference reads an uninitialised permutation for it).
    GIVEN (a caller's permutation) is this interface's addition."""

    NESTED_DISSECTION = 0
    AMD = 1
    MULTI_COLOR = 2
    NATURAL = 3
    GIVEN = 4
    DEFAULT = 1


_ORDER_NAMES = {"NESTED_DISSECTION": OrderingType.NESTED_DISSECTION,
                "AMD": OrderingType.AMD, "MULTI_COLOR": OrderingType.MULTI_COLOR,
                "NATURAL": OrderingType.NATURAL, "DEFAULT": OrderingType.DEFAULT,
                "GIVEN": OrderingType.GIVEN}


def ordering_code(order):
    """ABI code of an ordering given as this module's int, the reference's
    pybind enum (matched by name: its ``.name`` attribute or its ``str``
    'OrderingType.X'), or a name string.  Raises ValueError for unknown
    values and for MULTI_COLOR."""
    name = getattr(order, "name", None)
    if name is None and not isinstance(order, (int, np.integer)):
        name = str(order)
    if name is not None and name.startswith("OrderingType."):
        name = name.split(".", 1)[1]
    if name is not None:
        if name not in _ORDER_NAMES:
            raise ValueError(f"unknown ordering {order!r}")
        code = _ORDER_NAMES[name]
    else:
        code = int(order)
        if code not in _ORDER_NAMES.values():
            raise ValueError(f"unknown ordering {order!r}")
    if code == OrderingType.MULTI_COLOR:
        raise ValueError("MULTI_COLOR ordering is not supported")
    return code


class SolverType:
    LDL = 0
    CHOLESKY = 1


FP64 = 0
FP32_REFINE = 1
HOST = 0
DEVICE = 1
WARN_SINGULAR = 1
WARN_NOT_CONVERGED = 2
WARN_NONFINITE = 3


class Options(ctypes.Structure):
    _fields_ = [
        ("order", ctypes.c_int),
        ("solver_type", ctypes.c_int),
        ("ustab", ctypes.c_double),
        ("pivot_tol", ctypes.c_double),
        ("delay_growth", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("refine_tol", ctypes.c_double),
        ("max_refine", ctypes.c_int),
        ("perm", ctypes.POINTER(ctypes.c_int)),
        ("vals_hint", ctypes.POINTER(ctypes.c_double)),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("num_nodes", ctypes.c_int),
        ("num_snodes", ctypes.c_int),
        ("num_levels", ctypes.c_int),
        ("max_front", ctypes.c_int),
        ("nnz_l", ctypes.c_longlong),
        ("flops_factor", ctypes.c_double),
        ("stored_factor", ctypes.c_longlong),
        ("num_delayed", ctypes.c_int),
        ("num_2x2", ctypes.c_int),
        ("replans", ctypes.c_int),
        ("refine_steps", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("t_analyze_ms", ctypes.c_double),
        ("t_factor_ms", ctypes.c_double),
        ("t_solve_ms", ctypes.c_double),
        ("front_bytes", ctypes.c_longlong),
        ("num_zero_pivots", ctypes.c_int),
        ("num_nonfinite_pivots", ctypes.c_int),
        ("crit_path_pivots", ctypes.c_longlong),
        ("level_panel_sum", ctypes.c_longlong),
        ("panel_launches", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every symbol include/kkt_ldl.h declares.
EXPORTS = (
    "kkt_ldl_default_options", "kkt_ldl_create", "kkt_ldl_factor",
    "kkt_ldl_refactor", "kkt_ldl_set_pivot_threshold",
    "kkt_ldl_solve", "kkt_ldl_solve_refined", "kkt_ldl_solve_refined_noalloc",
    "kkt_ldl_inertia",
    "kkt_ldl_get_perm", "kkt_ldl_get_sizes", "kkt_ldl_get_supernodes",
    "kkt_ldl_get_node_info", "kkt_ldl_get_stats", "kkt_ldl_last_times",
    "kkt_ldl_set_profiling", "kkt_ldl_get_kernel_times",
    "kkt_ldl_last_error", "kkt_ldl_destroy",
)

KERNEL_CLASSES = ("prepare", "assemble", "front_lds", "factor_small", "panel",
                  "update", "forward", "backward", "residual", "panel_inverse")

_lib = None
_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def load_library(path=LIB_PATH):
    """Load libkktldl.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} not built; run `python -m amigo_amd.build`")
    # torch ships its own HIP runtime; loading it first makes the library
    # bind to that same runtime instead of a second copy from /opt/rocm
    # (two runtimes in one process do not see the device consistently)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    lib.kkt_ldl_default_options.argtypes = [ctypes.POINTER(Options)]
    lib.kkt_ldl_create.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.POINTER(Options),
                                   ctypes.POINTER(_vp)]
    lib.kkt_ldl_factor.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
    lib.kkt_ldl_refactor.argtypes = [_vp, _vp, ctypes.c_int, _vp]
    lib.kkt_ldl_set_pivot_threshold.argtypes = [_vp, ctypes.c_double, ctypes.c_double]
    lib.kkt_ldl_solve.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp]
    lib.kkt_ldl_solve_refined.argtypes = lib.kkt_ldl_solve.argtypes
    lib.kkt_ldl_solve_refined_noalloc.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp]
    lib.kkt_ldl_inertia.argtypes = [_vp, _i32p, _i32p, _i32p]
    lib.kkt_ldl_get_perm.argtypes = [_vp, _i32p, _i32p]
    lib.kkt_ldl_get_sizes.argtypes = [_vp, _i32p, _i32p]
    lib.kkt_ldl_get_supernodes.argtypes = [_vp] + [_i32p] * 6
    lib.kkt_ldl_get_node_info.argtypes = [_vp] + [_i32p] * 4
    lib.kkt_ldl_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
    lib.kkt_ldl_last_times.argtypes = [_vp, _f64p, _f64p]
    lib.kkt_ldl_set_profiling.argtypes = [_vp, ctypes.c_int]
    lib.kkt_ldl_get_kernel_times.argtypes = [_vp, _f64p, _i32p, ctypes.c_int]
    lib.kkt_ldl_last_error.restype = ctypes.c_char_p
    lib.kkt_ldl_destroy.argtypes = [_vp]
    for name in EXPORTS:
        if name not in ("kkt_ldl_default_options", "kkt_ldl_last_error",
                        "kkt_ldl_destroy"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(status, what):
    if status < 0:
        msg = _lib.kkt_ldl_last_error().decode()
        raise RuntimeError(f"{what} failed ({status}): {msg}")
    return status


def _ip(a):
    return a.ctypes.data_as(_i32p)


# ---------------------------------------------------------------------------
# Minimal CSRMat / Vector with the reference's method names.
# ---------------------------------------------------------------------------
class CSRMat:
    """n x n CSR, full symmetric pattern, sorted columns, diagonal present
    (include/csr_matrix.h:101-115).  ``vals`` may be a numpy array or a
    torch CUDA tensor (device-resident values)."""

    def __init__(self, nrows, ncols, rowp, cols, vals):
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.rowp = np.ascontiguousarray(rowp, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        if len(self.rowp) != self.nrows + 1:
            raise RuntimeError("rowp must have length nrows + 1")
        if _is_torch(vals):
            self.data = vals
        else:
            self.data = np.ascontiguousarray(vals, dtype=np.float64).copy()
        if len(self.cols) != len(self.data):
            raise RuntimeError("cols and vals must have the same length")

    def get_nonzero_structure(self):
        return self.nrows, self.ncols, len(self.cols), self.rowp.copy(), self.cols.copy()

    def get_data(self):
        return self.data

    def copy_data_device_to_host(self):
        pass


class Vector:
    """Host vector with the reference's accessors (include/vector.h:56-372)."""

    def __init__(self, size):
        self.array = np.zeros(int(size))

    def get_array(self):
        return self.array

    def get_size(self):
        return len(self.array)

    def __getitem__(self, k):
        return self.array[k]

    def __setitem__(self, k, v):
        self.array[k] = v

    def __len__(self):
        return len(self.array)

    def copy(self, other):
        self.array[:] = _host_array(other)

    def zero(self):
        self.array[:] = 0.0

    def add_scalar_at(self, indices, value):
        """array[indices[i]] += value, one index at a time as the reference
        does (include/vector.h:253-265), so repeated indices add repeatedly."""
        np.add.at(self.array, np.asarray(_host_array(indices), dtype=np.intp), value)

    def copy_device_to_host(self):
        pass

    def copy_host_to_device(self):
        pass


def _is_torch(x):
    return hasattr(x, "data_ptr") and hasattr(x, "is_cuda")


def _host_array(x):
    if hasattr(x, "get_array"):
        return x.get_array()
    return np.asarray(x)


def _current_stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class SparseLDL:
    """Supernodal LDL^T on the GPU behind the reference's SparseLDL
    interface (include/sparse_ldl.h:65-213, amigo/amigo.cpp:648-732).

    Extra keyword arguments beyond the reference's: ``precision``
    (FP64 or FP32_REFINE), ``refine_tol``, ``max_refine`` and ``perm`` (for
    ``order=OrderingType.GIVEN``).
    """

    def __init__(self, mat, solver_type=SolverType.LDL, ustab=0.01, pivot_tol=1e-14,
                 delay_growth=2.0, order=OrderingType.NATURAL, precision=FP64,
                 refine_tol=1e-10, max_refine=10, perm=None):
        order = ordering_code(order)
        self.lib = load_library()
        self.mat = mat
        nrows, ncols, nnz, rowp, cols = mat.get_nonzero_structure()
        if nrows != ncols:
            raise ValueError("matrix must be square")
        self.n = int(nrows)
        self.nnz = int(nnz)
        self._rowp = np.ascontiguousarray(rowp, dtype=np.int32)
        self._cols = np.ascontiguousarray(cols, dtype=np.int32)
        opts = Options()
        self.lib.kkt_ldl_default_options(ctypes.byref(opts))
        opts.order = order
        opts.solver_type = int(solver_type)
        opts.ustab = float(ustab)
        opts.pivot_tol = float(pivot_tol)
        opts.delay_growth = float(delay_growth)
        opts.precision = int(precision)
        opts.refine_tol = float(refine_tol)
        opts.max_refine = int(max_refine)
        self._perm = None
        self._hint = None
        if order == OrderingType.NESTED_DISSECTION and perm is None:
            # the matrix values at construction order the columns inside
            # each supernode (predicted 2x2 partners adjacent)
            vals = mat.get_data()
            if _is_torch(vals):
                vals = vals.detach().cpu().numpy()
            self._hint = np.ascontiguousarray(vals, dtype=np.float64)
            opts.vals_hint = self._hint.ctypes.data_as(_f64p)
        if perm is not None:
            self._perm = np.ascontiguousarray(perm, dtype=np.int32)
            opts.perm = _ip(self._perm)
            opts.order = OrderingType.GIVEN
        self._opts = opts
        h = _vp()
        _check(self.lib.kkt_ldl_create(self.n, _ip(self._rowp), _ip(self._cols),
                                       ctypes.byref(opts), ctypes.byref(h)),
               "kkt_ldl_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "lib", None):
            self.lib.kkt_ldl_destroy(self.h)
            self.h = None

    # -- reference interface -------------------------------------------------
    def factor(self, diagonal=None):
        """Numeric factorization of the matrix's current values plus
        ``diagonal``.  Returns 0, or a warning: WARN_SINGULAR (zero pivots,
        counted positive in the inertia as the reference does) or
        WARN_NONFINITE (NaN / infinite pivots).  Raises RuntimeError on
        errors."""
        vals = self.mat.get_data()
        d = None
        if diagonal is not None:
            d = diagonal.get_array() if hasattr(diagonal, "get_array") else diagonal
        if _is_torch(vals) or _is_torch(d):
            import torch
            if not _is_torch(vals):
                vals = torch.as_tensor(np.asarray(vals), device="cuda")
            if d is not None and not _is_torch(d):
                d = torch.as_tensor(np.asarray(d), device="cuda")
            vals = vals.to(torch.float64).contiguous()
            d = None if d is None else d.to(torch.float64).contiguous()
            self._keep = (vals, d)
            st = self.lib.kkt_ldl_factor(self.h, _vp(vals.data_ptr()),
                                         _vp(d.data_ptr()) if d is not None else None,
                                         DEVICE, _current_stream())
        else:
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            d = None if d is None else np.ascontiguousarray(d, dtype=np.float64)
            st = self.lib.kkt_ldl_factor(self.h, vals.ctypes.data_as(_vp),
                                         d.ctypes.data_as(_vp) if d is not None else None,
                                         HOST, None)
        return _check(st, "kkt_ldl_factor")

    def refactor(self, diagonal=None):
        """Refactorization with a new diagonal and the matrix values of the
        last ``factor`` call (kept on the device; they are not re-read).  For
        inertia-correction retries, which change only the diagonal."""
        d = None
        if diagonal is not None:
            d = diagonal.get_array() if hasattr(diagonal, "get_array") else diagonal
        if _is_torch(d):
            import torch
            d = d.to(torch.float64).contiguous()
            self._keep_diag = d
            st = self.lib.kkt_ldl_refactor(self.h, _vp(d.data_ptr()), DEVICE,
                                           _current_stream())
        else:
            d = None if d is None else np.ascontiguousarray(d, dtype=np.float64)
            st = self.lib.kkt_ldl_refactor(self.h, d.ctypes.data_as(_vp) if d is not None
                                           else None, HOST, None)
        return _check(st, "kkt_ldl_refactor")

    def set_pivot_threshold(self, ustab=-1.0, pivot_tol=-1.0):
        """Threshold-pivot stability factor and absolute pivot tolerance for
        the next factorizations (negative: unchanged)."""
        _check(self.lib.kkt_ldl_set_pivot_threshold(self.h, float(ustab), float(pivot_tol)),
               "kkt_ldl_set_pivot_threshold")

    def inertia_and_zero(self):
        """(npos, nneg, nzero): nzero counts the zero pivots, which are
        included in npos."""
        a, b, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self.lib.kkt_ldl_inertia(self.h, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(z)), "kkt_ldl_inertia")
        return a.value, b.value, z.value

    def solve(self, x, refine=False):
        """x <- K^{-1} x in place: a Vector, a float64 ndarray (1-D, or 2-D
        n x nrhs; Fortran order is solved in place, C order through a
        Fortran copy, as amigo/amigo.cpp:665-727) or a torch CUDA tensor."""
        fn = self.lib.kkt_ldl_solve_refined if refine else self.lib.kkt_ldl_solve
        if hasattr(x, "get_array"):
            arr = x.get_array()
            return self._solve_host(fn, arr)
        if _is_torch(x):
            if x.dtype != __import__("torch").float64:
                raise TypeError("tensor must be float64")
            nrhs = 1 if x.dim() == 1 else x.shape[1]
            if x.dim() == 2 and not (x.stride(0) == 1 and x.stride(1) == self.n):
                y = x.t().contiguous().t()  # column-major copy
                st = fn(self.h, nrhs, _vp(y.data_ptr()), self.n, DEVICE, _current_stream())
                x.copy_(y)
            else:
                st = fn(self.h, nrhs, _vp(x.data_ptr()), self.n, DEVICE, _current_stream())
            return _check(st, "kkt_ldl_solve")
        if isinstance(x, np.ndarray):
            if x.dtype != np.float64:
                raise TypeError("NumPy array must have dtype=float64")
            if not x.flags.writeable:
                raise TypeError("NumPy array must be writeable")
            return self._solve_host(fn, x)
        raise RuntimeError("Expected either a NumPy array, a Vector or a CUDA tensor")

    def _solve_host(self, fn, arr):
        if arr.ndim not in (1, 2):
            raise TypeError("NumPy array must be 1D or 2D")
        nrhs = 1 if arr.ndim == 1 else arr.shape[1]
        if arr.ndim == 1 or arr.flags.f_contiguous:
            st = fn(self.h, nrhs, arr.ctypes.data_as(_vp), self.n, HOST, None)
        else:
            f = np.asfortranarray(arr)
            st = fn(self.h, nrhs, f.ctypes.data_as(_vp), self.n, HOST, None)
            np.copyto(arr, f)
        return _check(st, "kkt_ldl_solve")

    def get_inertia(self):
        a, b, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self.lib.kkt_ldl_inertia(self.h, ctypes.byref(a), ctypes.byref(b),
                                        ctypes.byref(z)), "kkt_ldl_inertia")
        return a.value, b.value

    # -- introspection ---------------------------------------------------------
    def get_perm(self):
        p = np.zeros(self.n, np.int32)
        ip = np.zeros(self.n, np.int32)
        _check(self.lib.kkt_ldl_get_perm(self.h, _ip(p), _ip(ip)), "get_perm")
        return p, ip

    def get_supernodes(self):
        N, nc = ctypes.c_int(), ctypes.c_int()
        _check(self.lib.kkt_ldl_get_sizes(self.h, ctypes.byref(N), ctypes.byref(nc)),
               "get_sizes")
        N, nc = N.value, nc.value
        out = dict(
            node_ptr=np.zeros(N + 1, np.int32),
            node_to_var=np.zeros(self.n, np.int32),
            children_ptr=np.zeros(N + 2, np.int32),
            children=np.zeros(N, np.int32),
            contrib_ptr=np.zeros(N + 1, np.int32),
            contrib_rows=np.zeros(max(nc, 1), np.int32),
        )
        _check(self.lib.kkt_ldl_get_supernodes(
            self.h, *[_ip(out[k]) for k in ("node_ptr", "node_to_var", "children_ptr",
                                            "children", "contrib_ptr", "contrib_rows")]),
            "get_supernodes")
        out["contrib_rows"] = out["contrib_rows"][:nc]
        return out

    def node_info(self):
        N = ctypes.c_int()
        self.lib.kkt_ldl_get_sizes(self.h, ctypes.byref(N), None)
        arrs = [np.zeros(N.value, np.int32) for _ in range(4)]
        _check(self.lib.kkt_ldl_get_node_info(self.h, *[_ip(a) for a in arrs]),
               "get_node_info")
        return dict(zip(("npiv", "ndel", "n2x2", "root"), arrs))

    def stats(self):
        s = Stats()
        _check(self.lib.kkt_ldl_get_stats(self.h, ctypes.byref(s)), "get_stats")
        return s.as_dict()

    def set_profiling(self, on=True):
        _check(self.lib.kkt_ldl_set_profiling(self.h, int(bool(on))), "set_profiling")

    def kernel_times(self):
        """{class: (device ms, launches)} of the last factor or solve."""
        nc = len(KERNEL_CLASSES)
        ms = (ctypes.c_double * nc)()
        nl = (ctypes.c_int * nc)()
        _check(self.lib.kkt_ldl_get_kernel_times(self.h, ms, nl, nc), "get_kernel_times")
        return {k: (ms[i], nl[i]) for i, k in enumerate(KERNEL_CLASSES)}

    def last_times(self):
        f, s = ctypes.c_double(), ctypes.c_double()
        self.lib.kkt_ldl_last_times(self.h, ctypes.byref(f), ctypes.byref(s))
        return f.value, s.value
