"""KKT value assembly on the GPU (libkktldl.so, include/kkt_asm.h).

``KKTAssembler`` replaces the value half of the reference's
``OptimizationProblem.hessian`` (include/optimization_problem.h:1063-1087):
zero the CSR values, add every element's Hessian columns into them
(include/component_group.h:491-528, include/cuda/component_group_backend.cuh:
159-189) and replace fixed-variable rows and columns by identity rows
(include/fixed_variables.h:66-69).  The pattern work (the CSR slot of every
contribution, the fixed-row mask) is done once on the host; each
``values()`` call is one kernel that writes every CSR value from the
contributions, with the additions of one entry in contribution order, so the
result is bitwise equal to the reference's serial host assembly.
"""

import ctypes

import numpy as np

from . import ldl as _ldl

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_longlong)

EXPORTS = ("kkt_asm_create", "kkt_asm_values", "kkt_asm_get_info",
           "kkt_asm_last_error", "kkt_asm_destroy")


def _lib():
    lib = _ldl.load_library()
    if not getattr(lib, "_asm_ready", False):
        lib.kkt_asm_create.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_longlong,
                                       _i32p, _i32p, ctypes.c_int, _i32p, ctypes.POINTER(_vp)]
        lib.kkt_asm_create.restype = ctypes.c_int
        lib.kkt_asm_values.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        lib.kkt_asm_values.restype = ctypes.c_int
        lib.kkt_asm_get_info.argtypes = [_vp, _i64p, _i64p, _i64p]
        lib.kkt_asm_get_info.restype = ctypes.c_int
        lib.kkt_asm_last_error.restype = ctypes.c_char_p
        lib.kkt_asm_destroy.argtypes = [_vp]
        lib._asm_ready = True
    return lib


def _check(lib, status, what):
    if status < 0:
        raise RuntimeError(f"{what} failed ({status}): {lib.kkt_asm_last_error().decode()}")


def element_contributions(idx):
    """(row, col) of each contribution of an element group whose variable
    indices are idx (nelem x ncomp), in the reference's order: for element
    e and Hessian column k, the ncomp entries (idx[e][k], idx[e][i]) of one
    add_row call (include/component_group.h:516-523), i.e. contribution
    q = ncomp*(ncomp*e + k) + i."""
    idx = np.asarray(idx, np.int32)
    if idx.ndim != 2:
        raise ValueError("element indices must be (nelem, ncomp)")
    nc = idx.shape[1]
    return np.repeat(idx, nc, axis=1).ravel(), np.tile(idx, (1, nc)).ravel()


def element_values(H):
    """Contribution values of an element group from its Hessian blocks H
    (nelem x ncomp x ncomp): value q = ncomp*(ncomp*e + k) + i is entry i of
    the Hessian-vector product H_e e_k, i.e. H[e, i, k]."""
    H = np.asarray(H, np.float64)
    return np.ascontiguousarray(H.transpose(0, 2, 1)).ravel()


class KKTAssembler:
    """CSR values of the KKT matrix from element contributions, on the GPU.

    n, rowp, cols: the CSR pattern (sorted columns per row).
    crow, ccol:    row and column of every contribution (int32 arrays), or
                   use :meth:`from_elements`.
    fixed:         fixed variables (identity rows/columns); out-of-range
                   entries are ignored, as FixedVariables ignores them.
    Contributions outside the pattern are dropped (counted in ``ndropped``),
    as CSRMat::add_row drops them (include/csr_matrix.h:555-567).
    """

    def __init__(self, n, rowp, cols, crow, ccol, fixed=None):
        self.lib = _lib()
        self.n = int(n)
        rowp = np.ascontiguousarray(rowp, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        crow = np.ascontiguousarray(crow, np.int32)
        ccol = np.ascontiguousarray(ccol, np.int32)
        if len(rowp) != self.n + 1 or len(cols) != rowp[-1]:
            raise ValueError("rowp/cols do not describe an n x n CSR pattern")
        if crow.shape != ccol.shape:
            raise ValueError("crow and ccol differ in length")
        fx = np.ascontiguousarray([] if fixed is None else fixed, np.int32)
        fx = fx[(fx >= 0) & (fx < self.n)]
        self.h = _vp()
        _check(self.lib, self.lib.kkt_asm_create(
            self.n, rowp.ctypes.data_as(_i32p), cols.ctypes.data_as(_i32p), len(crow),
            crow.ctypes.data_as(_i32p), ccol.ctypes.data_as(_i32p), len(fx),
            fx.ctypes.data_as(_i32p), ctypes.byref(self.h)), "kkt_asm_create")
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self.lib.kkt_asm_get_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.nnz, self.ncontrib, self.ndropped = a.value, b.value, c.value

    @classmethod
    def from_elements(cls, n, rowp, cols, groups, fixed=None):
        """groups: element index arrays (nelem_g x ncomp_g), one per component
        group, in the order the reference loops over its components; the
        contribution values of :meth:`values` follow the same order (see
        :func:`element_values`)."""
        rc = [element_contributions(g) for g in groups]
        crow = np.concatenate([r for r, _ in rc]) if rc else np.zeros(0, np.int32)
        ccol = np.concatenate([c for _, c in rc]) if rc else np.zeros(0, np.int32)
        return cls(n, rowp, cols, crow, ccol, fixed)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.kkt_asm_destroy(self.h)
            self.h = None

    def values(self, contrib, out=None):
        """CSR values (nnz) from contribution values (ncontrib).

        A CUDA float64 tensor in gives a CUDA tensor out (written on torch's
        current stream, nothing copied through the host); a numpy array in
        gives a numpy array out."""
        if _ldl._is_torch(contrib) and contrib.is_cuda:
            import torch
            cv = contrib.to(torch.float64).contiguous()
            if cv.numel() != self.ncontrib:
                raise ValueError(f"expected {self.ncontrib} contributions, got {cv.numel()}")
            if out is None:
                out = torch.empty(self.nnz, dtype=torch.float64, device=cv.device)
            if not (out.is_cuda and out.dtype == torch.float64 and out.is_contiguous()
                    and out.numel() == self.nnz):
                raise ValueError("out must be a contiguous CUDA float64 tensor of nnz entries")
            _check(self.lib, self.lib.kkt_asm_values(
                self.h, _vp(cv.data_ptr()), _vp(out.data_ptr()), _ldl.DEVICE,
                _ldl._current_stream()), "kkt_asm_values")
            return out
        cv = np.ascontiguousarray(contrib, np.float64)
        if cv.size != self.ncontrib:
            raise ValueError(f"expected {self.ncontrib} contributions, got {cv.size}")
        res = np.empty(self.nnz) if out is None else out
        if res.dtype != np.float64 or not res.flags.c_contiguous or res.size != self.nnz:
            raise ValueError("out must be a contiguous float64 array of nnz entries")
        _check(self.lib, self.lib.kkt_asm_values(
            self.h, _vp(cv.ctypes.data), _vp(res.ctypes.data), _ldl.HOST, None),
            "kkt_asm_values")
        return res
