"""Interior-point condensation on the GPU (libkktldl.so, include/kkt_ipm.h).

``InteriorPointKernels`` mirrors the vector methods of the reference's
``InteriorPointOptimizer`` (include/interior_point_optimizer.h:233-433, bound
in amigo/amigo.cpp:770-800) that produce the KKT solve's inputs and consume
its output:

* ``compute_diagonal(vars, diag)``: the barrier diagonal Sigma, the ``diag``
  argument of ``SparseLDL.factor`` (interior_point_backend.h:179-199);
* ``compute_residual(mu, vars, grad, res) -> float``: the condensed
  right-hand side and its 2-norm (interior_point_backend.h:146-176);
* ``compute_update(mu, vars, px, update)``: the bound-dual back-substitution
  after the solve (interior_point_backend.h:213-236);
* ``compute_max_step(tau, vars, update) -> (ax, xi, az, zi)``: the
  fraction-to-the-boundary rule (interior_point_backend.h:242-287);
* ``compute_complementarity(vars) -> (avg, xi)`` and
  ``compute_kkt_error(mu, vars, grad) -> (dual, primal, comp)``.

Vectors are CUDA float64 tensors, so an IPM iteration stays on the device.
``vars`` / ``update`` are objects with ``get_solution()``, ``get_zl()`` and
``get_zu()`` (the reference's ``OptVector`` accessors) or tuples
``(x, zl, zu)``.  The product path has no CPU fallback: without the built
library every call raises.
"""

import ctypes

import numpy as np

from . import ldl as _ldl

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)
_dbl = ctypes.c_double

EXPORTS = ("kkt_ipm_create", "kkt_ipm_diagonal", "kkt_ipm_residual",
           "kkt_ipm_bound_dual_step", "kkt_ipm_max_step", "kkt_ipm_complementarity",
           "kkt_ipm_kkt_error", "kkt_ipm_last_error", "kkt_ipm_destroy")


def _lib():
    lib = _ldl.load_library()
    if not getattr(lib, "_ipm_ready", False):
        lib.kkt_ipm_create.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, _f64p,
                                       ctypes.c_int, _i32p, _f64p, ctypes.POINTER(_vp)]
        lib.kkt_ipm_diagonal.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        lib.kkt_ipm_residual.argtypes = [_vp, _dbl, _vp, _vp, _vp, _vp, _vp, _f64p, _vp]
        lib.kkt_ipm_bound_dual_step.argtypes = [_vp, _dbl, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        lib.kkt_ipm_max_step.argtypes = [_vp, _dbl, _vp, _vp, _vp, _vp, _vp, _vp,
                                         _f64p, _i32p, _f64p, _i32p, _vp]
        lib.kkt_ipm_complementarity.argtypes = [_vp, _vp, _vp, _vp, _f64p, _f64p, _vp]
        lib.kkt_ipm_kkt_error.argtypes = [_vp, _dbl, _vp, _vp, _vp, _vp, _f64p, _f64p,
                                          _f64p, _vp]
        lib.kkt_ipm_last_error.restype = ctypes.c_char_p
        lib.kkt_ipm_destroy.argtypes = [_vp]
        for name in EXPORTS:
            if name not in ("kkt_ipm_last_error", "kkt_ipm_destroy"):
                getattr(lib, name).restype = ctypes.c_int
        lib._ipm_ready = True
    return lib


def _unpack(v):
    if isinstance(v, (tuple, list)):
        return tuple(v)
    return v.get_solution(), v.get_zl(), v.get_zu()


def _ptr(t, n, what):
    if not (hasattr(t, "is_cuda") and t.is_cuda):
        raise TypeError(f"{what}: expected a CUDA tensor")
    import torch
    if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() < n:
        raise TypeError(f"{what}: expected a contiguous float64 tensor of >= {n} entries")
    return _vp(t.data_ptr())


class InteriorPointKernels:
    """The condensation of the reference's IPM Newton system on the GPU.

    n:                  length of the full KKT vector.
    primal_indices:     positions of the primals (design variables and
                        slacks), with bounds lbx, ubx (+-inf = no bound).
    constraint_indices: positions of the multipliers, with targets lbh.
    """

    def __init__(self, n, primal_indices, lbx, ubx, constraint_indices=(), lbh=()):
        self.lib = _lib()
        pi = np.ascontiguousarray(primal_indices, np.int32)
        lb = np.ascontiguousarray(lbx, np.float64)
        ub = np.ascontiguousarray(ubx, np.float64)
        ci = np.ascontiguousarray(constraint_indices, np.int32)
        lh = np.ascontiguousarray(lbh, np.float64)
        if not (len(pi) == len(lb) == len(ub)) or len(ci) != len(lh):
            raise ValueError("index and bound arrays differ in length")
        self.n, self.num_primals, self.num_constraints = int(n), len(pi), len(ci)
        h = _vp()
        st = self.lib.kkt_ipm_create(
            self.n, self.num_primals, pi.ctypes.data_as(_i32p), lb.ctypes.data_as(_f64p),
            ub.ctypes.data_as(_f64p), self.num_constraints, ci.ctypes.data_as(_i32p),
            lh.ctypes.data_as(_f64p), ctypes.byref(h))
        self._check(st, "kkt_ipm_create")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.kkt_ipm_destroy(self.h)
            self.h = None

    def _check(self, st, what):
        if st < 0:
            raise RuntimeError(f"{what} failed ({st}): {self.lib.kkt_ipm_last_error().decode()}")

    def _state(self, vars):
        x, zl, zu = _unpack(vars)
        return (_ptr(x, self.n, "x"), _ptr(zl, self.num_primals, "zl"),
                _ptr(zu, self.num_primals, "zu"))

    def compute_diagonal(self, vars, diag):
        """diag <- Sigma (interior_point_optimizer.h:267-285)."""
        x, zl, zu = self._state(vars)
        self._check(self.lib.kkt_ipm_diagonal(self.h, x, zl, zu, _ptr(diag, self.n, "diag"),
                                              _ldl._current_stream()), "kkt_ipm_diagonal")

    def compute_residual(self, mu, vars, grad, res):
        """res <- condensed right-hand side; returns its 2-norm
        (interior_point_optimizer.h:233-257)."""
        x, zl, zu = self._state(vars)
        nrm = ctypes.c_double()
        self._check(self.lib.kkt_ipm_residual(
            self.h, float(mu), x, zl, zu, _ptr(grad, self.n, "grad"), _ptr(res, self.n, "res"),
            ctypes.byref(nrm), _ldl._current_stream()), "kkt_ipm_residual")
        return nrm.value

    def compute_update(self, mu, vars, px, update):
        """update <- (px, dzl, dzu) (interior_point_optimizer.h:296-324)."""
        x, zl, zu = self._state(vars)
        ux, dzl, dzu = _unpack(update)
        if ux.data_ptr() != px.data_ptr():
            ux.copy_(px)
        self._check(self.lib.kkt_ipm_bound_dual_step(
            self.h, float(mu), x, zl, zu, _ptr(px, self.n, "px"),
            _ptr(dzl, self.num_primals, "dzl"), _ptr(dzu, self.num_primals, "dzu"),
            _ldl._current_stream()), "kkt_ipm_bound_dual_step")

    def compute_max_step(self, tau, vars, update):
        """(ax, xi, az, zi) by the fraction-to-the-boundary rule
        (amigo/amigo.cpp:788-794)."""
        x, zl, zu = self._state(vars)
        dx, dzl, dzu = _unpack(update)
        ax, az = ctypes.c_double(), ctypes.c_double()
        xi, zi = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.kkt_ipm_max_step(
            self.h, float(tau), x, zl, zu, _ptr(dx, self.n, "dx"),
            _ptr(dzl, self.num_primals, "dzl"), _ptr(dzu, self.num_primals, "dzu"),
            ctypes.byref(ax), ctypes.byref(xi), ctypes.byref(az), ctypes.byref(zi),
            _ldl._current_stream()), "kkt_ipm_max_step")
        return ax.value, xi.value, az.value, zi.value

    def compute_complementarity(self, vars):
        """(average complementarity, uniformity xi)
        (interior_point_optimizer.h:399-433)."""
        x, zl, zu = self._state(vars)
        avg, xi = ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.kkt_ipm_complementarity(
            self.h, x, zl, zu, ctypes.byref(avg), ctypes.byref(xi), _ldl._current_stream()),
            "kkt_ipm_complementarity")
        return avg.value, xi.value

    def compute_kkt_error(self, mu, vars, grad):
        """(dual, primal, complementarity) infinity-norm errors
        (interior_point_backend.h:345-369)."""
        x, zl, zu = self._state(vars)
        d, p, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self.lib.kkt_ipm_kkt_error(
            self.h, float(mu), x, zl, zu, _ptr(grad, self.n, "grad"), ctypes.byref(d),
            ctypes.byref(p), ctypes.byref(c), _ldl._current_stream()), "kkt_ipm_kkt_error")
        return d.value, p.value, c.value
