"""Drop-in linear solver for the reference's interior-point optimizer.

``MI355XSolver`` has the interface of the reference's ``AmigoSolver``
(amigo/algorithm/solvers/amigo_solver.py:11-64) over the ``LinearSolver``
plugin API (amigo/algorithm/solvers/linear_solver.py:5-35): the IPM calls
``factor(hessian, diagonal)``, ``solve(b, x)``, ``inertia_enabled()``,
``get_inertia()`` and ``set_pivot_tolerance()``.  It is injected without
editing the reference by passing an instance as ``options["solver"]``,
which ``make_solver`` returns as is (amigo/algorithm/solvers/__init__.py:18-19).

When the reference package is importable, the class derives from its
``LinearSolver`` so that ``make_solver``'s isinstance check accepts it;
otherwise from the local restatement below.
"""

from abc import ABC, abstractmethod

import numpy as np

from .ldl import (FP32_REFINE, FP64, WARN_NONFINITE, WARN_NOT_CONVERGED, OrderingType,
                  SolverType, SparseLDL)

try:  # the reference's ABC, when amigo is installed
    from amigo.algorithm.solvers.linear_solver import LinearSolver  # type: ignore
except Exception:  # pragma: no cover - the reference is not installed here

    class LinearSolver(ABC):
        """Restatement of amigo/algorithm/solvers/linear_solver.py:5-35."""

        @abstractmethod
        def factor(self, hess, diag):
            return

        @abstractmethod
        def solve(self, b, x):
            return

        def inertia_enabled(self):
            return False

        def get_inertia(self):
            return None

        def set_pivot_tolerance(self, pivtol):
            pass

        @staticmethod
        def find_diag_indices(rowp, cols, nrows):
            diag_idx = np.empty(nrows, dtype=np.intp)
            for i in range(nrows):
                start, end = rowp[i], rowp[i + 1]
                pos = np.searchsorted(cols[start:end], i)
                if pos < end - start and cols[start + pos] == i:
                    diag_idx[i] = start + pos
                else:
                    diag_idx[i] = start
            return diag_idx


class MI355XSolver(LinearSolver):
    """GPU supernodal LDL^T with inertia, as AmigoSolver but on the MI355X.

    Makes the same host/device synchronisation calls as AmigoSolver
    (amigo_solver.py:31-33, 47-54), which are no-ops for host-only vectors.

    ``options`` keys read (all optional, beyond the IPM's own):
    ``"ldl_precision"``: "fp64" (default) or "fp32" (FP32 factor + FP64
    refinement); ``"ldl_refine_tol"`` (1e-10); ``"ldl_order"``
    (OrderingType.DEFAULT, or the reference's enum); ``"ldl_use_pivtol"``
    (False): when True, ``set_pivot_tolerance(pivtol)`` sets the
    threshold-pivot stability factor, the role MUMPS gives it
    (mumps_solver.py:323-325); by default the call is ignored, as
    AmigoSolver ignores it (amigo_solver.py:63-64).

    Failures are not silent: a factorization with NaN or infinite pivots
    raises RuntimeError; in "fp32" mode a refined solve that misses the
    residual target is redone with an FP64 factorization of the same matrix
    (same ordering), counted in ``fallbacks``.
    """

    solver_name = "MI355XSolver"

    def __init__(self, options, state):
        self.options = options if options is not None else {}
        self.hessian = state.hessian
        ustab = 0.01
        pivot_tol = 1e-14
        self.precision = self.options.get("ldl_precision", "fp64")
        if self.precision not in ("fp64", "fp32"):
            raise ValueError(f"ldl_precision must be 'fp64' or 'fp32', not {self.precision!r}")
        self.refine = self.precision == "fp32" or bool(self.options.get("ldl_refine", False))
        self.refine_tol = self.options.get("ldl_refine_tol", 1e-10)
        self.use_pivtol = bool(self.options.get("ldl_use_pivtol", False))
        self.ldl = SparseLDL(
            self.hessian,
            SolverType.LDL,
            ustab=ustab,
            pivot_tol=pivot_tol,
            order=self.options.get("ldl_order", OrderingType.DEFAULT),
            precision=FP32_REFINE if self.precision == "fp32" else FP64,
            refine_tol=self.refine_tol,
        )
        self.pivtol = None
        self.fallbacks = 0
        self.last_factor_status = 0
        self.last_solve_status = 0
        self._fp64 = None
        self._fp64_current = False
        self._diag = None

    def factor(self, hessian, diagonal):
        if hessian is not self.hessian and hessian != self.hessian:
            raise ValueError("Hessian instance must be the same")
        # Copy the Hessian and diagonal entries to the host, as AmigoSolver
        _call(diagonal, "copy_device_to_host")
        _call(self.hessian, "copy_data_device_to_host")
        flag = self.ldl.factor(diagonal)
        self.last_factor_status = flag
        if flag == WARN_NONFINITE:
            raise RuntimeError(f"{self.solver_name} factorization failed: NaN or infinite "
                               f"pivots (non-finite values in the KKT matrix or diagonal)")
        if self.precision == "fp32":
            # kept for an FP64 refactorization should the refinement stall
            self._diag = _copy_array(diagonal)
            self._fp64_current = False
        # a singular matrix is not an error (zero pivots count positive in
        # the inertia, as the reference's factor, which always returns 0)
        return 0

    def solve(self, b, x):
        x.copy(b)
        _call(x, "copy_device_to_host")
        st = self.ldl.solve(x, refine=self.refine)
        if st == WARN_NOT_CONVERGED and self.precision == "fp32":
            st = self._solve_fp64(b, x)
        self.last_solve_status = st
        _call(x, "copy_host_to_device")

    def _solve_fp64(self, b, x):
        if self._fp64 is None:
            perm = self.ldl.get_perm()[0]
            self._fp64 = SparseLDL(self.hessian, SolverType.LDL, ustab=self.ldl._opts.ustab,
                                   pivot_tol=self.ldl._opts.pivot_tol, order=OrderingType.GIVEN,
                                   perm=perm, precision=FP64, refine_tol=self.refine_tol)
        if not self._fp64_current:
            self._fp64.set_pivot_threshold(self.ldl._opts.ustab, self.ldl._opts.pivot_tol)
            flag = self._fp64.factor(self._diag)
            if flag == WARN_NONFINITE:
                raise RuntimeError(f"{self.solver_name} FP64 refactorization failed: NaN or "
                                   f"infinite pivots")
            self._fp64_current = True
        self.fallbacks += 1
        x.copy(b)
        return self._fp64.solve(x, refine=True)

    def inertia_enabled(self):
        return True

    def get_inertia(self):
        return self.ldl.get_inertia()

    def set_pivot_tolerance(self, pivtol):
        self.pivtol = pivtol
        if self.use_pivtol:
            self.ldl.set_pivot_threshold(ustab=float(pivtol))
            self.ldl._opts.ustab = float(pivtol)

    def stats(self):
        """Engine statistics of the last factor / solve plus this adapter's
        counters (FP64 fallbacks, last statuses)."""
        st = self.ldl.stats()
        st.update(fallbacks=self.fallbacks, last_factor_status=self.last_factor_status,
                  last_solve_status=self.last_solve_status)
        return st


def _call(obj, name):
    fn = getattr(obj, name, None)
    if fn is not None:
        fn()


def _copy_array(v):
    if v is None:
        return None
    a = v.get_array() if hasattr(v, "get_array") else v
    if hasattr(a, "clone"):
        return a.clone()
    return np.array(a, dtype=np.float64, copy=True)
