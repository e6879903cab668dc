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

from .ldl import FP32_REFINE, FP64, OrderingType, SolverType, SparseLDL

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

    ``options`` keys read (all optional, beyond the IPM's own):
    ``"ldl_precision"``: "fp64" (default) or "fp32" (FP32 factor + FP64
    refinement); ``"ldl_refine_tol"`` (1e-10); ``"ldl_order"``
    (OrderingType.DEFAULT).
    """

    solver_name = "MI355XSolver"

    def __init__(self, options, state):
        self.options = options if options is not None else {}
        self.hessian = state.hessian
        ustab = 0.01
        pivot_tol = 1e-14
        prec = self.options.get("ldl_precision", "fp64")
        self.refine = prec == "fp32" or bool(self.options.get("ldl_refine", False))
        self.ldl = SparseLDL(
            self.hessian,
            SolverType.LDL,
            ustab=ustab,
            pivot_tol=pivot_tol,
            order=self.options.get("ldl_order", OrderingType.DEFAULT),
            precision=FP32_REFINE if prec == "fp32" else FP64,
            refine_tol=self.options.get("ldl_refine_tol", 1e-10),
        )

    def factor(self, hessian, diagonal):
        if hessian is not self.hessian and hessian != self.hessian:
            raise ValueError("Hessian instance must be the same")
        flag = self.ldl.factor(diagonal)
        if flag < 0:
            raise RuntimeError(f"{self.solver_name} factorization failed with flag = {flag}")
        return 0

    def solve(self, b, x):
        x.copy(b)
        self.ldl.solve(x, refine=self.refine)

    def inertia_enabled(self):
        return True

    def get_inertia(self):
        return self.ldl.get_inertia()

    def set_pivot_tolerance(self, pivtol):
        # The reference's AmigoSolver ignores this call too (amigo_solver.py:63-64).
        pass
