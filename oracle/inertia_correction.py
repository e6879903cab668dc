"""CPU restatement of the reference's inertia-correction loop (Algorithm IC of
Waechter & Biegler 2006, Table 3) -- test infrastructure, not product code.
Only tests/ import it.

The reference's ``InertiaCorrector`` (amigo/algorithm/solvers/
inertia_correction.py:11-416) drives any ``LinearSolver`` through
``set_pivot_tolerance``, ``factor(hessian, diagonal)`` and ``get_inertia``.
The reference package cannot be imported here (its ``amigo.amigo`` pybind
extension needs A2D), so the tests drive the GPU solver and the oracle with
this restatement and compare what the loop sees at every attempt.

Restated behaviour (line numbers in inertia_correction.py):
  constants                      61-84  (dw 1e-4 first, x100 first growth,
                                        x8 later, /3 from the last value,
                                        [1e-20, 1e20]; dc = 1e-8 mu^0.25;
                                        degeneracy after 3 iterations;
                                        pivtol 1e-6 -> sqrt up to 0.1)
  next delta_w                   95-120
  wrong inertia / singular       122-195
  degeneracy test bookkeeping    197-238, 240-276
  inertia check                  278-287
  perturbed diagonal             289-304  (diag + 1e-12 + dw on primal
                                        indices, -dc on constraint indices)
  main loop                      315-395  (at most 11 factorizations)
  increase quality               397-403
"""

import numpy as np

NOT_YET, NOT_DEGEN, DEGENERATE = 0, 1, 2
NO_TEST, T_C0X0, T_C1X0, T_C0X1, T_C1X1 = 0, 1, 2, 3, 4


class InertiaCorrection:
    """The IC state machine over host arrays.

    primal_idx / dual_idx: the variables that receive +delta_w / -delta_c;
    num_primal / num_dual: the inertia the loop accepts (default: the index
    counts); tol: options["inertia_tolerance"]."""

    DW_FIRST, DW_MIN, DW_MAX = 1e-4, 1e-20, 1e20
    GROW_FIRST, GROW, SHRINK = 100.0, 8.0, 1.0 / 3.0
    DC_VAL, DC_EXP = 1e-8, 0.25
    EPS = 1e-12
    DEGEN_ITERS = 3
    PIVTOL_MAX = 0.1
    MAX_CORRECTIONS = 10

    def __init__(self, primal_idx, dual_idx, num_primal=None, num_dual=None, tol=0):
        self.primal_idx = np.asarray(primal_idx)
        self.dual_idx = np.asarray(dual_idx)
        self.num_primal = len(self.primal_idx) if num_primal is None else num_primal
        self.num_dual = len(self.dual_idx) if num_dual is None else num_dual
        self.tol = tol
        self.dw = self.dc = 0.0            # current perturbations
        self.dw_last = self.dc_last = 0.0  # of the previous system
        self.hess_degen = self.jac_degen = NOT_YET
        self.degen_count = 0
        self.test = NO_TEST
        self.pivtol = 1e-6
        self.log = []                      # (dw, dc, npos, nneg) per attempt

    # -- helpers ---------------------------------------------------------------
    def _dc_value(self, mu):
        return self.DC_VAL * mu ** self.DC_EXP

    def _grow_dw(self):
        """Next delta_w; False (and delta_w unchanged) past the maximum."""
        old = self.dw
        if self.dw == 0.0:
            self.dw = self.DW_FIRST if self.dw_last == 0.0 else max(
                self.DW_MIN, self.dw_last * self.SHRINK)
        elif self.dw_last == 0.0 or self.dw > 1e5 * self.dw_last:
            self.dw *= self.GROW_FIRST
        else:
            self.dw *= self.GROW
        if self.dw > self.DW_MAX:
            self.dw = old
            self.dw_last = 0.0
            return False
        return True

    def _conclude_test(self):
        t = self.test
        if t == NO_TEST:
            return
        if t == T_C0X0:
            self.hess_degen = NOT_DEGEN if self.hess_degen == NOT_YET else self.hess_degen
            self.jac_degen = NOT_DEGEN if self.jac_degen == NOT_YET else self.jac_degen
        elif t == T_C1X0:
            if self.hess_degen == NOT_YET:
                self.hess_degen = NOT_DEGEN
            if self.jac_degen == NOT_YET:
                self.degen_count += 1
                if self.degen_count >= self.DEGEN_ITERS:
                    self.jac_degen = DEGENERATE
        elif t == T_C0X1:
            if self.jac_degen == NOT_YET:
                self.jac_degen = NOT_DEGEN
            if self.hess_degen == NOT_YET:
                self.degen_count += 1
                if self.degen_count >= self.DEGEN_ITERS:
                    self.hess_degen = DEGENERATE
        else:  # T_C1X1
            self.degen_count += 1
            if self.degen_count >= self.DEGEN_ITERS:
                self.hess_degen = self.jac_degen = DEGENERATE
        self.test = NO_TEST

    def _new_system(self, mu):
        self._conclude_test()
        if self.dw > 0.0:
            self.dw_last = self.dw
        if self.dc > 0.0:
            self.dc_last = self.dc
        undecided = NOT_YET in (self.hess_degen, self.jac_degen)
        self.test = T_C0X0 if undecided else NO_TEST
        self.dc = self._dc_value(mu) if self.jac_degen == DEGENERATE else 0.0
        self.dw = 0.0
        if self.hess_degen == DEGENERATE:
            self._grow_dw()

    def _on_wrong_inertia(self, mu):
        self._conclude_test()
        if self._grow_dw():
            return True
        if self.dc == 0.0:
            self.dc = self._dc_value(mu)
            self.dw = 0.0
            if self.hess_degen == DEGENERATE:
                self.hess_degen = NOT_YET
            self.test = NO_TEST
            return self._grow_dw()
        return False

    def _on_singular(self, mu):
        if NOT_YET in (self.hess_degen, self.jac_degen):
            t = self.test
            if t == T_C0X0:
                if self.jac_degen == NOT_YET:
                    self.dc = self._dc_value(mu)
                    self.test = T_C1X0
                else:
                    if not self._grow_dw():
                        return False
                    self.test = T_C0X1
            elif t == T_C1X0:
                self.dc = 0.0
                if not self._grow_dw():
                    return False
                self.test = T_C0X1
            elif t == T_C0X1:
                self.dc = self._dc_value(mu)
                if not self._grow_dw():
                    return False
                self.test = T_C1X1
            elif t == T_C1X1:
                if not self._grow_dw():
                    return False
        elif self.dc > 0.0:
            if not self._grow_dw():
                return False
        else:
            self.dc = self._dc_value(mu)
        return True

    def _inertia_ok(self, npos, nneg):
        tot = self.num_primal + self.num_dual
        return (abs(npos - self.num_primal) <= self.tol and abs(nneg - self.num_dual) <= self.tol
                and npos + nneg >= tot - self.tol)

    def perturbed(self, diag):
        d = np.array(diag, dtype=np.float64, copy=True)
        d[self.primal_idx] += self.EPS + self.dw
        if self.dc > 0:
            d[self.dual_idx] -= self.dc
        return d

    # -- the loop ----------------------------------------------------------------
    def factor(self, factor_fn, diag, mu, set_pivtol=None):
        """factor_fn(d) factors K + diag(d) and returns (npos, nneg).
        Returns True when an accepted inertia was reached."""
        self._new_system(mu)
        if set_pivtol:
            set_pivtol(self.pivtol)
        improved = False
        for _ in range(self.MAX_CORRECTIONS + 1):
            npos, nneg = factor_fn(self.perturbed(diag))
            self.log.append((self.dw, self.dc, int(npos), int(nneg)))
            if self._inertia_ok(npos, nneg):
                return True
            if nneg < self.num_dual:
                singular = True
                if not improved and self.pivtol < self.PIVTOL_MAX:
                    self.pivtol = min(self.PIVTOL_MAX, self.pivtol ** 0.5)
                    if set_pivtol:
                        set_pivtol(self.pivtol)
                    improved = True
                    singular = False
                if singular and not self._on_singular(mu):
                    break
            elif not self._on_wrong_inertia(mu):
                break
        return False
