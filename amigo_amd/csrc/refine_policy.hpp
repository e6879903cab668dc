// Host-side decisions of the iterative refinement in kkt_ldl_solve_refined:
// the worst relative residual over the right-hand sides, and whether to stop,
// go on, or take back the last correction.  Plain C++ (no HIP), so that the
// CPU tests can compile and exercise it on its own.
#ifndef KKT_REFINE_POLICY_HPP
#define KKT_REFINE_POLICY_HPP

#include <cmath>
#include <limits>

namespace kktldl {

// max_c ||r_c|| / ||b_c|| from the squared norms bn2[c] = ||b_c||^2 and
// rn2[c] = ||r_c||^2 (||r_c|| alone where b_c = 0).  A NaN or infinite
// residual in any column gives +infinity, so that it never reads as
// converged (std::max would drop a NaN).
inline double refine_worst(const double* bn2, const double* rn2, int nrhs) {
  const double inf = std::numeric_limits<double>::infinity();
  double w = 0.0;
  for (int c = 0; c < nrhs; c++) {
    const double q = bn2[c] > 0.0 ? std::sqrt(rn2[c] / bn2[c]) : std::sqrt(rn2[c]);
    if (!(q <= inf)) return inf;  // NaN
    if (q > w) w = q;
  }
  return w;
}

enum { REFINE_DONE = 0, REFINE_CONTINUE = 1, REFINE_UNDO_LAST = 2 };

// Decision after the residual check of iterate `it` (0 = the plain solve):
// w is its worst residual, best the smallest one of the iterates before it.
// A step that does not lower the worst residual (a diverging refinement, or a
// non-finite residual) is taken back and the refinement stops with the best
// iterate; a non-finite first residual cannot be refined.
inline int refine_next(int it, double w, double best, double tol, int max_refine) {
  if (it > 0 && !(w < best)) return REFINE_UNDO_LAST;
  if (w <= tol || it >= max_refine || !std::isfinite(w)) return REFINE_DONE;
  return REFINE_CONTINUE;
}

}  // namespace kktldl

#endif  // KKT_REFINE_POLICY_HPP
