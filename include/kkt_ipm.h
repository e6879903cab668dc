/*
 * kkt_ipm.h -- C ABI of the MI355X interior-point condensation kernels: the
 * barrier diagonal Sigma and the condensed right-hand side that feed the KKT
 * factorization and solve, the bound-dual back-substitution after the solve,
 * the fraction-to-the-boundary step and the complementarity / KKT-error
 * reductions.
 *
 * The problem layout is the reference's OptProblemInfo
 * (include/interior_point_backend.h:54-62): a full KKT vector of length n
 * (primals, slacks and multipliers), num_primals primal positions
 * primal_indices[i] with bounds lbx[i], ubx[i] (an infinite bound is absent),
 * num_constraints constraint positions constraint_indices[j] with targets
 * lbh[j]; the bound duals zl[i], zu[i] are indexed by primal position i.
 *
 * All vector arguments are DEVICE pointers (float64); stream is a
 * hipStream_t (NULL = default stream).  Calls that return scalars copy them
 * to the host and synchronize the stream.  Every call returns 0 on success
 * and < 0 on error (message from kkt_ipm_last_error(), thread-local).  The
 * elementwise outputs are bitwise equal to the reference's serial loops
 * (same operation order, no contraction); sums are deterministic two-stage
 * reductions, equal to the serial sums up to rounding.
 */
#ifndef KKT_IPM_H
#define KKT_IPM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kkt_ipm_handle* kkt_ipm_t;

/* Set-up, once per problem: host arrays, copied to the device.  Replaces the
 * construction of detail::OptProblemInfo in InteriorPointOptimizer
 * (include/interior_point_optimizer.h:98-140).  Indices must lie in [0, n);
 * a position listed twice takes its last occurrence, and a position that is
 * both a primal and a constraint takes the constraint in the residual, as the
 * reference's loops (primals first, then constraints) leave them. */
int kkt_ipm_create(int n, int num_primals, const int* primal_indices, const double* lbx,
                   const double* ubx, int num_constraints, const int* constraint_indices,
                   const double* lbh, kkt_ipm_t* out);

/* diag[n] <- Sigma: zl/(x - lb) + zu/(ub - x) at the primal positions, 0
 * elsewhere.  Replaces InteriorPointOptimizer::compute_diagonal
 * (include/interior_point_optimizer.h:267-285) and detail::compute_diagonal
 * (include/interior_point_backend.h:179-199). */
int kkt_ipm_diagonal(kkt_ipm_t h, const double* x, const double* zl, const double* zu,
                     double* diag, void* stream);

/* res[n] <- the condensed right-hand side: at primal positions
 * -(grad - zl + zu + (gap_l zl - mu)/gap_l - (gap_u zu - mu)/gap_u), at
 * constraint positions -(grad - lbh), 0 elsewhere; *norm <- ||res||_2 (host).
 * Replaces InteriorPointOptimizer::compute_residual
 * (include/interior_point_optimizer.h:233-257) and detail::compute_residual
 * (include/interior_point_backend.h:146-176). */
int kkt_ipm_residual(kkt_ipm_t h, double mu, const double* x, const double* zl,
                     const double* zu, const double* grad, double* res, double* norm,
                     void* stream);

/* dzl, dzu[num_primals] <- bound-dual steps from the primal step px[n]:
 * -(gap_l zl - mu + zl dx)/gap_l and -(gap_u zu - mu - zu dx)/gap_u, 0 where
 * the bound is absent.  Replaces detail::compute_bound_dual_step
 * (include/interior_point_backend.h:213-236), the dual part of
 * InteriorPointOptimizer::compute_update (interior_point_optimizer.h:
 * 296-324). */
int kkt_ipm_bound_dual_step(kkt_ipm_t h, double mu, const double* x, const double* zl,
                            const double* zu, const double* px, double* dzl, double* dzu,
                            void* stream);

/* Fraction to the boundary: the largest primal step ax and dual step az in
 * (0, 1] that keep every gap at >= (1 - tau) of its value, and the KKT-vector
 * index (xi) / primal's index (zi) of the first binding term in primal order,
 * -1 when the step is 1.  Replaces detail::compute_max_step
 * (include/interior_point_backend.h:242-287) as called with ax = az = 1 and
 * xi = zi = -1 (amigo/amigo.cpp:788-794). */
int kkt_ipm_max_step(kkt_ipm_t h, double tau, const double* x, const double* zl,
                     const double* zu, const double* dx, const double* dzl,
                     const double* dzu, double* ax, int* xi, double* az, int* zi,
                     void* stream);

/* Average complementarity avg = sum(gap z) / (number of finite bounds) and the
 * uniformity xi = clamp(min(gap z) / avg, 0, 1) (1 when avg <= 0).  Replaces
 * InteriorPointOptimizer::compute_complementarity
 * (include/interior_point_optimizer.h:399-433) and
 * detail::compute_complementarity (include/interior_point_backend.h:
 * 313-339). */
int kkt_ipm_complementarity(kkt_ipm_t h, const double* x, const double* zl,
                            const double* zu, double* avg, double* xi, void* stream);

/* Infinity-norm optimality errors: dual = max|grad - zl + zu| over primals,
 * primal = max|grad - lbh| over constraints, comp = max|gap z - mu| over
 * finite bounds.  Replaces detail::compute_kkt_error
 * (include/interior_point_backend.h:345-369). */
int kkt_ipm_kkt_error(kkt_ipm_t h, double mu, const double* x, const double* zl,
                      const double* zu, const double* grad, double* dual, double* primal,
                      double* comp, void* stream);

const char* kkt_ipm_last_error(void);

void kkt_ipm_destroy(kkt_ipm_t h);

#ifdef __cplusplus
}
#endif

#endif /* KKT_IPM_H */
