/*
 * kkt_ldl.h -- C ABI of the MI355X supernodal LDL^T engine for interior-point
 * KKT systems.  Plain pointers and sizes only; no C++ or torch types.
 *
 * Every call returns an int status: 0 ok, < 0 error (message from
 * kkt_ldl_last_error(), thread-local), > 0 warning (see KKT_LDL_WARN_*).
 * No C++ exception crosses this boundary.
 *
 * Threading: one handle per thread.  kkt_ldl_solve may be called
 * concurrently on one handle from one thread's stream only; kkt_ldl_factor
 * may not overlap any other call on the same handle.
 *
 * Each entry point replaces a piece of the reference's SparseLDL binding
 * (amigo/amigo.cpp:648-732 over include/sparse_ldl.h); the reference file and
 * line is given next to each declaration.
 */
#ifndef KKT_LDL_H
#define KKT_LDL_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kkt_ldl_handle* kkt_ldl_t;

/* Orderings, numbered as the reference's OrderingType (include/
 * ordering_utils.h:21: NESTED_DISSECTION, AMD, MULTI_COLOR, NATURAL), so that
 * int(amigo.OrderingType.X) can be passed unchanged.  MULTI_COLOR is
 * rejected (the reference reads an uninitialised permutation for it,
 * ordering_utils.h:280-291); GIVEN (a caller's permutation) is this ABI's
 * own addition.  Unknown values are rejected. */
enum { KKT_LDL_ORDER_ND = 0, KKT_LDL_ORDER_AMD = 1, KKT_LDL_ORDER_MULTI_COLOR = 2,
       KKT_LDL_ORDER_NATURAL = 3, KKT_LDL_ORDER_GIVEN = 4 };
/* Factorization kinds (include/sparse_ldl.h:23 SolverType). */
enum { KKT_LDL_LDL = 0, KKT_LDL_CHOLESKY = 1 };
/* Arithmetic of the stored factor. */
enum { KKT_LDL_FP64 = 0, KKT_LDL_FP32_REFINE = 1 };
/* Memory space of caller buffers. */
enum { KKT_LDL_HOST = 0, KKT_LDL_DEVICE = 1 };
/* Positive status codes (warnings).  A singular matrix factors to the end,
 * as in the reference, with zero pivots counted positive in the inertia
 * (include/sparse_ldl.h:1958); kkt_ldl_factor then returns
 * KKT_LDL_WARN_SINGULAR and kkt_ldl_inertia reports the zero pivots in
 * nzero.  A NaN or infinite pivot (non-finite values in K or the diagonal,
 * or an indefinite matrix in Cholesky mode) returns KKT_LDL_WARN_NONFINITE.
 * kkt_ldl_solve_refined returns KKT_LDL_WARN_NOT_CONVERGED when the residual
 * target is not met, including a non-finite residual. */
enum { KKT_LDL_WARN_SINGULAR = 1, KKT_LDL_WARN_NOT_CONVERGED = 2,
       KKT_LDL_WARN_NONFINITE = 3 };

typedef struct {
  int order;            /* KKT_LDL_ORDER_*                                  */
  int solver_type;      /* KKT_LDL_LDL or KKT_LDL_CHOLESKY                  */
  double ustab;         /* threshold-pivot stability factor (ref. 0.01)     */
  double pivot_tol;     /* absolute pivot tolerance (ref. 1e-14)            */
  double delay_growth;  /* initial front slack for delayed pivots (ref. 2)  */
  int precision;        /* KKT_LDL_FP64 or KKT_LDL_FP32_REFINE              */
  double refine_tol;    /* ||Kx-b||/||b|| target of kkt_ldl_solve_refined   */
  int max_refine;       /* refinement step limit                            */
  const int* perm;      /* KKT_LDL_ORDER_GIVEN: perm[new] = old             */
  const double* vals_hint; /* optional values of the pattern (CSR order), read
                            * only by kkt_ldl_create with KKT_LDL_ORDER_ND: the
                            * columns of each separator are then ordered so
                            * that predicted 2x2 pivot partners are adjacent
                            * (the fill does not change).  The factorization's
                            * pivot choices depend on it only through that
                            * order.  NULL: separators in index order.         */
} kkt_ldl_options;

typedef struct {
  int n, num_nodes, num_snodes, num_levels, max_front;
  long long nnz_l;          /* fundamental nnz(L) incl. diagonal            */
  double flops_factor;      /* fundamental sum c_j (c_j + 2)                */
  long long stored_factor;  /* executed: sum over nodes rows x pivots       */
  int num_delayed;          /* delayed pivots summed over nodes (last factor)*/
  int num_2x2;              /* 2x2 pivot blocks (last factor)               */
  int replans;              /* front re-layouts caused by delayed pivots    */
  int refine_steps;         /* last kkt_ldl_solve_refined                   */
  double residual;          /* last relative residual ||Kx-b||/||b||        */
  double t_analyze_ms, t_factor_ms, t_solve_ms;
  long long front_bytes;    /* device arena holding every front             */
  int num_zero_pivots;      /* last factor: exact zero 1x1 pivots / 2x2 dets */
  int num_nonfinite_pivots; /* last factor: NaN or infinite pivots          */
  long long crit_path_pivots; /* largest sum of supernode columns on a
                               * leaf-to-root path of the assembly tree     */
  long long level_panel_sum;  /* sum over tree levels of the largest number
                               * of 16-column panels of a front of the level */
  long long panel_launches;   /* panel launches of the factorization's
                               * schedule: a front starts once its children
                               * are factored, so this is close to the
                               * panels on the tree's longest path          */
} kkt_ldl_stats;

/* Defaults matching AmigoSolver (amigo/algorithm/solvers/amigo_solver.py:15-25). */
void kkt_ldl_default_options(kkt_ldl_options* opts);

/* Pattern + symbolic analysis.  rowp[n+1], cols[nnz] (symmetric, sorted,
 * diagonal present).  Replaces SparseLDL(mat, solver_type, ustab, pivot_tol,
 * delay_growth, order) -- include/sparse_ldl.h:65-118, amigo/amigo.cpp:655-662. */
int kkt_ldl_create(int n, const int* rowp, const int* cols,
                   const kkt_ldl_options* opts, kkt_ldl_t* out);

/* Numeric factorization of K + diag(d).  vals[nnz] in CSR order, diag[n] or
 * NULL.  mem = KKT_LDL_HOST or KKT_LDL_DEVICE for both pointers; stream is a
 * hipStream_t (NULL = default stream).  Replaces SparseLDL.factor(diagonal)
 * -- include/sparse_ldl.h:136-158, amigo/amigo.cpp:663-664. */
int kkt_ldl_factor(kkt_ldl_t h, const double* vals, const double* diag, int mem,
                   void* stream);

/* Numeric refactorization with a new diagonal only: K's values are those
 * of the last kkt_ldl_factor call, kept on the device, so only diag[n] is
 * read (mem as in kkt_ldl_factor).  For the inertia-correction retries, which
 * change only the diagonal perturbation (amigo/algorithm/solvers/
 * inertia_correction.py:280-317, 324-383); the reference re-reads the whole
 * matrix on each of them (include/sparse_ldl.h:136-158).  Same status codes
 * as kkt_ldl_factor; -3 before any kkt_ldl_factor. */
int kkt_ldl_refactor(kkt_ldl_t h, const double* diag, int mem, void* stream);

/* Threshold-pivot stability factor (ustab) and absolute pivot tolerance
 * used from the next factorization on; a negative value leaves that setting
 * unchanged.  The reference's AmigoSolver accepts set_pivot_tolerance and
 * ignores it (amigo/algorithm/solvers/amigo_solver.py:63-64); MUMPS maps the
 * IPM's pivtol to its relative threshold CNTL(1) (mumps_solver.py:323-325),
 * which is the role ustab plays here. */
int kkt_ldl_set_pivot_threshold(kkt_ldl_t h, double ustab, double pivot_tol);

/* x <- (K + diag d)^{-1} x for nrhs columns of length n with leading
 * dimension ldx (column major, FP64).  Replaces SparseLDL.solve -- include/
 * sparse_ldl.h:169-197, amigo/amigo.cpp:665-727. */
int kkt_ldl_solve(kkt_ldl_t h, int nrhs, double* x, int ldx, int mem,
                  void* stream);

/* Solve followed by FP64 iterative refinement on the GPU against the stored
 * K + diag until ||Kx-b||/||b|| <= refine_tol (or max_refine steps).  No
 * counterpart in the reference (it never refines, amigo/algorithm/solvers/
 * linear_solver.py:5-35); used by the FP32 factor and by the IPM drop-in.
 * The refinement stops when the worst column's residual no longer decreases
 * and then returns the best iterate.  Returns KKT_LDL_WARN_NOT_CONVERGED if
 * the target is not met (a NaN or infinite residual counts as not met). */
int kkt_ldl_solve_refined(kkt_ldl_t h, int nrhs, double* x, int ldx, int mem,
                          void* stream);

/* kkt_ldl_solve_refined for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): x is a DEVICE pointer (n x nrhs, leading dimension ldx), solved
 * and refined in place, with the refinement's buffers in a static per-GPU
 * scratch instead of handle buffers allocated on demand.  Limits:
 * ldx * nrhs <= 1048576 and nrhs <= 64 (error -7 beyond); the handle's solve
 * buffers must already hold nrhs right-hand sides (every kkt_ldl_factor
 * sizes them for 1; a kkt_ldl_solve with a larger nrhs grows them), else
 * error -7.  Calls no hipMalloc or hipFree; no per-kernel timing.  Calls that
 * overlap on one GPU share the scratch and race.  No counterpart in the
 * reference (which always allocates its solve temporaries,
 * include/sparse_ldl.h:1829); it exists for the callers that cannot
 * allocate.  Same results and status codes as kkt_ldl_solve_refined. */
int kkt_ldl_solve_refined_noalloc(kkt_ldl_t h, int nrhs, double* x, int ldx, void* stream);

/* Inertia from D of the last factorization; zero pivots count positive as in
 * the reference.  Replaces SparseLDL.get_inertia -- include/sparse_ldl.h:
 * 205-213, 1935-1987, amigo/amigo.cpp:728-731.  nzero (may be NULL) is the
 * number of zero pivots, already counted in npos; the reference does not
 * report it. */
int kkt_ldl_inertia(kkt_ldl_t h, int* npos, int* nneg, int* nzero);

/* Fill-reducing permutation (perm[new] = old, iperm[old] = new); identity for
 * NATURAL.  Exposes SparseLDL::invperm (include/sparse_ldl.h:3265). */
int kkt_ldl_get_perm(kkt_ldl_t h, int* perm, int* iperm);

/* Number of relaxed supernodes and total contribution rows, to size the
 * arrays of kkt_ldl_get_supernodes. */
int kkt_ldl_get_sizes(kkt_ldl_t h, int* num_nodes, int* contrib_total);

/* Supernode partition and tree exactly as SparseLDL stores them
 * (node_ptr[N+1], node_to_var[n], children_ptr[N+2], children[N],
 * contrib_ptr[N+1], contrib_rows[contrib_total]) -- include/sparse_ldl.h:
 * 3268-3281. */
int kkt_ldl_get_supernodes(kkt_ldl_t h, int* node_ptr, int* node_to_var,
                           int* children_ptr, int* children, int* contrib_ptr,
                           int* contrib_rows);

/* Per-node pivot outcome of the last factorization: pivots, delayed-out
 * count, 2x2 blocks, root flag (arrays of num_nodes). */
int kkt_ldl_get_node_info(kkt_ldl_t h, int* npiv, int* ndel, int* n2x2,
                          int* root);

int kkt_ldl_get_stats(kkt_ldl_t h, kkt_ldl_stats* st);

/* Device time of the last factor / solve kernels on the caller's stream, ms
 * (hipEvent based). */
int kkt_ldl_last_times(kkt_ldl_t h, double* factor_ms, double* solve_ms);

/* Per-kernel-class device timing: when on, every launch of the next
 * factor / solve is timed with hipEvents on the caller's stream (this adds
 * launch latency; use it for attribution, not for the headline time).  The
 * panel and trailing-update launches carry the events in the launch itself
 * (hipExtLaunchKernelGGL), so their times are the kernels' own durations;
 * the other classes are bracketed, dispatch included.
 * Classes: 0 prepare, 1 assembly, 2 LDS fronts, 3 small fronts, 4 panel,
 * 5 MFMA trailing update, 6 forward solve, 7 backward solve, 8 residual,
 * 9 panel inverses of the large fronts (after the factorization).
 * No counterpart in the reference, which has no kernel instrumentation
 * (amigo/algorithm/iteration_logger.py:28-39 only timestamps iterations). */
int kkt_ldl_set_profiling(kkt_ldl_t h, int on);
int kkt_ldl_get_kernel_times(kkt_ldl_t h, double* ms, int* launches, int nclass);

const char* kkt_ldl_last_error(void);

void kkt_ldl_destroy(kkt_ldl_t h);

#ifdef __cplusplus
}
#endif
#endif /* KKT_LDL_H */
