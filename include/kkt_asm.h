/*
 * kkt_asm.h -- C ABI of the MI355X KKT value assembly: element Hessian
 * contributions scatter-added into the CSR values of the KKT matrix, with the
 * rows and columns of fixed variables replaced by identity rows.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success and < 0 on
 * error (message from kkt_asm_last_error(), thread-local).
 */
#ifndef KKT_ASM_H
#define KKT_ASM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kkt_asm_handle* kkt_asm_t;

/* Pattern set-up, once per CSR pattern.  Contribution q adds to entry
 * (crow[q], ccol[q]) of the CSR (rowp[n+1], cols[nnz], columns sorted in each
 * row).  A contribution outside the pattern is dropped, as the reference
 * drops it (CSRMat::add_row, include/csr_matrix.h:555-567; hess_pos = -1 in
 * include/cuda/component_group_backend.cuh:236-250).  fixed[nfixed] lists
 * fixed variables: their rows and columns become 0 with 1 on the diagonal
 * (FixedVariables::zero_rows_and_columns, include/fixed_variables.h:66-69,
 * 88-130); indices outside [0, n) are ignored, as FixedVariables ignores
 * them (include/fixed_variables.h:36-41).  For an element group of ncomp variables with indices idx[e][.],
 * the reference's contribution order is q = ncomp*(ncomp*e + k) + i for the
 * entry (idx[e][k], idx[e][i]) (include/cuda/component_group_backend.cuh:
 * 159-189, include/component_group.h:491-528). */
int kkt_asm_create(int n, const int* rowp, const int* cols, long long ncontrib,
                   const int* crow, const int* ccol, int nfixed, const int* fixed,
                   kkt_asm_t* out);

/* vals[nnz] <- assembled CSR values from contrib_vals[ncontrib].  Every
 * entry of vals is written (entries without contributions become 0), so no
 * zeroing pass is needed.  mem = 0 host pointers, 1 device pointers; stream
 * is a hipStream_t (NULL = default stream).  Replaces the reference's
 * zero + add_hessian + zero_rows_and_columns sequence of
 * OptimizationProblem::hessian (include/optimization_problem.h:1063-1087;
 * include/component_group.h:491-528; include/cuda/
 * component_group_backend.cuh:364-395).  The sum per
 * entry is taken in a fixed order, so the result is bitwise repeatable. */
int kkt_asm_values(kkt_asm_t h, const double* contrib_vals, double* vals, int mem,
                   void* stream);

/* Pattern sizes: nnz of the CSR, number of contributions, and how many of
 * them fall outside the pattern (dropped). */
int kkt_asm_get_info(kkt_asm_t h, long long* nnz, long long* ncontrib,
                     long long* ndropped);

const char* kkt_asm_last_error(void);

void kkt_asm_destroy(kkt_asm_t h);

#ifdef __cplusplus
}
#endif
#endif /* KKT_ASM_H */
