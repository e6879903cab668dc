// Golden-vector generator -- test infrastructure, not product code.
//
// Compiles the reference's own SparseLDL (include/sparse_ldl.h and the
// headers it includes, read from the reference tree where it lies) into a
// small C ABI so that tests/golden/gen_golden.py can record its outputs
// (ordering, supernodes, row patterns, per-node pivot counts, inertia,
// solutions) as fixtures.  Built by `make -C oracle ref` into oracle/_ref/.
//
// Private members of SparseLDL are read through `private` -> `public`
// after every standard header it needs is already included.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <deque>
#include <iomanip>
#include <iostream>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>
#include <mpi.h>
#include <omp.h>
#define private public
#include "sparse_ldl.h"
#include "fixed_variables.h"
#undef private

using Ldl = amigo::SparseLDL<double>;

struct RefHandle {
  std::shared_ptr<amigo::CSRMat<double>> mat;
  std::unique_ptr<Ldl> ldl;
  int n;
  int nnz;
};

extern "C" {

// order: 0 NATURAL, 1 AMD.  solver_type: 0 LDL, 1 CHOLESKY.
void* ref_analyze(int n, const int* rowp, const int* cols, int order,
                  double ustab, double pivot_tol, double delay_growth,
                  int solver_type) {
  RefHandle* h = new RefHandle();
  h->n = n;
  h->nnz = rowp[n];
  h->mat = amigo::CSRMat<double>::create_from_csr_data(n, n, rowp[n], rowp, cols);
  h->ldl.reset(new Ldl(h->mat,
                       solver_type == 1 ? Ldl::SolverType::CHOLESKY
                                        : Ldl::SolverType::LDL,
                       ustab, pivot_tol, delay_growth,
                       order == 1 ? amigo::OrderingType::AMD
                                  : amigo::OrderingType::NATURAL));
  return h;
}

void ref_destroy(void* h) { delete (RefHandle*)h; }

int ref_num_nodes(void* h) { return ((RefHandle*)h)->ldl->num_nodes; }

int ref_contrib_total(void* hh) {
  Ldl& L = *((RefHandle*)hh)->ldl;
  return L.contrib_ptr[L.num_nodes];
}

void ref_get_perm(void* hh, int* perm, int* iperm) {
  RefHandle* h = (RefHandle*)hh;
  Ldl& L = *h->ldl;
  for (int i = 0; i < h->n; i++) {
    iperm[i] = L.invperm ? L.invperm[i] : i;
  }
  for (int i = 0; i < h->n; i++) perm[iperm[i]] = i;
}

void ref_get_nodes(void* hh, int* node_ptr, int* node_to_var, int* children_ptr,
                   int* children, int* contrib_ptr, int* contrib_rows) {
  RefHandle* h = (RefHandle*)hh;
  Ldl& L = *h->ldl;
  int N = L.num_nodes;
  std::copy(L.node_ptr, L.node_ptr + N + 1, node_ptr);
  std::copy(L.node_to_var, L.node_to_var + h->n, node_to_var);
  std::copy(L.node_children_ptr, L.node_children_ptr + N + 2, children_ptr);
  std::copy(L.node_children, L.node_children + N, children);
  std::copy(L.contrib_ptr, L.contrib_ptr + N + 1, contrib_ptr);
  std::copy(L.contrib_rows, L.contrib_rows + L.contrib_ptr[N], contrib_rows);
}

int ref_factor(void* hh, const double* vals, const double* diag) {
  RefHandle* h = (RefHandle*)hh;
  double* data = h->mat->get_data_ptr();
  std::copy(vals, vals + h->nnz, data);
  std::shared_ptr<amigo::Vector<double>> d;
  if (diag) {
    d = std::make_shared<amigo::Vector<double>>(h->n);
    std::copy(diag, diag + h->n, d->get_array());
  }
  return h->ldl->factor(d);
}

void ref_solve(void* hh, int nrhs, double* x, int ldx) {
  ((RefHandle*)hh)->ldl->solve(nrhs, x, ldx);
}

void ref_inertia(void* hh, int* npos, int* nneg) {
  ((RefHandle*)hh)->ldl->get_inertia(npos, nneg);
}

// Per-node pivot counts of the last factorization (same layout as the
// oracle's orc_node_info).
void ref_node_info(void* hh, int* npiv, int* ndel, int* n2x2, int* root) {
  Ldl& L = *((RefHandle*)hh)->ldl;
  for (int k = 0; k < L.num_nodes; k++) {
    int np, nd, ni;
    const int *piv, *ipiv;
    const double* Lv;
    L.fact.get_factor(k, &np, &piv, &nd, nullptr, &Lv, &ni, &ipiv);
    npiv[k] = np;
    ndel[k] = nd;
    root[k] = ipiv != nullptr;
    const int* p = ipiv ? ipiv : piv;
    int c = 0;
    for (int i = 0; i < np;) {
      if (p[i] < 0) { c++; i += 2; } else { i++; }
    }
    n2x2[k] = c;
  }
}

// KKT value assembly as OptimizationProblem::hessian does it on the host
// (include/optimization_problem.h:1063-1087): zero the values, one
// CSRMat::add_row per (element, column) call, then
// FixedVariables::zero_rows_and_columns.  Call c adds call_vals[p] at
// (call_row[c], call_cols[p]) for p in [call_ptr[c], call_ptr[c+1]).
void ref_assemble(int n, const int* rowp, const int* cols, int ncalls,
                  const int* call_row, const int* call_ptr, const int* call_cols,
                  const double* call_vals, int nfixed, const int* fixed,
                  double* vals_out) {
  auto mat = amigo::CSRMat<double>::create_from_csr_data(n, n, rowp[n], rowp, cols);
  mat->zero();
  for (int c = 0; c < ncalls; c++)
    mat->add_row(call_row[c], call_ptr[c + 1] - call_ptr[c], call_cols + call_ptr[c],
                 call_vals + call_ptr[c]);
  if (nfixed > 0) {
    auto dofs = std::make_shared<amigo::Vector<int>>(nfixed);
    std::copy(fixed, fixed + nfixed, dofs->get_array());
    amigo::FixedVariables fv(dofs, mat);
    fv.zero_rows_and_columns<amigo::ExecPolicy::SERIAL>(mat);
  }
  std::copy(mat->get_data_ptr(), mat->get_data_ptr() + rowp[n], vals_out);
}

}  // extern "C"
