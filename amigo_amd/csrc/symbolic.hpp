// Host-side symbolic analysis of the KKT pattern: fill-reducing ordering,
// elimination tree, supernodes and the per-node row structure.  The results
// are bit-identical to the reference's SparseLDL constructor
// (include/sparse_ldl.h:2222-2376) for NATURAL and AMD orderings, because
// the numeric factor and the solve index everything by this structure.
#pragma once
#include <cstdint>
#include <vector>

namespace kktldl {

struct SymbolicResult {
  int n = 0;
  bool permuted = false;
  std::vector<int> perm, iperm;       // perm[new] = old
  std::vector<int> parent, post;      // elimination tree (permuted space)
  std::vector<int> colcount;          // fundamental column counts incl. diag
  int num_snodes = 0;                 // fundamental supernodes before merging
  int num_nodes = 0;                  // relaxed supernodes
  std::vector<int> node_ptr;          // [N+1]
  std::vector<int> node_to_var;       // [n], original indices
  std::vector<int> node_parent;       // [N], -1 for roots
  std::vector<int> children_ptr;      // [N+2], roots after position N
  std::vector<int> children;          // [N]
  std::vector<int> contrib_ptr;       // [N+1]
  std::vector<int> contrib_rows;      // original indices, ascending permuted order
};

// Reference BlockAMD (include/block_amd.h:27-452) with no multiplier
// constraint, on the pattern with the diagonal added (ordering_utils.h:313-360).
void amd_ordering(int n, const int* rowp, const int* cols, int* perm);

// Nested dissection on the graph of the pattern (no METIS in this build;
// nd_ordering.cpp): multilevel node bisection against BFS level-set
// separators, both refined by a band max-flow cut and node
// Fiduccia-Mattheyses, leaves ordered by AMD.  perm[new] = old.
void nd_ordering(int n, const int* rowp, const int* cols, int* perm);

// From an analysed order S, a new order (perm[new] = old) with the columns of
// every fundamental supernode rearranged so that columns predicted from the
// values `vals` (threshold ustab) to form 2x2 pivots are adjacent.  Each
// supernode keeps its dense diagonal block and the structure below it; the
// new order is analysed again by the caller.
void pair_partners(int n, const int* rowp, const int* cols, const double* vals, double ustab,
                   const SymbolicResult& S, int* perm);

// Full symbolic analysis.  `perm` may be null (natural order).
void analyze(int n, const int* rowp, const int* cols, const int* perm,
             SymbolicResult& S);

}  // namespace kktldl
