// Host symbolic analysis for the KKT LDL^T engine.
//
// The ordering and the supernode partition must be bit-identical to the
// reference's SparseLDL (amigo include/sparse_ldl.h, include/block_amd.h),
// so every tie-breaking rule of those routines is kept: LIFO degree
// buckets, the in-place quotient-graph layout that the approximate degrees
// are computed from, the post-order that visits children in increasing
// index, and the supernode merge that sorts child positions by the width of
// the root of the supernode whose id equals that position.
#include "symbolic.hpp"

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <queue>

namespace kktldl {

using std::vector;

// ===========================================================================
// Approximate minimum degree (include/block_amd.h:27-452, nmult = 0).
// ===========================================================================
namespace {

// Quotient graph stored the way the reference stores it: row r occupies
// adj[start[r] .. start[r+1]) of which the first nel[r] entries are element
// ids (sorted) and the next (len[r] - nel[r]) entries are variable ids
// (sorted).  Rows are compacted in place and borrow free space from their
// neighbours when they grow.
class QuotientGraph {
 public:
  int n;
  vector<int> start, adj, len, nel;

  QuotientGraph(int n_, const int* rowp, const int* cols) : n(n_) {
    // Diagonal added, rows sorted and de-duplicated (ordering_utils.h:313-360,
    // 812-838); the buffer keeps the reference's nnz + n allocation.
    start.assign(n + 1, 0);
    adj.assign((size_t)rowp[n] + n + 8, 0);
    size_t w = 0;
    for (int r = 0; r < n; r++) {
      size_t r0 = w;
      adj[w++] = r;
      for (int p = rowp[r]; p < rowp[r + 1]; p++) adj[w++] = cols[p];
      std::sort(adj.begin() + r0, adj.begin() + w);
      w = r0 + (std::unique(adj.begin() + r0, adj.begin() + w) - (adj.begin() + r0));
      start[r + 1] = (int)w;
    }
    len.resize(n);
    nel.assign(n, 0);
    for (int r = 0; r < n; r++) len[r] = start[r + 1] - start[r];
  }

  int* row(int r) { return adj.data() + start[r]; }
  int vbeg(int r) const { return start[r] + nel[r]; }
  int vend(int r) const { return start[r] + len[r]; }
  int slack(int r) const { return start[r + 1] - start[r] - len[r]; }

  // Obtain `need` free slots at the end of row r (block_amd.h:531-581):
  // first squeeze rows r-1, r-2, ... towards the front, then rows r+1, ...
  // towards the back.
  void grow(int r, int need) {
    int got = 0;
    if (r > 0) {
      int s = r - 1;
      while (got < need && s >= 0) got += slack(s--);
      s++;
      int dst = start[s] + len[s];
      for (int i = s + 1; i <= r; i++) {
        int src = start[i];
        start[i] = dst;
        for (int k = 0; k < len[i]; k++) adj[dst++] = adj[src + k];
      }
    }
    if (got < need) {
      int s = r + 1;
      while (got < need && s < n) got += slack(s++);
      s--;
      if (s >= n) s = n - 1;
      int dst = start[s + 1] - 1;
      for (int i = s; i > r; i--) {
        for (int k = len[i] - 1; k >= 0; k--) adj[dst--] = adj[start[i] + k];
        start[i] = start[i + 1] - len[i];
      }
    }
  }
};

class DegreeLists {  // LIFO buckets, block_amd.h:687-812
 public:
  int n, low;
  vector<int> deg, head, fwd, back;
  DegreeLists(const QuotientGraph& g) : n(g.n) {
    deg.resize(n);
    head.assign(n + 1, -1);
    fwd.assign(n + 1, -1);
    back.assign(n + 1, -1);
    low = 0;
    for (int i = 0; i < n; i++) {
      deg[i] = g.len[i];
      low = (i == 0) ? deg[i] : std::min(low, deg[i]);
    }
    for (int i = 0; i < n; i++) link_front(i, deg[i]);
  }
  void link_front(int v, int d) {
    int h = head[d];
    if (h != -1) {
      back[h] = v;
      fwd[v] = h;
    }
    head[d] = v;
  }
  void unlink(int v) {
    int d = deg[v];
    if (head[d] == v) head[d] = fwd[v];
    if (fwd[v] != -1) back[fwd[v]] = back[v];
    if (back[v] != -1) fwd[back[v]] = fwd[v];
    fwd[v] = back[v] = -1;
  }
  void move(int v, int d) {
    low = std::min(low, d);
    unlink(v);
    link_front(v, d);
    deg[v] = d;
  }
  int first_nonempty() const {
    for (int d = low; d <= n; d++)
      if (head[d] != -1) return head[d];
    return -1;
  }
};

// Sort, drop negatives and duplicates (block_amd.h:459-477).
int sorted_set(int count, int* a) {
  std::sort(a, a + count);
  int lo = 0;
  while (lo < count && a[lo] < 0) lo++;
  int out = 0;
  for (int i = lo; i < count; i++) {
    if (i + 1 < count && a[i] == a[i + 1]) continue;
    a[out++] = a[i];
  }
  return out;
}

}  // namespace

void amd_ordering(int n, const int* rowp, const int* cols, int* perm) {
  QuotientGraph g(n, rowp, cols);
  DegreeLists dl(g);
  vector<int> weight(n, 1);      // supervariable size; -(p+1) once absorbed by p
  vector<int> elem_w(n, -1);     // cached |L_e \ L_p| work array
  vector<int> lp(n);             // pattern of the new element
  vector<int> slot(n, 0);        // used at the end to place absorbed variables
  vector<int>& A = g.adj;

  int order_len = 0;
  for (int eliminated = 0; eliminated < n; order_len++) {
    const int piv = dl.first_nonempty();
    dl.unlink(piv);
    perm[order_len] = piv;

    // --- new element pattern L_p = {piv} + adj vars + vars of adj elements
    int m = 0;
    lp[m++] = piv;
    for (int q = g.vbeg(piv); q < g.vend(piv); q++)
      if (A[q] != piv) lp[m++] = A[q];
    for (int q = g.start[piv]; q < g.vbeg(piv); q++) {
      int e = A[q];
      for (int t = g.vbeg(e); t < g.vend(e); t++) {
        if (m >= n) m = sorted_set(m, lp.data());
        if (A[t] != piv) lp[m++] = A[t];
      }
    }
    m = sorted_set(m, lp.data());

    // --- prune every neighbour: absorbed elements and members of L_p
    for (int j = 0; j < m; j++) {
      const int v = lp[j];
      if (v == piv) continue;
      int dead_e = 0, dead_v = 0;
      const int pe_end = g.vbeg(piv);
      int a = g.start[piv];
      for (int c = g.start[v]; a < pe_end && c < g.vbeg(v); c++) {
        while (a < pe_end && A[a] < A[c]) a++;
        if (a < pe_end && A[a] == A[c]) {
          A[c] = -1;
          dead_e++;
        }
      }
      int b = 0;
      for (int c = g.vbeg(v); b < m && c < g.vend(v); c++) {
        while (b < m && lp[b] < A[c]) b++;
        if (b < m && A[c] == lp[b]) {
          A[c] = -1;
          dead_v++;
        }
      }
      if (dead_e || dead_v) {
        const int end = g.vend(v);
        int rd = g.start[v], wr = rd;
        for (; rd < end; rd++, wr++) {
          while (rd < end && A[rd] == -1) rd++;
          if (rd < end) A[wr] = A[rd];
        }
        g.nel[v] -= dead_e;
        g.len[v] -= dead_e + dead_v;
      }
    }

    // --- make piv an element of every neighbour (sorted insert)
    for (int j = 0; j < m; j++) {
      const int v = lp[j];
      if (v == piv) continue;
      if (g.slack(v) == 0) g.grow(v, 1);
      int pos = g.start[v];
      const int e_end = g.vbeg(v);
      while (pos < e_end && A[pos] < piv) pos++;
      if (A[pos] != piv) {
        for (int q = g.vend(v) - 1; q >= pos; q--) A[q + 1] = A[q];
        A[pos] = piv;
        g.len[v]++;
        g.nel[v]++;
      }
    }

    // --- elements adjacent to piv are absorbed into it
    for (int q = g.start[piv]; q < g.vbeg(piv); q++) {
      int e = A[q];
      g.len[e] = 0;
      g.nel[e] = 0;
    }

    // --- row piv now stores L_p \ {piv}
    if (m - 1 > g.start[piv + 1] - g.start[piv])
      g.grow(piv, m - 1 - (g.start[piv + 1] - g.start[piv]));
    g.nel[piv] = 0;
    g.len[piv] = m - 1;
    {
      int w = g.start[piv];
      for (int k = 0; k < m; k++) {
        if (lp[k] == piv) {
          k++;
          if (k == m) break;
        }
        A[w++] = lp[k];
      }
    }

    // --- approximate degrees (block_amd.h:282-375)
    for (int j = 0; j < m; j++) {
      const int v = lp[j];
      for (int q = g.start[v]; q < g.vbeg(v); q++) {
        int e = A[q];
        if (elem_w[e] < 0) {
          elem_w[e] = 0;
          for (int t = g.start[e]; t < g.vend(e); t++) elem_w[e] += weight[A[t]];
        }
        elem_w[e] -= weight[v];
      }
    }
    int wl = 0;
    for (int j = 0; j < m; j++)
      if (lp[j] != piv) wl += weight[lp[j]];
    for (int j = 0; j < m; j++) {
      const int v = lp[j];
      if (v == piv) continue;
      int est = wl;
      for (int q = g.vbeg(v); q < g.vend(v); q++) est += weight[A[q]];
      const int nv = g.len[v] - g.nel[v];
      if (nv > 0 && std::binary_search(&A[g.vbeg(v)], &A[g.vbeg(v)] + nv, v))
        est -= weight[v];
      if (std::binary_search(lp.data(), lp.data() + m, v)) est -= weight[v];
      for (int q = g.start[v]; q < g.vbeg(v); q++) {
        int e = A[q];
        if (e == piv) continue;
        if (elem_w[e] < 0) {
          elem_w[e] = 0;
          for (int t = g.start[e]; t < g.vend(e); t++) elem_w[e] += weight[A[t]];
        }
        est += elem_w[e];
      }
      est = std::min(est, n - eliminated);
      dl.move(v, std::min(est, dl.deg[v] + wl));
    }
    for (int j = 0; j < m; j++) {
      const int v = lp[j];
      for (int q = g.start[v]; q < g.vbeg(v); q++) elem_w[A[q]] = -1;
    }

    // --- indistinguishable variables collapse into supervariables
    for (int j = 0; j < m; j++) {
      const int a = lp[j];
      if (a == piv || weight[a] < 0) continue;
      for (int k = j + 1; k < m; k++) {
        const int b = lp[k];
        if (b == piv || weight[b] < 0) continue;
        if (a == b || g.len[a] != g.len[b] || g.nel[a] != g.nel[b]) continue;
        // walk both rows, each skipping the other's id (block_amd.h:600-630)
        int ia = g.start[a], ib = g.start[b];
        const int ea = ia + g.len[a], eb = ib + g.len[b];
        for (int t = 0; t < g.len[a]; t++, ia++, ib++) {
          if (A[ia] == b) ia++;
          if (A[ib] == a) ib++;
          if (A[ia] != A[ib]) break;
        }
        if (!(ia == ea && ib == eb)) continue;
        weight[a] += weight[b];
        dl.move(a, dl.deg[a] - weight[b]);
        weight[b] = -(a + 1);
        // remove b from its elements' and neighbours' lists
        int q = g.start[b];
        for (; q < g.vbeg(b); q++) {
          int e = A[q];
          int x = g.start[e], xe = g.vend(e) - 1;
          while (x < xe && A[x] != b) x++;
          for (; x < xe; x++) A[x] = A[x + 1];
          g.len[e]--;
        }
        for (; q < g.vend(b); q++) {
          int w = A[q];
          int x = g.vbeg(w), xe = g.vend(w) - 1;
          while (x < xe && A[x] != b) x++;
          for (; x < xe; x++) A[x] = A[x + 1];
          g.len[w]--;
        }
        g.nel[b] = 0;
        g.len[b] = 0;
        dl.unlink(b);
      }
    }
    eliminated += weight[piv];
  }

  // Expand supervariables: principal variables in elimination order, each
  // followed by the variables it absorbed (block_amd.h:414-437).
  int end = n;
  for (int j = order_len - 1; j >= 0; j--) {
    int sv = perm[j];
    end -= weight[sv];
    perm[end] = sv;
    slot[sv] = end + 1;
  }
  for (int k = 0; k < n; k++) {
    if (weight[k] >= 0) continue;
    int p = k;
    while (weight[p] < 0) p = -(weight[p] + 1);
    perm[slot[p]++] = k;
  }
}

// ===========================================================================
// Elimination tree, post-order, column counts, supernodes.
// ===========================================================================
namespace {

// Liu's algorithm with path compression over the permuted pattern
// (sparse_ldl.h:2393-2458).
void elimination_tree(int n, const int* colp, const int* rows, const int* perm,
                      const int* iperm, vector<int>& parent) {
  vector<int> anc(n, -1);
  parent.assign(n, -1);
  for (int k = 0; k < n; k++) {
    const int ka = perm ? perm[k] : k;
    for (int p = colp[ka]; p < colp[ka + 1]; p++) {
      int i = iperm ? iperm[rows[p]] : rows[p];
      while (i < k) {
        int up = anc[i];
        anc[i] = k;
        if (up == -1) {
          parent[i] = k;
          break;
        }
        i = up;
      }
    }
  }
}

// Depth-first post-order, children visited in increasing index, roots in
// increasing index (sparse_ldl.h:2473-2536).
void tree_postorder(int n, const vector<int>& parent, vector<int>& post) {
  vector<int> first(n, -1), sib(n, -1), stk;
  for (int v = n - 1; v >= 0; v--)
    if (parent[v] != -1) {
      sib[v] = first[parent[v]];
      first[parent[v]] = v;
    }
  post.resize(n);
  int k = 0;
  stk.reserve(n);
  for (int r = 0; r < n; r++) {
    if (parent[r] != -1) continue;
    stk.push_back(r);
    while (!stk.empty()) {
      int p = stk.back();
      int c = first[p];
      if (c == -1) {
        post[k++] = p;
        stk.pop_back();
      } else {
        first[p] = sib[c];
        stk.push_back(c);
      }
    }
  }
}

// Column counts via the skeleton / least-common-ancestor method
// (Gilbert-Ng-Peyton; sparse_ldl.h:2545-2719).
void column_counts(int n, const int* colp, const int* rows, const int* perm,
                   const int* iperm, const vector<int>& parent,
                   const vector<int>& post, vector<int>& cc) {
  vector<int> first(n, -1), maxfirst(n, -1), prevleaf(n, -1), uf(n);
  cc.assign(n, 0);
  for (int k = 0; k < n; k++) {
    int j = post[k];
    if (first[j] == -1) cc[j] = 1;  // leaf of the etree
    for (; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
  }
  std::iota(uf.begin(), uf.end(), 0);
  auto root_of = [&](int p) {
    int r = p;
    while (uf[r] != r) r = uf[r];
    while (uf[p] != p) {
      int nx = uf[p];
      uf[p] = r;
      p = nx;
    }
    return r;
  };
  for (int k = 0; k < n; k++) {
    const int j = post[k];
    if (parent[j] != -1) cc[parent[j]]--;
    const int ja = perm ? perm[j] : j;
    for (int p = colp[ja]; p < colp[ja + 1]; p++) {
      const int i = iperm ? iperm[rows[p]] : rows[p];
      if (i <= j || first[j] <= maxfirst[i]) continue;
      cc[j]++;
      maxfirst[i] = first[j];
      const int jprev = prevleaf[i];
      prevleaf[i] = j;
      if (jprev != -1) cc[root_of(jprev)]--;
    }
    if (parent[j] != -1) uf[j] = parent[j];
  }
  for (int k = 0; k < n; k++) {
    int j = post[k];
    if (parent[j] != -1) cc[parent[j]] += cc[j];
  }
}

// Children lists with roots appended after position N (sparse_ldl.h:2978-3014).
void build_children(int N, const vector<int>& par, vector<int>& ptr,
                    vector<int>& ch) {
  ptr.assign(N + 2, 0);
  ch.assign(N, 0);
  for (int i = 0; i < N; i++) ptr[(par[i] != -1 ? par[i] : N) + 1]++;
  for (int i = 0; i <= N; i++) ptr[i + 1] += ptr[i];
  vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < N; i++) ch[fill[par[i] != -1 ? par[i] : N]++] = i;
}

}  // namespace

void analyze(int n, const int* rowp, const int* cols, const int* perm_in,
             SymbolicResult& S) {
  S = SymbolicResult();
  S.n = n;
  const int* perm = nullptr;
  const int* iperm = nullptr;
  if (perm_in) {
    S.permuted = true;
    S.perm.assign(perm_in, perm_in + n);
    S.iperm.assign(n, 0);
    for (int i = 0; i < n; i++) S.iperm[S.perm[i]] = i;
    perm = S.perm.data();
    iperm = S.iperm.data();
  }
  elimination_tree(n, rowp, cols, perm, iperm, S.parent);
  tree_postorder(n, S.parent, S.post);
  column_counts(n, rowp, cols, perm, iperm, S.parent, S.post, S.colcount);
  const vector<int>& parent = S.parent;
  const vector<int>& post = S.post;
  const vector<int>& cc = S.colcount;

  // Fundamental supernodes: maximal post-order chains v -> parent(v) with
  // colcount(parent) = colcount(v) - 1 (sparse_ldl.h:2734-2765).
  vector<int> sn_of(n), sn_vars(n), sn_ptr(1, 0);
  for (int i = 0; i < n;) {
    const int s = (int)sn_ptr.size() - 1;
    int v = post[i];
    sn_of[v] = s;
    sn_vars[i++] = v;
    while (i < n && parent[v] == post[i] && cc[post[i]] == cc[v] - 1) {
      v = post[i];
      sn_of[v] = s;
      sn_vars[i++] = v;
    }
    sn_ptr.push_back(i);
  }
  const int ns = (int)sn_ptr.size() - 1;
  S.num_snodes = ns;
  vector<int> sn_par(ns);
  for (int s = 0; s < ns; s++) {
    int p = parent[sn_vars[sn_ptr[s + 1] - 1]];
    sn_par[s] = p >= 0 ? sn_of[p] : -1;
  }
  vector<int> sn_cptr, sn_ch;
  build_children(ns, sn_par, sn_cptr, sn_ch);

  // Relaxed amalgamation: width <= 64, <= 256 explicit zeros per merge
  // chain (sparse_ldl.h:2786-2967).
  const int kMaxWidth = 64, kMaxZeros = 256;
  vector<int> rep(ns), rows_of(ns), width(ns), zeros(ns, 0);
  for (int s = 0; s < ns; s++) {
    rep[s] = s;
    width[s] = sn_ptr[s + 1] - sn_ptr[s];
    rows_of[s] = cc[sn_vars[sn_ptr[s]]];
  }
  auto top = [&](int s) {
    while (rep[s] != s) {
      rep[s] = rep[rep[s]];
      s = rep[s];
    }
    return s;
  };
  vector<int> order;
  for (int s = 0; s < ns; s++) {
    order.clear();
    for (int q = sn_cptr[s]; q < sn_cptr[s + 1]; q++) order.push_back(q);
    // Positions in the child array, keyed by width[top(position)].
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return width[top(a)] < width[top(b)]; });
    for (int q : order) {
      const int rp = top(s), rc = top(sn_ch[q]);
      const int w = width[rc] + width[rp];
      if (rc == rp || w > kMaxWidth) continue;
      const int extra = std::max(0, rows_of[rp] - (rows_of[rc] - width[rp]));
      const int z = extra * width[rc];
      if (zeros[rp] + zeros[rc] + z > kMaxZeros) continue;
      rows_of[rp] = w + rows_of[rp] - width[rp];
      width[rp] = w;
      rep[rc] = rp;
      zeros[rp] += z + zeros[rc];
    }
  }
  vector<int> node_of_sn(ns), new_id(ns, -1);
  int N = 0;
  for (int s = 0; s < ns; s++)
    if (top(s) == s) new_id[s] = N++;
  for (int s = 0; s < ns; s++) node_of_sn[s] = new_id[top(s)];
  S.num_nodes = N;
  S.node_ptr.assign(N + 1, 0);
  for (int s = 0; s < ns; s++) S.node_ptr[node_of_sn[s] + 1] += sn_ptr[s + 1] - sn_ptr[s];
  for (int k = 0; k < N; k++) S.node_ptr[k + 1] += S.node_ptr[k];
  S.node_to_var.assign(n, 0);
  vector<int> node_of_var(n), ncc(N), fill(S.node_ptr.begin(), S.node_ptr.end() - 1);
  for (int s = 0; s < ns; s++)
    for (int q = sn_ptr[s]; q < sn_ptr[s + 1]; q++) {
      int v = sn_vars[q];
      S.node_to_var[fill[node_of_sn[s]]++] = v;
      node_of_var[v] = node_of_sn[s];
    }
  for (int s = 0; s < ns; s++) ncc[node_of_sn[s]] = rows_of[top(s)];
  S.node_parent.assign(N, -1);
  for (int s = 0; s < ns; s++) {
    if (top(s) != s || sn_par[s] < 0) continue;
    int kp = node_of_sn[top(sn_par[s])];
    if (kp != node_of_sn[s]) S.node_parent[node_of_sn[s]] = kp;
  }
  build_children(N, S.node_parent, S.children_ptr, S.children);

  // Row structure below each node (sparse_ldl.h:2324-2342, 3036-3122):
  // rows discovered in increasing permuted index, stored as original ids.
  S.contrib_ptr.assign(N + 1, 0);
  for (int k = 0; k < N; k++)
    S.contrib_ptr[k + 1] = S.contrib_ptr[k] + ncc[k] - (S.node_ptr[k + 1] - S.node_ptr[k]);
  S.contrib_rows.assign(S.contrib_ptr[N], 0);
  {
    vector<int> used(N, 0), seen(N, -1), up(N, -1);
    for (int k = 0; k < N; k++) {
      int p = parent[S.node_to_var[S.node_ptr[k + 1] - 1]];
      up[k] = p >= 0 ? node_of_var[p] : -1;
    }
    for (int k = 0; k < n; k++) {
      const int kn = node_of_var[k];
      const int ka = perm ? perm[k] : k;
      for (int p = rowp[ka]; p < rowp[ka + 1]; p++) {
        const int i = iperm ? iperm[cols[p]] : cols[p];
        if (i >= k) continue;
        for (int s = node_of_var[i]; s >= 0 && s != kn && seen[s] != k; s = up[s]) {
          seen[s] = k;
          S.contrib_rows[S.contrib_ptr[s] + used[s]++] = ka;
        }
      }
    }
  }
  if (perm)
    for (int i = 0; i < n; i++) S.node_to_var[i] = S.perm[S.node_to_var[i]];
}

}  // namespace kktldl
