// ORACLE -- test infrastructure, not product code.
//
// A CPU restatement of the reference's KKT LDL^T path (amigo `SparseLDL`),
// used only as the checker: by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py. Nothing in amigo_amd/ may load or call it.
//
// What it restates (file:line in the reference):
//   ordering graph        include/ordering_utils.h:313-360, 812-838
//   BlockAMD (nmult = 0)  include/block_amd.h:27-452, 459-477, 531-681, 687-812
//   symbolic analysis     include/sparse_ldl.h:2222-2376 (etree 2393-2458,
//                         post-order 2473-2536, column counts 2545-2719,
//                         supernodes 2734-2765, merge 2786-2967, children
//                         2978-3014, row pattern 3036-3122)
//   numeric factor        include/sparse_ldl.h:767-1174 (multifrontal driver,
//                         frontal variables, assembly incl. delayed pivots),
//                         1270-1553 (blocked threshold 1x1/2x2 pivoting with
//                         delays), 1226-1248 (root: LAPACK dsytrf 'L'),
//                         2010-2044 (Cholesky mode)
//   solve                 include/sparse_ldl.h:1568-1926, 2053-2133
//   inertia               include/sparse_ldl.h:205-213, 1935-1987
// LAPACK dsytf2/dsytrs ('L', Bunch-Kaufman, alpha = (1+sqrt(17))/8) are
// restated from the published LAPACK algorithm; the reference calls them
// through MKL/OpenBLAS for root fronts.
//
// How each part is pinned (tests/test_oracle.py):
//   ordering / supernodes / row pattern: bit-exact against the compiled
//     reference (oracle/_ref) on the reference's own test matrices, on
//     seeded Poisson grids and on the Rosenbrock KKT;
//   etree / column counts: brute-force dense symbolic elimination;
//   numeric factor / solve / inertia: reference outputs, eigh inertia,
//     scipy.linalg.ldl on dense cases, and residual ||Kx-b|| checks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

using std::vector;

// ---------------------------------------------------------------------------
// Ordering graph: add the diagonal, sort and de-duplicate every row
// (ordering_utils.h:313-360 for AMD, 812-838).  The buffer keeps the
// reference's allocation size (nnz + n) because BlockAMD works in place.
// ---------------------------------------------------------------------------
static void amd_graph(int n, const int* rowp, const int* cols, vector<int>& gp,
                      vector<int>& gc) {
  gp.assign(n + 1, 0);
  gc.assign((size_t)rowp[n] + n + 8, 0);  // +8: slack for end-of-array reads
  int ptr = 0;
  for (int i = 0; i < n; i++) {
    gc[ptr++] = i;
    for (int jp = rowp[i]; jp < rowp[i + 1]; jp++) gc[ptr++] = cols[jp];
    gp[i + 1] = ptr;
  }
  int start = gp[0];
  for (int i = 0; i < n; i++) {
    int size = gp[i + 1] - start;
    int* a = &gc[start];
    std::sort(a, a + size);
    int m = 0;
    if (size > 0) {
      gc[gp[i]] = a[0];
      m = 1;
      for (int r = 1; r < size; r++)
        if (a[r] != gc[gp[i] + m - 1]) gc[gp[i] + m++] = a[r];
    }
    start = gp[i + 1];
    gp[i + 1] = gp[i] + m;
  }
}

// ---------------------------------------------------------------------------
// BlockAMD restatement.  Every step acts on the same in-place quotient-graph
// layout as the reference (element list first, then variable list, inside
// each row's slot of `cols`), because the degree estimates and the LIFO
// degree buckets depend on that exact state.
// ---------------------------------------------------------------------------
struct DegreeBuckets {  // block_amd.h:687-812
  int size, min_degree;
  vector<int> deg, head, nxt, prv;
  DegreeBuckets(int n, const int* rowp) : size(n) {
    deg.resize(n);
    head.assign(n + 1, -1);
    nxt.assign(n + 1, -1);
    prv.assign(n + 1, -1);
    min_degree = 0;
    for (int i = 0; i < n; i++) {
      deg[i] = rowp[i + 1] - rowp[i];
      if (i == 0 || deg[i] < min_degree) min_degree = deg[i];
    }
    for (int i = 0; i < n; i++) push(i, deg[i]);
  }
  void push(int i, int d) {
    if (head[d] != -1) {
      prv[head[d]] = i;
      nxt[i] = head[d];
    }
    head[d] = i;
  }
  void unlink(int i) {
    int d = deg[i];
    if (head[d] == i) head[d] = nxt[i];
    if (nxt[i] != -1) prv[nxt[i]] = prv[i];
    if (prv[i] != -1) nxt[prv[i]] = nxt[i];
    prv[i] = nxt[i] = -1;
  }
  void set(int i, int d) {
    if (d < min_degree) min_degree = d;
    unlink(i);
    push(i, d);
    deg[i] = d;
  }
  int pick() const {
    for (int d = min_degree; d <= size; d++)
      if (head[d] != -1) return head[d];
    return -1;
  }
};

static int sort_unique_nonneg(int len, int* a) {  // block_amd.h:459-477
  std::sort(a, a + len);
  int i = 0, j = 0;
  while (i < len && a[i] < 0) i++;
  for (; i < len; i++, j++) {
    while (i < len - 1 && a[i] == a[i + 1]) i++;
    if (i != j) a[j] = a[i];
  }
  return j;
}

// Free `need` slots at the end of row r by compacting neighbouring rows
// (block_amd.h:531-581).
static void make_room(int n, int* rowp, int* cols, const int* alen, int r,
                      int need) {
  int got = 0;
  if (r > 0) {
    int s = r - 1;
    for (; got < need && s >= 0; s--) got += (rowp[s + 1] - rowp[s]) - alen[s];
    s++;
    int dst = rowp[s] + alen[s];
    for (int i = s + 1; i <= r; i++) {
      int src = rowp[i];
      int nr = dst;
      int end = src + alen[i];
      for (; src < end; src++, dst++) cols[dst] = cols[src];
      rowp[i] = nr;
    }
  }
  if (got < need) {
    int s = r + 1;
    for (; got < need && s < n; s++) got += (rowp[s + 1] - rowp[s]) - alen[s];
    s--;
    if (s >= n) s = n - 1;
    int dst = rowp[s + 1] - 1;
    for (int i = s; i > r; i--) {
      int lo = rowp[i];
      int src = lo + alen[i] - 1;
      for (; src >= lo; src--, dst--) cols[dst] = cols[src];
      rowp[i] = rowp[i + 1] - alen[i];
    }
  }
}

// Indistinguishable-variable test (block_amd.h:600-630), including its
// reads one past a row end, which the padded buffer keeps in bounds.
static bool same_adjacency(int i, int j, const int* elen, const int* alen,
                           const int* rowp, const int* cols) {
  if (i == j) return false;
  if (alen[i] != alen[j] || elen[i] != elen[j]) return false;
  int ip = rowp[i], iend = ip + alen[i];
  int jp = rowp[j], jend = jp + alen[j];
  for (int k = 0; k < alen[i]; k++, ip++, jp++) {
    if (cols[ip] == j) ip++;
    if (cols[jp] == i) jp++;
    if (cols[ip] != cols[jp]) break;
  }
  return ip == iend && jp == jend;
}

static void drop_variable(int v, int* elen, int* alen, const int* rowp,
                          int* cols) {  // block_amd.h:637-681
  int p = rowp[v];
  for (; p < rowp[v] + elen[v]; p++) {
    int e = cols[p];
    int q = rowp[e], qend = rowp[e] + alen[e] - 1;
    for (; q < qend; q++)
      if (cols[q] == v) break;
    for (; q < qend; q++) cols[q] = cols[q + 1];
    alen[e]--;
  }
  for (; p < rowp[v] + alen[v]; p++) {
    int w = cols[p];
    int q = rowp[w] + elen[w], qend = rowp[w] + alen[w] - 1;
    for (; q < qend; q++)
      if (cols[q] == v) break;
    for (; q < qend; q++) cols[q] = cols[q + 1];
    alen[w]--;
  }
  elen[v] = 0;
  alen[v] = 0;
}

// AMD with approximate degrees and no multiplier constraint, as SparseLDL
// calls it (ordering_utils.h:202-221 with nmult = 0).  perm[new] = old.
static void amd_order(int n, vector<int>& gp, vector<int>& gc, int* perm) {
  int* rowp = gp.data();
  int* cols = gc.data();
  vector<int> alen(n), elen(n, 0), slen(n, 1), edeg(n, -1), Lp(n), state(n, 1);
  for (int i = 0; i < n; i++) {
    perm[i] = i;
    alen[i] = rowp[i + 1] - rowp[i];
  }
  DegreeBuckets dl(n, rowp);

  int nsv = 0;
  for (int i = 0; i < n; nsv++) {
    int piv = dl.pick();
    dl.unlink(piv);
    perm[nsv] = piv;
    state[piv] = 0;

    // Pattern of the new element: row piv's variables plus the variables
    // of every element adjacent to piv.
    int len = 0;
    Lp[len++] = piv;
    for (int j = rowp[piv] + elen[piv]; j < rowp[piv] + alen[piv]; j++)
      if (cols[j] != piv) Lp[len++] = cols[j];
    for (int j = rowp[piv]; j < rowp[piv] + elen[piv]; j++) {
      int e = cols[j];
      for (int k = rowp[e] + elen[e]; k < rowp[e] + alen[e]; k++) {
        if (len >= n) len = sort_unique_nonneg(len, Lp.data());
        if (cols[k] != piv) Lp[len++] = cols[k];
      }
    }
    len = sort_unique_nonneg(len, Lp.data());

    // Absorb piv's elements and drop Lp from every neighbour's lists.
    for (int j = 0; j < len; j++) {
      int v = Lp[j];
      if (v == piv) continue;
      int nre = 0;
      int pend = rowp[piv] + elen[piv];
      for (int k = rowp[piv], c = rowp[v]; k < pend && c < rowp[v] + elen[v];
           c++) {
        while (k < pend && cols[k] < cols[c]) k++;
        if (k < pend && cols[k] == cols[c]) {
          cols[c] = -1;
          nre++;
        }
      }
      int nrv = 0;
      for (int k = 0, c = rowp[v] + elen[v]; k < len && c < rowp[v] + alen[v];
           c++) {
        while (k < len && Lp[k] < cols[c]) k++;
        if (k < len && cols[c] == Lp[k]) {
          cols[c] = -1;
          nrv++;
        }
      }
      if (nre > 0 || nrv > 0) {
        int end = rowp[v] + alen[v];
        int k = rowp[v], w = k;
        for (; k < end; k++, w++) {
          while (k < end && cols[k] == -1) k++;
          if (k < end) cols[w] = cols[k];
        }
        elen[v] -= nre;
        alen[v] -= nre + nrv;
      }
    }

    // piv becomes an element of every neighbour (sorted insertion).
    for (int j = 0; j < len; j++) {
      int v = Lp[j];
      if (v == piv) continue;
      if (alen[v] == rowp[v + 1] - rowp[v])
        make_room(n, rowp, cols, alen.data(), v, 1);
      int k = rowp[v], end = k + elen[v];
      while (k < end && cols[k] < piv) k++;
      if (cols[k] != piv) {
        for (int p = rowp[v] + alen[v] - 1; p >= k; p--) cols[p + 1] = cols[p];
        cols[k] = piv;
        alen[v]++;
        elen[v]++;
      }
    }

    for (int j = rowp[piv]; j < rowp[piv] + elen[piv]; j++) {
      int e = cols[j];
      alen[e] = 0;
      elen[e] = 0;
      state[e] = -1;
    }

    int rsize = rowp[piv + 1] - rowp[piv];
    if (len - 1 > rsize)
      make_room(n, rowp, cols, alen.data(), piv,
                len - 1 - (rowp[piv + 1] - rowp[piv]));
    elen[piv] = 0;
    alen[piv] = len - 1;
    for (int j = rowp[piv], k = 0; k < len; k++, j++) {
      if (Lp[k] == piv) {
        k++;
        if (k == len) break;
      }
      cols[j] = Lp[k];
    }

    // Approximate degree update (block_amd.h:282-375).
    for (int j = 0; j < len; j++) {
      int v = Lp[j];
      for (int k = rowp[v]; k < rowp[v] + elen[v]; k++) {
        int e = cols[k];
        if (edeg[e] < 0) {
          edeg[e] = 0;
          for (int p = rowp[e]; p < rowp[e] + alen[e]; p++) edeg[e] += slen[cols[p]];
        }
        edeg[e] -= slen[v];
      }
    }
    int degLp = 0;
    for (int j = 0; j < len; j++)
      if (Lp[j] != piv) degLp += slen[Lp[j]];
    for (int j = 0; j < len; j++) {
      int v = Lp[j];
      if (v == piv) continue;
      int est = degLp;
      for (int k = rowp[v] + elen[v]; k < rowp[v] + alen[v]; k++) est += slen[cols[k]];
      int s0 = rowp[v] + elen[v], sl = alen[v] - elen[v];
      if (sl > 0 && std::binary_search(&cols[s0], &cols[s0] + sl, v)) est -= slen[v];
      if (std::binary_search(Lp.data(), Lp.data() + len, v)) est -= slen[v];
      for (int k = rowp[v]; k < rowp[v] + elen[v]; k++) {
        int e = cols[k];
        if (e == piv) continue;
        if (edeg[e] < 0) {
          edeg[e] = 0;
          for (int p = rowp[e]; p < rowp[e] + alen[e]; p++) edeg[e] += slen[cols[p]];
        }
        est += edeg[e];
      }
      est = est < n - i ? est : n - i;
      int d = dl.deg[v];
      d = est < d + degLp ? est : d + degLp;
      dl.set(v, d);
    }
    for (int j = 0; j < len; j++) {
      int v = Lp[j];
      for (int k = rowp[v]; k < rowp[v] + elen[v]; k++) edeg[cols[k]] = -1;
    }

    // Supervariable detection among the members of Lp.
    for (int j = 0; j < len; j++) {
      int a = Lp[j];
      if (a == piv || slen[a] < 0) continue;
      for (int k = j + 1; k < len; k++) {
        int b = Lp[k];
        if (b == piv || slen[b] < 0) continue;
        if (same_adjacency(a, b, elen.data(), alen.data(), rowp, cols)) {
          slen[a] += slen[b];
          dl.set(a, dl.deg[a] - slen[b]);
          slen[b] = -(a + 1);
          state[b] = -1;
          drop_variable(b, elen.data(), alen.data(), rowp, cols);
          dl.unlink(b);
        }
      }
    }
    i += slen[piv];
  }

  // Expand supervariables (block_amd.h:414-437).
  int end = n;
  for (int j = nsv - 1; j >= 0; j--) {
    int sv = perm[j];
    end -= slen[sv];
    perm[end] = sv;
    state[sv] = end + 1;
  }
  for (int k = 0; k < n; k++) {
    if (slen[k] < 0) {
      int j = k;
      while (slen[j] < 0) j = -(slen[j] + 1);
      perm[state[j]] = k;
      state[j]++;
    }
  }
}

// ---------------------------------------------------------------------------
// Symbolic analysis (sparse_ldl.h:2222-2376).  Index conventions: the
// elimination tree, post-order and column counts live in the permuted index
// space; node_to_var and contrib_rows hold original indices.
// ---------------------------------------------------------------------------
struct Symbolic {
  int n = 0;
  bool permuted = false;
  vector<int> perm, iperm;
  vector<int> parent, post, colcount;  // fundamental structure
  int num_snodes = 0;
  int num_nodes = 0;
  vector<int> node_ptr, node_to_var, children_ptr, children;
  vector<int> contrib_ptr, contrib_rows;
  int max_contrib = 0;
};

static void etree(int n, const int* colp, const int* rows, const int* perm,
                  const int* iperm, int* parent) {
  vector<int> anc(n, -1);
  std::fill(parent, parent + n, -1);
  for (int k = 0; k < n; k++) {
    int ka = perm ? perm[k] : k;
    for (int p = colp[ka]; p < colp[ka + 1]; p++) {
      int i = iperm ? iperm[rows[p]] : rows[p];
      while (i < k) {
        int t = anc[i];
        anc[i] = k;
        if (t == -1) {
          parent[i] = k;
          break;
        }
        i = t;
      }
    }
  }
}

static void postorder(int n, const int* parent, int* post) {
  vector<int> head(n, -1), next(n, -1), stack(n);
  for (int j = n - 1; j >= 0; j--)
    if (parent[j] != -1) {
      next[j] = head[parent[j]];
      head[parent[j]] = j;
    }
  int k = 0;
  for (int r = 0; r < n; r++) {
    if (parent[r] != -1) continue;
    int top = 0;
    stack[0] = r;
    while (top >= 0) {
      int p = stack[top];
      int c = head[p];
      if (c == -1) {
        post[k++] = p;
        top--;
      } else {
        head[p] = next[c];
        stack[++top] = c;
      }
    }
  }
}

static int find_root_compress(int p, int* anc) {
  int r = p;
  while (anc[r] != r) r = anc[r];
  while (anc[p] != p) {
    int nx = anc[p];
    anc[p] = r;
    p = nx;
  }
  return r;
}

// Column counts of L including the diagonal (sparse_ldl.h:2545-2719).
static void column_counts(int n, const int* colp, const int* rows,
                          const int* post, const int* perm, const int* iperm,
                          const int* parent, int* cc) {
  vector<int> first(n, -1), prevleaf(n, -1), maxfirst(n, -1), anc(n);
  std::fill(cc, cc + n, 0);
  for (int k = 0; k < n; k++) {
    int j = post[k];
    if (first[j] == -1) cc[j] = 1;
    while (j != -1 && first[j] == -1) {
      first[j] = k;
      j = parent[j];
    }
  }
  std::iota(anc.begin(), anc.end(), 0);
  for (int k = 0; k < n; k++) {
    int j = post[k];
    int ja = perm ? perm[j] : j;
    if (parent[j] != -1) cc[parent[j]]--;
    for (int p = colp[ja]; p < colp[ja + 1]; p++) {
      int i = iperm ? iperm[rows[p]] : rows[p];
      if (i <= j) continue;
      if (first[j] <= maxfirst[i]) continue;
      cc[j]++;
      maxfirst[i] = first[j];
      int jprev = prevleaf[i];
      prevleaf[i] = j;
      if (jprev != -1) cc[find_root_compress(jprev, anc.data())]--;
    }
    if (parent[j] != -1) anc[j] = parent[j];
  }
  for (int k = 0; k < n; k++) {
    int j = post[k];
    if (parent[j] != -1) cc[parent[j]] += cc[j];
  }
}

static void children_of(int nn, const int* par, vector<int>& cptr,
                        vector<int>& ch) {  // sparse_ldl.h:2978-3014
  cptr.assign(nn + 2, 0);
  ch.assign(nn, 0);
  for (int i = 0; i < nn; i++) {
    if (par[i] != -1) cptr[par[i] + 1]++;
    else cptr[nn + 1]++;
  }
  for (int i = 0; i < nn + 1; i++) cptr[i + 1] += cptr[i];
  for (int i = 0; i < nn; i++) {
    int p = par[i] != -1 ? par[i] : nn;
    ch[cptr[p]++] = i;
  }
  for (int i = nn; i >= 0; i--) cptr[i + 1] = cptr[i];
  cptr[0] = 0;
}

static void analyze(Symbolic& S, int n, const int* colp, const int* rows,
                    const int* perm_in) {
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
  S.parent.assign(n, -1);
  S.post.assign(n, 0);
  S.colcount.assign(n, 0);
  etree(n, colp, rows, perm, iperm, S.parent.data());
  postorder(n, S.parent.data(), S.post.data());
  column_counts(n, colp, rows, S.post.data(), perm, iperm, S.parent.data(),
                S.colcount.data());
  const int* parent = S.parent.data();
  const int* post = S.post.data();
  const int* cc = S.colcount.data();

  // Fundamental supernodes along post-order chains (sparse_ldl.h:2734-2765).
  vector<int> v2s(n), s2v(n);
  int ns = 0;
  for (int i = 0; i < n; ns++) {
    int v = post[i];
    v2s[v] = ns;
    s2v[i] = v;
    i++;
    int nv = i < n ? post[i] : n;
    while (i < n && parent[v] == nv && cc[nv] == cc[v] - 1) {
      v = nv;
      v2s[v] = ns;
      s2v[i] = v;
      i++;
      if (i < n) nv = post[i];
    }
  }
  S.num_snodes = ns;
  vector<int> sptr(ns + 1, 0);
  for (int i = 0; i < n; i++) sptr[v2s[i] + 1]++;
  for (int i = 0; i < ns; i++) sptr[i + 1] += sptr[i];
  vector<int> spar(ns);
  for (int s = 0; s < ns; s++) {
    int p = parent[s2v[sptr[s + 1] - 1]];
    spar[s] = p >= 0 ? v2s[p] : -1;
  }
  vector<int> scp, sch;
  children_of(ns, spar.data(), scp, sch);

  // Relaxed amalgamation (sparse_ldl.h:2786-2967), max width 64, at most
  // 256 added zeros.  The children are sorted as *positions* into the child
  // array, compared by the width of the root of the supernode whose id
  // equals that position -- the reference's comparator, kept as is.
  const int max_width = 64, max_zeros = 256;
  vector<int> root_of(ns), rcc(ns), width(ns), nz(ns, 0);
  for (int s = 0; s < ns; s++) {
    width[s] = sptr[s + 1] - sptr[s];
    rcc[s] = cc[s2v[sptr[s]]];
    root_of[s] = s;
  }
  auto find = [&](int s) {
    while (root_of[s] != s) {
      root_of[s] = root_of[root_of[s]];
      s = root_of[s];
    }
    return s;
  };
  for (int s = 0; s < ns; s++) {
    vector<int> kids;
    for (int ip = scp[s]; ip < scp[s + 1]; ip++) kids.push_back(ip);
    std::sort(kids.begin(), kids.end(),
              [&](int a, int b) { return width[find(a)] < width[find(b)]; });
    for (int ip : kids) {
      int child = sch[ip];
      int rp = find(s), rc = find(child);
      int mw = width[rc] + width[rp];
      if (rc != rp && mw <= max_width) {
        int extra = std::max(0, rcc[rp] - (rcc[rc] - width[rp]));
        int zadd = extra * width[rc];
        if (nz[rp] + nz[rc] + zadd <= max_zeros) {
          rcc[rp] = mw + rcc[rp] - width[rp];
          width[rp] = mw;
          root_of[rc] = rp;
          nz[rp] += zadd + nz[rc];
        }
      }
    }
  }
  vector<int> newid(ns, -1), s2n(ns);
  int nn = 0;
  for (int s = 0; s < ns; s++)
    if (find(s) == s) newid[s] = nn++;
  for (int s = 0; s < ns; s++) s2n[s] = newid[find(s)];
  S.num_nodes = nn;
  S.node_ptr.assign(nn + 1, 0);
  S.node_to_var.assign(n, 0);
  vector<int> var2node(n), ncc(nn), npar(nn, -1);
  for (int s = 0; s < ns; s++) S.node_ptr[s2n[s] + 1] += sptr[s + 1] - sptr[s];
  for (int k = 0; k < nn; k++) S.node_ptr[k + 1] += S.node_ptr[k];
  {
    vector<int> fill(S.node_ptr.begin(), S.node_ptr.end() - 1);
    for (int s = 0; s < ns; s++)
      for (int p = sptr[s]; p < sptr[s + 1]; p++) {
        int v = s2v[p];
        S.node_to_var[fill[s2n[s]]++] = v;
        var2node[v] = s2n[s];
      }
  }
  for (int s = 0; s < ns; s++) ncc[s2n[s]] = rcc[find(s)];
  for (int s = 0; s < ns; s++) {
    if (find(s) != s) continue;
    int k = s2n[s];
    if (spar[s] >= 0) {
      int kp = s2n[find(spar[s])];
      if (kp != k) npar[k] = kp;
    }
  }
  children_of(nn, npar.data(), S.children_ptr, S.children);

  S.contrib_ptr.assign(nn + 1, 0);
  for (int k = 0; k < nn; k++)
    S.contrib_ptr[k + 1] = ncc[k] - (S.node_ptr[k + 1] - S.node_ptr[k]);
  for (int k = 0; k < nn; k++) S.contrib_ptr[k + 1] += S.contrib_ptr[k];
  S.max_contrib = 0;
  for (int k = 0; k < nn; k++)
    S.max_contrib = std::max(S.max_contrib, S.contrib_ptr[k + 1] - S.contrib_ptr[k]);

  // Row pattern below each node, discovered in increasing permuted row
  // order (sparse_ldl.h:3036-3122), stored as original indices.
  S.contrib_rows.assign(S.contrib_ptr[nn], 0);
  {
    vector<int> cnt(nn, 0), flag(nn, -1), npar2(nn, -1);
    for (int k = 0; k < nn; k++) {
      int last = S.node_to_var[S.node_ptr[k + 1] - 1];
      int p = parent[last];
      npar2[k] = p >= 0 ? var2node[p] : -1;
    }
    for (int k = 0; k < n; k++) {
      int ks = var2node[k];
      int ka = perm ? perm[k] : k;
      for (int p = colp[ka]; p < colp[ka + 1]; p++) {
        int i = iperm ? iperm[rows[p]] : rows[p];
        if (i >= k) continue;
        for (int s = var2node[i]; s >= 0 && s != ks; s = npar2[s]) {
          if (flag[s] == k) break;
          flag[s] = k;
          S.contrib_rows[S.contrib_ptr[s] + cnt[s]++] = ka;
        }
      }
    }
  }
  if (perm)
    for (int i = 0; i < n; i++) S.node_to_var[i] = S.perm[S.node_to_var[i]];
}

// ---------------------------------------------------------------------------
// Numeric multifrontal factorization.
// ---------------------------------------------------------------------------
struct NodeFactor {
  int npiv = 0, ndel = 0, csize = 0;
  bool root = false;      // factored by the Bunch-Kaufman root path
  vector<int> pivots;     // negative-encoded -(v+1) for 2x2 (non-root)
  vector<int> delayed;
  vector<int> ipiv;       // LAPACK convention, roots only
  vector<double> L;       // (npiv + ndel + csize) x npiv, column major
};

struct Contribution {
  int ndel = 0, size = 0;
  vector<int> vars;       // delayed first, then the nominal rows
  vector<double> block;   // packed lower triangle, column major
};

struct Solver {
  Symbolic S;
  bool cholesky = false;
  double ustab = 0.01, pivot_tol = 1e-14;
  vector<NodeFactor> fac;
  vector<Contribution> con;
  vector<int> findex;  // front position of each original variable, -1 if none
  int info = 0;
};

// Symmetric interchange of rows/columns k < c of the lower triangle of F
// (sparse_ldl.h:1270-1292).
static void sym_swap(int* vars, double* F, int ldf, int k, int c) {
  std::swap(vars[k], vars[c]);
  for (int j = 0; j < k; j++) std::swap(F[k + j * ldf], F[c + j * ldf]);
  std::swap(F[k + k * ldf], F[c + c * ldf]);
  for (int j = k + 1; j < c; j++) std::swap(F[j + k * ldf], F[c + j * ldf]);
  for (int i = c + 1; i < ldf; i++) std::swap(F[i + k * ldf], F[i + c * ldf]);
}

static int first_absmax(int n, const double* x, int inc) {  // idamax, 0-based
  int im = 0;
  double m = -1.0;
  for (int i = 0; i < n; i++) {
    double a = std::fabs(x[(size_t)i * inc]);
    if (a > m) {
      m = a;
      im = i;
    }
  }
  return im;
}

// Blocked threshold LDL^T of one front (sparse_ldl.h:1316-1553).
// Pivot kinds chosen in fronts of more than 96 rows by the last factor:
// 1x1 at the candidate, 2x2 with the next candidate, 1x1 at a later row,
// 2x2 with a later row, delay (diagnostic counters for orderings).
static long long g_pivot_kinds[5];

static void factor_front(Solver& s, int ks, int fs, int m, int* vars,
                         double* F) {
  const int ldf = m, ldw = m, n = m;
  const int nb = std::min(64, m);
  vector<double> Wv((size_t)m * nb, 0.0);
  double* W = Wv.data();
  const double ustab = s.ustab, ptol = s.pivot_tol;
  int nc = fs;
  for (int k = 0; k < nc;) {
    int kstart = k, kw = 0;
    for (; kw < nb - 1 && k < nc;) {
      int step = 0, kpiv = -1;
      for (int i = k; i < n; i++) W[i + kw * ldw] = F[i + k * ldf];
      if (kw > 0)
        for (int c = 0; c < kw; c++) {
          double w = W[k + c * ldw];
          const double* Fc = &F[(size_t)(kstart + c) * ldf];
          for (int i = k; i < n; i++) W[i + kw * ldw] -= Fc[i] * w;
        }
      double awkk = std::fabs(W[k + kw * ldw]);
      double awik = 0.0;
      int imax = 0;
      if (k + 1 < nc) {
        imax = k + 1 + first_absmax(nc - (k + 1), &W[k + 1 + kw * ldw], 1);
        awik = std::fabs(W[imax + kw * ldw]);
      }
      int ismax = nc + first_absmax(n - nc, &W[nc + kw * ldw], 1);
      double colmax = std::max(awik, std::fabs(W[ismax + kw * ldw]));
      if (awkk >= ptol && awkk >= ustab * colmax) {
        step = 1;
        kpiv = k;
      } else if (k + 1 < nc) {
        double* W1 = &W[(kw + 1) * ldw];
        for (int j = k; j < imax; j++) W1[j] = F[imax + j * ldf];
        for (int i = imax; i < n; i++) W1[i] = F[i + imax * ldf];
        if (kw > 0)
          for (int c = 0; c < kw; c++) {
            double w = W[imax + c * ldw];
            const double* Fc = &F[(size_t)(kstart + c) * ldf];
            for (int i = k; i < n; i++) W1[i] -= Fc[i] * w;
          }
        int jmax = k + first_absmax(imax - k, &W1[k], 1);
        double rowmax = std::fabs(W1[jmax]);
        if (imax + 1 < n) {
          jmax = imax + 1 + first_absmax(n - (imax + 1), &W1[imax + 1], 1);
          rowmax = std::max(rowmax, std::fabs(W1[jmax]));
        }
        double awii = std::fabs(W1[imax]);
        if (awii >= ptol && awii >= ustab * rowmax) {
          for (int i = k; i < n; i++) W[i + kw * ldw] = W1[i];
          step = 1;
          kpiv = imax;
        } else {
          double wkk = W[k + kw * ldw], wik = W[imax + kw * ldw], wii = W1[imax];
          double det = std::fabs(wkk * wii - wik * wik);
          if (det >= ptol) {
            double delta = det / std::max(awii + awik, awkk + awik);
            if (delta >= ustab * colmax && delta >= ustab * rowmax) {
              step = 2;
              kpiv = imax;
            }
          }
        }
      }
      if (step == 0) {  // delay candidate k
        if (m > 96) g_pivot_kinds[4]++;
        if (k < nc - 1) {
          sym_swap(vars, F, ldf, k, nc - 1);
          for (int c = 0; c < kw; c++) std::swap(W[k + c * ldw], W[nc - 1 + c * ldw]);
        }
        nc -= 1;
        continue;
      }
      if (m > 96) g_pivot_kinds[step == 1 ? (kpiv == k ? 0 : 2) : (kpiv == k + 1 ? 1 : 3)]++;
      int ksrc = k + step - 1;
      if (kpiv != ksrc) {
        sym_swap(vars, F, ldf, ksrc, kpiv);
        for (int c = 0; c < kw + step; c++)
          std::swap(W[ksrc + c * ldw], W[kpiv + c * ldw]);
      }
      if (step == 1) {
        for (int i = k; i < n; i++) F[i + k * ldf] = W[i + kw * ldw];
        double dinv = 1.0 / F[k * (ldf + 1)];
        for (int i = k + 1; i < n; i++) F[i + k * ldf] *= dinv;
      } else {
        double d11 = W[k + kw * ldw], d21 = W[k + 1 + kw * ldw],
               d22 = W[k + 1 + (kw + 1) * ldw];
        F[k * (ldf + 1)] = d11;
        F[k * (ldf + 1) + 1] = d21;
        F[(k + 1) * (ldf + 1)] = d22;
        double detinv = 1.0 / (d11 * d22 - d21 * d21);
        double i00 = d22 * detinv, i01 = -d21 * detinv, i11 = d11 * detinv;
        for (int i = k + 2; i < n; i++) {
          double a = W[i + kw * ldw], b = W[i + (kw + 1) * ldw];
          F[i + k * ldf] = a * i00 + b * i01;
          F[i + (k + 1) * ldf] = a * i01 + b * i11;
        }
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
      }
      k += step;
      kw += step;
    }
    // Trailing update of the lower triangle: F[k:,k:] -= L_panel * W_panel^T
    for (int j = k; j < n; j++)
      for (int c = 0; c < kw; c++) {
        double w = W[j + c * ldw];
        const double* Lc = &F[(size_t)(kstart + c) * ldf];
        double* Fj = &F[(size_t)j * ldf];
        for (int i = j; i < n; i++) Fj[i] -= Lc[i] * w;
      }
  }
  int npiv = nc, ndel = fs - npiv, csize = m - fs;
  Contribution& C = s.con[ks];
  C.ndel = ndel;
  C.size = m - npiv;
  C.vars.assign(vars + npiv, vars + m);
  C.block.resize((size_t)C.size * (C.size + 1) / 2);
  size_t p = 0;
  for (int j = npiv; j < m; j++)
    for (int i = j; i < m; i++) C.block[p++] = F[i + (size_t)j * m];
  NodeFactor& nf = s.fac[ks];
  nf.npiv = npiv;
  nf.ndel = ndel;
  nf.csize = csize;
  nf.pivots.assign(vars, vars + npiv);
  nf.delayed.assign(vars + npiv, vars + fs);
  nf.L.assign(F, F + (size_t)m * npiv);
}

// LAPACK dsytf2 'L' (Bunch-Kaufman), restated from the published algorithm.
static int bk_factor(int n, double* A, int lda, int* ipiv) {
  const double alpha = (1.0 + std::sqrt(17.0)) / 8.0;
  int info = 0;
  int k = 0;
  while (k < n) {
    int kstep = 1, kp;
    double absakk = std::fabs(A[k + k * lda]);
    int imax = k;
    double colmax = 0.0;
    if (k < n - 1) {
      imax = k + 1 + first_absmax(n - k - 1, &A[k + 1 + k * lda], 1);
      colmax = std::fabs(A[imax + k * lda]);
    }
    if (std::max(absakk, colmax) == 0.0 || std::isnan(absakk)) {
      if (info == 0) info = k + 1;
      kp = k;
    } else {
      if (absakk >= alpha * colmax) {
        kp = k;
      } else {
        int jmax = k + first_absmax(imax - k, &A[imax + k * lda], lda);
        double rowmax = std::fabs(A[imax + jmax * lda]);
        if (imax < n - 1) {
          jmax = imax + 1 + first_absmax(n - imax - 1, &A[imax + 1 + imax * lda], 1);
          rowmax = std::max(rowmax, std::fabs(A[jmax + imax * lda]));
        }
        if (absakk >= alpha * colmax * (colmax / rowmax)) {
          kp = k;
        } else if (std::fabs(A[imax + imax * lda]) >= alpha * rowmax) {
          kp = imax;
        } else {
          kp = imax;
          kstep = 2;
        }
      }
      int kk = k + kstep - 1;
      if (kp != kk) {
        for (int i = kp + 1; i < n; i++) std::swap(A[i + kk * lda], A[i + kp * lda]);
        for (int j = kk + 1; j < kp; j++) std::swap(A[j + kk * lda], A[kp + j * lda]);
        std::swap(A[kk + kk * lda], A[kp + kp * lda]);
        if (kstep == 2) std::swap(A[k + 1 + k * lda], A[kp + k * lda]);
      }
      if (kstep == 1) {
        if (k < n - 1) {
          double r1 = 1.0 / A[k + k * lda];
          for (int j = k + 1; j < n; j++) {
            double t = -r1 * A[j + k * lda];
            for (int i = j; i < n; i++) A[i + j * lda] += A[i + k * lda] * t;
          }
          for (int i = k + 1; i < n; i++) A[i + k * lda] *= r1;
        }
      } else if (k < n - 2) {
        double d21 = A[k + 1 + k * lda];
        double d11 = A[k + 1 + (k + 1) * lda] / d21;
        double d22 = A[k + k * lda] / d21;
        double t = 1.0 / (d11 * d22 - 1.0);
        d21 = t / d21;
        for (int j = k + 2; j < n; j++) {
          double wk = d21 * (d11 * A[j + k * lda] - A[j + (k + 1) * lda]);
          double wkp1 = d21 * (d22 * A[j + (k + 1) * lda] - A[j + k * lda]);
          for (int i = j; i < n; i++)
            A[i + j * lda] = A[i + j * lda] - A[i + k * lda] * wk -
                             A[i + (k + 1) * lda] * wkp1;
          A[j + k * lda] = wk;
          A[j + (k + 1) * lda] = wkp1;
        }
      }
    }
    if (kstep == 1) {
      ipiv[k] = kp + 1;
    } else {
      ipiv[k] = -(kp + 1);
      ipiv[k + 1] = -(kp + 1);
    }
    k += kstep;
  }
  return info;
}

// LAPACK dsytrs 'L' restated.
static void bk_solve(int n, int nrhs, const double* A, int lda, const int* ipiv,
                     double* B, int ldb) {
  auto swaprow = [&](int a, int b) {
    for (int j = 0; j < nrhs; j++) std::swap(B[a + j * ldb], B[b + j * ldb]);
  };
  int k = 0;
  while (k < n) {
    if (ipiv[k] > 0) {
      int kp = ipiv[k] - 1;
      if (kp != k) swaprow(k, kp);
      for (int j = 0; j < nrhs; j++) {
        double bk = B[k + j * ldb];
        for (int i = k + 1; i < n; i++) B[i + j * ldb] -= A[i + k * lda] * bk;
        B[k + j * ldb] = bk * (1.0 / A[k + k * lda]);
      }
      k += 1;
    } else {
      int kp = -ipiv[k] - 1;
      if (kp != k + 1) swaprow(k + 1, kp);
      for (int j = 0; j < nrhs; j++) {
        double b0 = B[k + j * ldb], b1 = B[k + 1 + j * ldb];
        for (int i = k + 2; i < n; i++)
          B[i + j * ldb] -= A[i + k * lda] * b0 + A[i + (k + 1) * lda] * b1;
      }
      double akm1k = A[k + 1 + k * lda];
      double akm1 = A[k + k * lda] / akm1k;
      double ak = A[k + 1 + (k + 1) * lda] / akm1k;
      double denom = akm1 * ak - 1.0;
      for (int j = 0; j < nrhs; j++) {
        double bkm1 = B[k + j * ldb] / akm1k;
        double bk = B[k + 1 + j * ldb] / akm1k;
        B[k + j * ldb] = (ak * bkm1 - bk) / denom;
        B[k + 1 + j * ldb] = (akm1 * bk - bkm1) / denom;
      }
      k += 2;
    }
  }
  k = n - 1;
  while (k >= 0) {
    if (ipiv[k] > 0) {
      for (int j = 0; j < nrhs; j++) {
        double s = 0.0;
        for (int i = k + 1; i < n; i++) s += A[i + k * lda] * B[i + j * ldb];
        B[k + j * ldb] -= s;
      }
      int kp = ipiv[k] - 1;
      if (kp != k) swaprow(k, kp);
      k -= 1;
    } else {
      for (int j = 0; j < nrhs; j++) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = k + 1; i < n; i++) {
          s0 += A[i + k * lda] * B[i + j * ldb];
          s1 += A[i + (k - 1) * lda] * B[i + j * ldb];
        }
        B[k + j * ldb] -= s0;
        B[k - 1 + j * ldb] -= s1;
      }
      int kp = -ipiv[k] - 1;
      if (kp != k) swaprow(k, kp);
      k -= 2;
    }
  }
}

static void factor_root(Solver& s, int ks, int m, const int* vars, double* F) {
  NodeFactor& nf = s.fac[ks];
  nf.root = true;
  nf.npiv = m;
  nf.ndel = 0;
  nf.csize = 0;
  nf.pivots.assign(vars, vars + m);
  nf.delayed.clear();
  nf.ipiv.assign(m, 0);
  int info = bk_factor(m, F, m, nf.ipiv.data());
  if (info && !s.info) s.info = info;
  nf.L.assign(F, F + (size_t)m * m);
}

static void factor_cholesky(Solver& s, int ks, int fs, int m, const int* vars,
                            double* F) {
  const int ld = m;
  for (int j = 0; j < fs; j++) {  // potrf 'L' on F11
    double d = F[j + j * ld];
    for (int c = 0; c < j; c++) d -= F[j + c * ld] * F[j + c * ld];
    if (d <= 0.0 && !s.info) s.info = j + 1;
    d = std::sqrt(d);
    F[j + j * ld] = d;
    for (int i = j + 1; i < fs; i++) {
      double v = F[i + j * ld];
      for (int c = 0; c < j; c++) v -= F[i + c * ld] * F[j + c * ld];
      F[i + j * ld] = v / d;
    }
  }
  for (int i = fs; i < m; i++)  // trsm: L21 = F21 L11^{-T}
    for (int j = 0; j < fs; j++) {
      double v = F[i + j * ld];
      for (int c = 0; c < j; c++) v -= F[i + c * ld] * F[j + c * ld];
      F[i + j * ld] = v / F[j + j * ld];
    }
  for (int j = fs; j < m; j++)  // syrk: F22 -= L21 L21^T
    for (int i = j; i < m; i++) {
      double v = 0.0;
      for (int c = 0; c < fs; c++) v += F[i + c * ld] * F[j + c * ld];
      F[i + j * ld] -= v;
    }
  Contribution& C = s.con[ks];
  C.ndel = 0;
  C.size = m - fs;
  C.vars.assign(vars + fs, vars + m);
  C.block.resize((size_t)C.size * (C.size + 1) / 2);
  size_t p = 0;
  for (int j = fs; j < m; j++)
    for (int i = j; i < m; i++) C.block[p++] = F[i + (size_t)j * m];
  NodeFactor& nf = s.fac[ks];
  nf.npiv = fs;
  nf.ndel = 0;
  nf.csize = m - fs;
  nf.pivots.assign(vars, vars + fs);
  nf.delayed.clear();
  nf.L.assign(F, F + (size_t)m * fs);
}

static void factor_node(Solver& s, int ks, const int* colp, const int* rows,
                        const double* data, const double* diag) {
  Symbolic& S = s.S;
  int c0 = S.children_ptr[ks], c1 = S.children_ptr[ks + 1];
  for (int c = c0; c < c1; c++) factor_node(s, S.children[c], colp, rows, data, diag);

  // Frontal variables: own columns, delayed pivots of the children (in child
  // order), then the nominal row pattern (sparse_ldl.h:970-1000, 619-636).
  int k0 = S.node_ptr[ks], ns = S.node_ptr[ks + 1] - k0;
  vector<int> vars;
  vars.reserve(ns + 16);
  for (int j = 0; j < ns; j++) {
    int v = S.node_to_var[k0 + j];
    s.findex[v] = j;
    vars.push_back(v);
  }
  for (int c = c0; c < c1; c++) {
    Contribution& C = s.con[S.children[c]];
    for (int j = 0; j < C.ndel; j++) {
      int v = C.vars[j];
      if (s.findex[v] == -1) {
        s.findex[v] = (int)vars.size();
        vars.push_back(v);
      }
    }
  }
  int fs = (int)vars.size();
  for (int p = S.contrib_ptr[ks]; p < S.contrib_ptr[ks + 1]; p++) {
    s.findex[S.contrib_rows[p]] = (int)vars.size();
    vars.push_back(S.contrib_rows[p]);
  }
  int m = (int)vars.size();

  // Assembly (sparse_ldl.h:1018-1174).
  vector<double> Fv((size_t)m * m, 0.0);
  double* F = Fv.data();
  for (int j = 0; j < ns; j++) {
    int v = S.node_to_var[k0 + j];
    double* Fj = &F[(size_t)m * j];
    if (!S.permuted) {
      for (int p = colp[v]; p < colp[v + 1]; p++)
        if (rows[p] >= v) Fj[s.findex[rows[p]]] += data[p];
    } else {
      int pj = S.iperm[v];
      for (int p = colp[v]; p < colp[v + 1]; p++) {
        int i = rows[p];
        int fi = s.findex[i];
        if (fi >= 0 && S.iperm[i] >= pj) Fj[fi] += data[p];
      }
    }
  }
  if (diag)
    for (int j = 0; j < ns; j++) F[(size_t)j * (m + 1)] += diag[S.node_to_var[k0 + j]];
  for (int c = c0; c < c1; c++) {
    Contribution& C = s.con[S.children[c]];
    int cs = C.size;
    vector<int> idx(cs);
    for (int i = 0; i < cs; i++) idx[i] = s.findex[C.vars[i]];
    for (int j = 0; j < cs; j++) {
      const double* Cj = &C.block[(size_t)j * cs - (size_t)j * (j + 1) / 2];
      int jf = idx[j];
      for (int i = j; i < cs; i++) {
        int fi = idx[i];
        if (fi >= jf) F[fi + (size_t)m * jf] += Cj[i];
        else F[jf + (size_t)m * fi] += Cj[i];
      }
    }
    C = Contribution();
  }

  if (s.cholesky) {
    factor_cholesky(s, ks, fs, m, vars.data(), F);
  } else if (fs < m) {
    factor_front(s, ks, fs, m, vars.data(), F);
  } else {
    factor_root(s, ks, m, vars.data(), F);
  }
  for (int v : vars) s.findex[v < 0 ? -v - 1 : v] = -1;
}

static void factor(Solver& s, const int* colp, const int* rows, const double* data,
                   const double* diag) {
  Symbolic& S = s.S;
  s.fac.assign(S.num_nodes, NodeFactor());
  s.con.assign(S.num_nodes, Contribution());
  s.findex.assign(S.n, -1);
  s.info = 0;
  for (int p = S.children_ptr[S.num_nodes]; p < S.children_ptr[S.num_nodes + 1]; p++)
    factor_node(s, S.children[p], colp, rows, data, diag);
}

static inline int unpiv(int v) { return v >= 0 ? v : -v - 1; }

// Supernodal solve (sparse_ldl.h:1825-1926, 2053-2133).
static void solve(const Solver& s, int nrhs, double* x, int ldx) {
  const Symbolic& S = s.S;
  int mt = 0;
  for (const NodeFactor& f : s.fac) mt = std::max(mt, f.npiv + f.ndel + f.csize);
  vector<double> t((size_t)mt * nrhs);
  for (int ks = 0; ks < S.num_nodes; ks++) {
    const NodeFactor& f = s.fac[ks];
    const int* cr = &S.contrib_rows[S.contrib_ptr[ks]];
    int np = f.npiv, nd = f.ndel, nc = f.csize, ld = np + nd + nc;
    for (int r = 0; r < nrhs; r++)
      for (int j = 0; j < np; j++) t[j + r * ld] = x[unpiv(f.pivots[j]) + (size_t)r * ldx];
    if (f.root) {
      bk_solve(np, nrhs, f.L.data(), ld, f.ipiv.data(), t.data(), ld);
    } else if (np > 0) {
      const double* L = f.L.data();
      for (int r = 0; r < nrhs; r++) {
        double* tr = &t[r * ld];
        if (s.cholesky) {
          for (int j = 0; j < np; j++) {
            tr[j] /= L[j + j * ld];
            for (int i = j + 1; i < np; i++) tr[i] -= L[i + j * ld] * tr[j];
          }
        } else {
          for (int j = 0; j < np;) {
            if (f.pivots[j] >= 0) {
              for (int i = j + 1; i < np; i++) tr[i] -= L[i + j * ld] * tr[j];
              j += 1;
            } else {
              for (int i = j + 2; i < np; i++)
                tr[i] -= L[i + j * ld] * tr[j] + L[i + (j + 1) * ld] * tr[j + 1];
              j += 2;
            }
          }
        }
        for (int i = np; i < ld; i++) {
          double v = 0.0;
          for (int j = 0; j < np; j++) v += L[i + j * ld] * tr[j];
          tr[i] = v;
        }
        for (int i = 0; i < nd; i++) x[f.delayed[i] + (size_t)r * ldx] -= tr[np + i];
        for (int i = 0; i < nc; i++) x[cr[i] + (size_t)r * ldx] -= tr[np + nd + i];
        if (!s.cholesky) {
          for (int j = 0; j < np;) {
            if (f.pivots[j] >= 0) {
              tr[j] /= L[j * (ld + 1)];
              j++;
            } else {
              double d11 = L[j * (ld + 1)], d21 = L[j * (ld + 1) + 1],
                     d22 = L[(j + 1) * (ld + 1)];
              double inv = 1.0 / (d11 * d22 - d21 * d21);
              double a = tr[j], b = tr[j + 1];
              tr[j] = inv * (d22 * a - d21 * b);
              tr[j + 1] = inv * (d11 * b - d21 * a);
              j += 2;
            }
          }
        }
      }
    }
    for (int r = 0; r < nrhs; r++)
      for (int j = 0; j < np; j++) x[unpiv(f.pivots[j]) + (size_t)r * ldx] = t[j + r * ld];
  }
  for (int ks = S.num_nodes - 1; ks >= 0; ks--) {
    const NodeFactor& f = s.fac[ks];
    if (f.root || f.npiv == 0) continue;
    const int* cr = &S.contrib_rows[S.contrib_ptr[ks]];
    int np = f.npiv, nd = f.ndel, nc = f.csize, ld = np + nd + nc;
    const double* L = f.L.data();
    for (int r = 0; r < nrhs; r++) {
      double* tr = &t[r * ld];
      for (int j = 0; j < np; j++) tr[j] = x[unpiv(f.pivots[j]) + (size_t)r * ldx];
      for (int i = 0; i < nd; i++) tr[np + i] = x[f.delayed[i] + (size_t)r * ldx];
      for (int i = 0; i < nc; i++) tr[np + nd + i] = x[cr[i] + (size_t)r * ldx];
      for (int j = 0; j < np; j++) {
        double v = 0.0;
        for (int i = np; i < ld; i++) v += L[i + j * ld] * tr[i];
        tr[j] -= v;
      }
      if (s.cholesky) {
        for (int j = np - 1; j >= 0; j--) {
          double v = tr[j];
          for (int i = j + 1; i < np; i++) v -= L[i + j * ld] * tr[i];
          tr[j] = v / L[j + j * ld];
        }
      } else {
        for (int i = np - 1; i >= 0;) {
          if (f.pivots[i] >= 0) {
            for (int j = i + 1; j < np; j++) tr[i] -= L[j + i * ld] * tr[j];
            i -= 1;
          } else {
            for (int j = i + 1; j < np; j++) {
              tr[i - 1] -= L[j + (i - 1) * ld] * tr[j];
              tr[i] -= L[j + i * ld] * tr[j];
            }
            i -= 2;
          }
        }
      }
      for (int j = 0; j < np; j++) x[unpiv(f.pivots[j]) + (size_t)r * ldx] = tr[j];
    }
  }
}

static void inertia(const Solver& s, int* npos, int* nneg) {
  *npos = *nneg = 0;
  if (s.cholesky) return;
  for (const NodeFactor& f : s.fac) {
    int ld = f.npiv + f.ndel + f.csize;
    const int* piv = f.root ? f.ipiv.data() : f.pivots.data();
    for (int i = 0; i < f.npiv;) {
      const double* L = f.L.data();
      if (piv[i] >= 0) {
        if (L[(size_t)i * (ld + 1)] >= 0.0) (*npos)++;
        else (*nneg)++;
        i++;
      } else {
        double d11 = L[(size_t)i * (ld + 1)], d21 = L[(size_t)i * (ld + 1) + 1],
               d22 = L[(size_t)(i + 1) * (ld + 1)];
        double det = d11 * d22 - d21 * d21;
        if (det >= 0.0) {
          if (d11 >= 0.0) *npos += 2;
          else *nneg += 2;
        } else {
          (*npos)++;
          (*nneg)++;
        }
        i += 2;
      }
    }
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI used by the tests (ctypes).
// ---------------------------------------------------------------------------
extern "C" {

struct OrcHandle {
  orc::Solver s;
  int n;
  std::vector<int> rowp, cols;
};

// order: 0 natural, 1 AMD (reference BlockAMD), 2 caller-given permutation
// (perm[new] = old).  solver_type: 0 LDL, 1 Cholesky.
void* orc_analyze(int n, const int* rowp, const int* cols, int order,
                  const int* perm_in, double ustab, double pivot_tol,
                  int solver_type) {
  OrcHandle* h = new OrcHandle();
  h->n = n;
  h->rowp.assign(rowp, rowp + n + 1);
  h->cols.assign(cols, cols + rowp[n]);
  h->s.ustab = ustab;
  h->s.pivot_tol = pivot_tol;
  h->s.cholesky = solver_type == 1;
  std::vector<int> perm;
  const int* p = nullptr;
  if (order == 1) {
    std::vector<int> gp, gc;
    orc::amd_graph(n, rowp, cols, gp, gc);
    perm.assign(n, 0);
    orc::amd_order(n, gp, gc, perm.data());
    p = perm.data();
  } else if (order == 2) {
    p = perm_in;
  }
  orc::analyze(h->s.S, n, rowp, cols, p);
  return h;
}

void orc_destroy(void* h) { delete (OrcHandle*)h; }

int orc_num_nodes(void* h) { return ((OrcHandle*)h)->s.S.num_nodes; }
int orc_num_snodes(void* h) { return ((OrcHandle*)h)->s.S.num_snodes; }
int orc_contrib_total(void* h) {
  auto& S = ((OrcHandle*)h)->s.S;
  return S.contrib_ptr[S.num_nodes];
}

void orc_get_perm(void* hh, int* perm, int* iperm) {
  OrcHandle* h = (OrcHandle*)hh;
  for (int i = 0; i < h->n; i++) {
    perm[i] = h->s.S.permuted ? h->s.S.perm[i] : i;
    iperm[i] = h->s.S.permuted ? h->s.S.iperm[i] : i;
  }
}

void orc_get_nodes(void* hh, int* node_ptr, int* node_to_var, int* children_ptr,
                   int* children, int* contrib_ptr, int* contrib_rows) {
  auto& S = ((OrcHandle*)hh)->s.S;
  std::copy(S.node_ptr.begin(), S.node_ptr.end(), node_ptr);
  std::copy(S.node_to_var.begin(), S.node_to_var.end(), node_to_var);
  std::copy(S.children_ptr.begin(), S.children_ptr.end(), children_ptr);
  std::copy(S.children.begin(), S.children.end(), children);
  std::copy(S.contrib_ptr.begin(), S.contrib_ptr.end(), contrib_ptr);
  std::copy(S.contrib_rows.begin(), S.contrib_rows.end(), contrib_rows);
}

// Elimination tree parent, post-order and fundamental column counts (incl.
// diagonal), all in the permuted index space.
void orc_get_etree(void* hh, int* parent, int* post, int* colcount) {
  auto& S = ((OrcHandle*)hh)->s.S;
  std::copy(S.parent.begin(), S.parent.end(), parent);
  std::copy(S.post.begin(), S.post.end(), post);
  std::copy(S.colcount.begin(), S.colcount.end(), colcount);
}

void orc_pivot_kinds(long long* out) {
  for (int i = 0; i < 5; i++) out[i] = orc::g_pivot_kinds[i];
}

int orc_factor(void* hh, const double* vals, const double* diag) {
  for (int i = 0; i < 5; i++) orc::g_pivot_kinds[i] = 0;
  OrcHandle* h = (OrcHandle*)hh;
  orc::factor(h->s, h->rowp.data(), h->cols.data(), vals, diag);
  return h->s.info;
}

void orc_solve(void* hh, int nrhs, double* x, int ldx) {
  orc::solve(((OrcHandle*)hh)->s, nrhs, x, ldx);
}

void orc_inertia(void* hh, int* npos, int* nneg) {
  orc::inertia(((OrcHandle*)hh)->s, npos, nneg);
}

// Per-node pivot counts after a factorization: npiv, ndel (delayed out of
// the node), n2x2 (number of 2x2 blocks), root flag.
void orc_node_info(void* hh, int* npiv, int* ndel, int* n2x2, int* root) {
  auto& F = ((OrcHandle*)hh)->s.fac;
  for (size_t k = 0; k < F.size(); k++) {
    npiv[k] = F[k].npiv;
    ndel[k] = F[k].ndel;
    root[k] = F[k].root;
    int c = 0;
    const int* piv = F[k].root ? F[k].ipiv.data() : F[k].pivots.data();
    for (int i = 0; i < F[k].npiv;) {
      if (piv[i] < 0) { c++; i += 2; } else { i++; }
    }
    n2x2[k] = c;
  }
}

// Stored factor entries (sum over nodes of rows x pivots).
long long orc_factor_nnz(void* hh) {
  long long t = 0;
  for (auto& f : ((OrcHandle*)hh)->s.fac)
    t += (long long)(f.npiv + f.ndel + f.csize) * f.npiv;
  return t;
}

}  // extern "C"
