// Nested dissection ordering for the KKT LDL^T engine.  The reference orders
// with METIS_NodeND (include/ordering_utils.h:177-197), which is not
// available in this build; this is an ordering of its own with the same
// goal: small vertex separators, numbered last, found recursively.
//
//  * Indistinguishable vertices (identical closed neighbourhoods, e.g. the
//    three displacement components of a mesh node) are merged first into
//    weighted vertices.
//  * Each dissection step computes two candidate vertex separators and keeps
//    the smaller one (weighted by the imbalance of the two sides):
//      - multilevel: heavy-edge-matching coarsening down to ~100 vertices,
//        separators grown from several seeds on the coarsest graph, then
//        projected back level by level, each level refined by a two-sided
//        node Fiduccia-Mattheyses pass (move a separator vertex to one side
//        and pull its neighbours on the other side into the separator;
//        hill-climbing with roll-back to the best state);
//      - a breadth-first level set from a pseudo-peripheral vertex, refined
//        the same way.
//  * Disconnected parts are dissected separately; parts of at most kLeaf
//    original vertices are ordered by the engine's AMD.
//
// Randomness (matching order, seeds) comes from a fixed-seed generator, so
// the ordering is a deterministic function of the pattern.
#include <algorithm>
#include <future>
#include <limits>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <queue>
#include <random>
#include <vector>

#include "symbolic.hpp"

namespace kktldl {

using std::vector;

namespace {

constexpr int kLeaf = 200;        // parts up to this many original vertices go to AMD
constexpr double kUbFactor = 1.2; // side weight bound: kUbFactor * total / 2
constexpr int kBand = 2;              // band half-width of the flow refinement
constexpr long long kFlowBudget = 3000000;  // adjacency entries of a band worth a max flow

struct Graph {
  int n = 0;
  vector<int> xadj, adj, ew;  // CSR adjacency without self loops, edge weights
  vector<int> vw;             // vertex weights (original vertices represented)
  long long total = 0;
  void finish() {
    total = 0;
    for (int v = 0; v < n; v++) total += vw[v];
  }
};

// --- coarsening -------------------------------------------------------------

// Heavy-edge matching in random vertex order; coarse vertices are capped in
// weight so that no coarse vertex dominates the bisection.
static Graph coarsen(const Graph& G, vector<int>& cmap, std::mt19937& rng) {
  const int n = G.n;
  vector<int> order(n), match(n, -1);
  std::iota(order.begin(), order.end(), 0);
  std::shuffle(order.begin(), order.end(), rng);
  const long long cap = std::max<long long>(1, (long long)(1.5 * G.total / 100));
  for (int v : order) {
    if (match[v] >= 0) continue;
    int best = -1, bw = -1;
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) {
      const int u = G.adj[p];
      if (match[u] >= 0 || (long long)G.vw[u] + G.vw[v] > cap) continue;
      if (G.ew[p] > bw) {
        bw = G.ew[p];
        best = u;
      }
    }
    if (best >= 0) {
      match[v] = best;
      match[best] = v;
    } else {
      match[v] = v;
    }
  }
  cmap.assign(n, -1);
  int nc = 0;
  for (int v : order)
    if (cmap[v] < 0) cmap[v] = cmap[match[v]] = nc++;
  vector<int> first(nc, -1), second(nc, -1);
  for (int v = 0; v < n; v++) {
    const int c = cmap[v];
    if (first[c] < 0) first[c] = v;
    else second[c] = v;
  }
  Graph C;
  C.n = nc;
  C.xadj.assign(nc + 1, 0);
  C.vw.assign(nc, 0);
  C.adj.reserve(G.adj.size());
  C.ew.reserve(G.adj.size());
  vector<int> pos(nc, -1);
  for (int c = 0; c < nc; c++) {
    const int start = (int)C.adj.size();
    for (int v : {first[c], second[c]}) {
      if (v < 0) continue;
      C.vw[c] += G.vw[v];
      for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) {
        const int cu = cmap[G.adj[p]];
        if (cu == c) continue;
        if (pos[cu] < 0) {
          pos[cu] = (int)C.adj.size();
          C.adj.push_back(cu);
          C.ew.push_back(G.ew[p]);
        } else {
          C.ew[pos[cu]] += G.ew[p];
        }
      }
    }
    for (int p = start; p < (int)C.adj.size(); p++) pos[C.adj[p]] = -1;
    C.xadj[c + 1] = (int)C.adj.size();
  }
  C.finish();
  return C;
}

// --- node separator refinement ---------------------------------------------

// where[v]: 0 = side A, 1 = side B, 2 = separator.
struct Sides {
  long long w[3] = {0, 0, 0};
};

static Sides side_weights(const Graph& G, const vector<int>& where) {
  Sides s;
  for (int v = 0; v < G.n; v++) s.w[where[v]] += G.vw[v];
  return s;
}

// Both sides non-empty and neither above kUbFactor * total / 2.
static bool balanced(const Graph& G, const Sides& s) {
  const double maxside = kUbFactor * (double)G.total / 2.0;
  return s.w[0] > 0 && s.w[1] > 0 && (double)s.w[0] <= maxside && (double)s.w[1] <= maxside;
}

// Score of a separator: its weight, inflated by the imbalance of the sides.
static double sep_score(const Sides& s) {
  const double a = (double)s.w[0], b = (double)s.w[1];
  const double imb = (a + b) > 0 ? std::fabs(a - b) / (a + b) : 1.0;
  return (double)s.w[2] * (1.0 + imb);
}

struct HeapItem {
  long long gain;
  int v;
  bool operator<(const HeapItem& o) const { return gain < o.gain || (gain == o.gain && v > o.v); }
};

// Two-sided node FM: moving separator vertex v to side t puts v in t and
// pulls its neighbours on side 1-t into the separator; gain = vw(v) -
// (weight pulled).  A pass moves every vertex at most once, accepts
// non-improving moves up to a limit and rolls back to the best state.
static void node_fm(const Graph& G, vector<int>& where, int npass) {
  const int n = G.n;
  if (n == 0) return;
  const long long maxside = (long long)std::ceil(kUbFactor * (double)G.total / 2.0);
  vector<long long> gain[2] = {vector<long long>(n, 0), vector<long long>(n, 0)};
  vector<int> locked(n, 0);
  int stamp = 0;
  auto compute = [&](int v, int t) {
    long long g = G.vw[v];
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++)
      if (where[G.adj[p]] == 1 - t) g -= G.vw[G.adj[p]];
    return g;
  };
  struct Move {
    int v, t, pulled_begin, pulled_end;
  };
  vector<Move> log;
  vector<int> pulled;
  const int limit = std::max(32, std::min(400, n / 50));
  for (int pass = 0; pass < npass; pass++) {
    Sides s = side_weights(G, where);
    stamp++;
    std::priority_queue<HeapItem> heap[2];
    for (int v = 0; v < n; v++)
      if (where[v] == 2)
        for (int t = 0; t < 2; t++) {
          gain[t][v] = compute(v, t);
          heap[t].push({gain[t][v], v});
        }
    log.clear();
    pulled.clear();
    const long long start_sep = s.w[2];
    long long best_sep = s.w[2];
    long long best_bal = std::llabs(s.w[0] - s.w[1]);
    size_t best_len = 0;
    int nbad = 0;
    vector<int> touched;
    while (nbad < limit) {
      int cand[2] = {-1, -1};
      for (int t = 0; t < 2; t++) {
        while (!heap[t].empty()) {
          const HeapItem h = heap[t].top();
          const int v = h.v;
          if (where[v] != 2 || locked[v] == stamp || h.gain != gain[t][v]) {
            heap[t].pop();
            continue;
          }
          if (s.w[t] + G.vw[v] > maxside) break;  // this side is full
          cand[t] = v;
          break;
        }
      }
      int t = -1;
      if (cand[0] >= 0 && cand[1] >= 0) {
        const long long g0 = gain[0][cand[0]], g1 = gain[1][cand[1]];
        t = g0 > g1 ? 0 : (g1 > g0 ? 1 : (s.w[0] <= s.w[1] ? 0 : 1));
      } else if (cand[0] >= 0) {
        t = 0;
      } else if (cand[1] >= 0) {
        t = 1;
      }
      if (t < 0) break;
      const int v = cand[t];
      heap[t].pop();
      // move v to side t
      const int o = 1 - t;
      where[v] = t;
      locked[v] = stamp;
      s.w[2] -= G.vw[v];
      s.w[t] += G.vw[v];
      const int pb = (int)pulled.size();
      for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) {
        const int u = G.adj[p];
        if (where[u] == o) {
          where[u] = 2;
          s.w[o] -= G.vw[u];
          s.w[2] += G.vw[u];
          pulled.push_back(u);
        }
      }
      log.push_back({v, t, pb, (int)pulled.size()});
      // gains change for separator vertices next to v or to a pulled vertex
      touched.clear();
      for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) touched.push_back(G.adj[p]);
      for (int q = pb; q < (int)pulled.size(); q++) {
        const int u = pulled[q];
        touched.push_back(u);
        for (int p = G.xadj[u]; p < G.xadj[u + 1]; p++) touched.push_back(G.adj[p]);
      }
      for (int w : touched) {
        if (where[w] != 2 || locked[w] == stamp) continue;
        for (int tt = 0; tt < 2; tt++) {
          gain[tt][w] = compute(w, tt);
          heap[tt].push({gain[tt][w], w});
        }
      }
      const long long bal = std::llabs(s.w[0] - s.w[1]);
      if (s.w[2] < best_sep || (s.w[2] == best_sep && bal < best_bal)) {
        best_sep = s.w[2];
        best_bal = bal;
        best_len = log.size();
        nbad = 0;
      } else {
        nbad++;
      }
    }
    // roll back the moves after the best state
    for (size_t i = log.size(); i > best_len; i--) {
      const Move& m = log[i - 1];
      for (int q = m.pulled_begin; q < m.pulled_end; q++) where[pulled[q]] = 1 - m.t;
      where[m.v] = 2;
    }
    if (best_sep >= start_sep) break;
  }
}


// --- minimum vertex cut in a band around the separator ----------------------

// Max flow (Dinic) on a small directed graph with integer capacities; edges
// are added first (add), then frozen into CSR form (build).
struct FlowNet {
  int nn = 0;
  vector<int> from, to, head, order;  // edge e: from[e] -> to[e], twin e ^ 1
  vector<long long> cap;
  vector<int> level, it, stack;
  void reset(int nodes) {
    nn = nodes;
    from.clear();
    to.clear();
    cap.clear();
  }
  void add(int a, int b, long long c) {
    from.push_back(a);
    to.push_back(b);
    cap.push_back(c);
    from.push_back(b);
    to.push_back(a);
    cap.push_back(0);
  }
  void build() {
    head.assign(nn + 1, 0);
    for (int a : from) head[a + 1]++;
    for (int v = 0; v < nn; v++) head[v + 1] += head[v];
    order.resize(from.size());
    vector<int> fill(head.begin(), head.end() - 1);
    for (int e = 0; e < (int)from.size(); e++) order[fill[from[e]]++] = e;
  }
  bool bfs_levels(int s, int t) {
    level.assign(nn, -1);
    stack.assign(1, s);
    level[s] = 0;
    for (size_t h = 0; h < stack.size(); h++) {
      const int v = stack[h];
      for (int q = head[v]; q < head[v + 1]; q++) {
        const int e = order[q];
        if (cap[e] > 0 && level[to[e]] < 0) {
          level[to[e]] = level[v] + 1;
          stack.push_back(to[e]);
        }
      }
    }
    return level[t] >= 0;
  }
  long long push(int v, int t, long long f) {
    if (v == t) return f;
    for (int& q = it[v]; q < head[v + 1]; q++) {
      const int e = order[q];
      const int u = to[e];
      if (cap[e] <= 0 || level[u] != level[v] + 1) continue;
      const long long d = push(u, t, std::min(f, cap[e]));
      if (d > 0) {
        cap[e] -= d;
        cap[e ^ 1] += d;
        return d;
      }
    }
    return 0;
  }
  long long maxflow(int s, int t, long long inf) {
    long long f = 0;
    while (bfs_levels(s, t)) {
      it.assign(head.begin(), head.end() - 1);
      while (long long d = push(s, t, inf)) f += d;
    }
    return f;
  }
  // nodes reachable from s (fwd) or reaching t (!fwd) in the residual graph
  void residual_reach(int s, bool fwd, vector<char>& mark) {
    mark.assign(nn, 0);
    stack.assign(1, s);
    mark[s] = 1;
    for (size_t h = 0; h < stack.size(); h++) {
      const int v = stack[h];
      for (int q = head[v]; q < head[v + 1]; q++) {
        const int e = order[q];
        const int u = to[e];
        // forward: edge v -> u usable; backward: u -> v (the twin) usable
        const long long c = fwd ? cap[e] : cap[e ^ 1];
        if (c > 0 && !mark[u]) {
          mark[u] = 1;
          stack.push_back(u);
        }
      }
    }
  }
};

// Replaces the separator by a minimum-weight vertex cut between the two
// sides, chosen among the vertices within distance d of the separator (the
// band).  Vertices of a side outside the band stay where they are; of the
// two extreme minimum cuts (nearest side A, nearest side B) the better
// balanced one is kept.  The result is a valid separator of weight at most
// the old one.
static void flow_refine(const Graph& G, vector<int>& where, int d) {
  const int n = G.n;
  vector<int> dist(n, -1), band;
  for (int v = 0; v < n; v++)
    if (where[v] == 2) {
      dist[v] = 0;
      band.push_back(v);
    }
  if (band.empty()) return;
  // the band takes at most 40% of each side, so that both terminals keep
  // vertices outside it
  const Sides before = side_weights(G, where);
  long long taken[2] = {0, 0};
  for (size_t h = 0; h < band.size(); h++) {
    const int v = band[h];
    if (dist[v] >= d) continue;
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) {
      const int u = G.adj[p];
      if (dist[u] >= 0 || where[u] == 2) continue;
      if (5 * (taken[where[u]] + G.vw[u]) > 2 * before.w[where[u]]) continue;
      taken[where[u]] += G.vw[u];
      dist[u] = dist[v] + 1;
      band.push_back(u);
    }
  }
  // the flow network grows with the band's adjacency; large bands (3D
  // separators) are cut back to the separator's first layer, and past a
  // budget the refinement is left to node FM
  {
    long long deg = 0;
    for (int v : band) deg += G.xadj[v + 1] - G.xadj[v];
    if (deg > kFlowBudget && d > 1) {
      band.resize(std::stable_partition(band.begin(), band.end(),
                                        [&](int v) { return dist[v] <= 1; }) - band.begin());
      for (int v = 0; v < n; v++)
        if (dist[v] > 1) dist[v] = -1;
      deg = 0;
      for (int v : band) deg += G.xadj[v + 1] - G.xadj[v];
    }
    if (deg > kFlowBudget) return;
  }
  const int nb = (int)band.size();
  vector<int> bid(n, -1);
  for (int i = 0; i < nb; i++) bid[band[i]] = i;
  const int S = 2 * nb, T = 2 * nb + 1;
  static thread_local FlowNet F;
  F.reset(2 * nb + 2);
  const long long inf = G.total + 1;
  for (int i = 0; i < nb; i++) {
    const int v = band[i];
    F.add(2 * i, 2 * i + 1, G.vw[v]);
    bool src = false, snk = false;
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) {
      const int u = G.adj[p];
      const int j = bid[u];
      if (j >= 0) {
        F.add(2 * i + 1, 2 * j, inf);
      } else if (where[u] == 0) {
        src = true;
      } else if (where[u] == 1) {
        snk = true;
      }
    }
    if (src) F.add(S, 2 * i, inf);
    if (snk) F.add(2 * i + 1, T, inf);
  }
  F.build();
  long long old_sep = 0;
  for (int v : band)
    if (where[v] == 2) old_sep += G.vw[v];
  const long long f = F.maxflow(S, T, inf);
  if (f >= old_sep) return;
  vector<char> fromS, toT;
  F.residual_reach(S, true, fromS);
  F.residual_reach(T, false, toT);
  vector<int> wa = where, wb = where;
  for (int i = 0; i < nb; i++) {
    const int v = band[i];
    // cut nearest A: in reachable from S, out not
    wa[v] = (fromS[2 * i] && !fromS[2 * i + 1]) ? 2 : (fromS[2 * i] ? 0 : 1);
    // cut nearest B: out reaches T, in does not
    wb[v] = (toT[2 * i + 1] && !toT[2 * i]) ? 2 : (toT[2 * i] ? 1 : 0);
  }
  const Sides xa = side_weights(G, wa), xb = side_weights(G, wb);
  const double inf_score = 1e300;
  const double sa = balanced(G, xa) ? sep_score(xa) : inf_score;
  const double sb = balanced(G, xb) ? sep_score(xb) : inf_score;
  const double sold = balanced(G, before) ? sep_score(before) : inf_score;
  if (std::min(sa, sb) >= sold && sold < inf_score) return;
  if (std::min(sa, sb) >= inf_score) return;
  where = sa <= sb ? wa : wb;
}

// --- initial separators and bisection ---------------------------------------

// Breadth-first search from src; dist[v] = level (-1 unreached).  Returns
// the number of levels and fills `order` with the visited vertices.
static int bfs(const Graph& G, int src, vector<int>& dist, vector<int>& order) {
  std::fill(dist.begin(), dist.end(), -1);
  order.clear();
  order.push_back(src);
  dist[src] = 0;
  int nl = 1;
  for (size_t h = 0; h < order.size(); h++) {
    const int v = order[h];
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++) {
      const int u = G.adj[p];
      if (dist[u] >= 0) continue;
      dist[u] = dist[v] + 1;
      nl = std::max(nl, dist[u] + 1);
      order.push_back(u);
    }
  }
  return nl;
}

// Pseudo-peripheral vertex: repeated BFS from the farthest vertex.
static int peripheral(const Graph& G, int s, vector<int>& dist, vector<int>& order) {
  int nl = bfs(G, s, dist, order), prev = -1;
  for (int it = 0; it < 4 && nl > prev; it++) {
    prev = nl;
    const int far = order.back();
    const int nl2 = bfs(G, far, dist, order);
    if (nl2 <= nl) {
      bfs(G, s, dist, order);
      break;
    }
    s = far;
    nl = nl2;
  }
  bfs(G, s, dist, order);
  return s;
}

// Side A grown breadth-first from seed until it holds half the weight; the
// separator is A's boundary.  The graph is connected.
static void grow_separator(const Graph& G, int seed, vector<int>& where, vector<int>& dist,
                           vector<int>& order) {
  bfs(G, seed, dist, order);
  where.assign(G.n, 1);
  long long wa = 0;
  for (int v : order) {
    if (2 * wa >= G.total) break;
    where[v] = 0;
    wa += G.vw[v];
  }
  for (int v = 0; v < G.n; v++) {
    if (where[v] != 0) continue;
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++)
      if (where[G.adj[p]] == 1) {
        where[v] = 2;
        break;
      }
  }
}

// Best of several grown and refined separators of a small graph.
static void initial_separator(const Graph& G, vector<int>& where, std::mt19937& rng) {
  vector<int> dist(G.n), order, w;
  const int s0 = peripheral(G, 0, dist, order);
  double best = 1e300;
  const int trials = G.total >= 20000 ? 8 : (G.total >= 4000 ? 4 : 2);
  for (int t = 0; t < trials; t++) {
    const int seed = t == 0 ? s0 : (int)(rng() % (unsigned)G.n);
    grow_separator(G, seed, w, dist, order);
    flow_refine(G, w, kBand);
    node_fm(G, w, 8);
    const Sides sw = side_weights(G, w);
    const double sc = sep_score(sw) * (balanced(G, sw) ? 1.0 : 1e6);
    if (sc < best) {
      best = sc;
      where = w;
    }
  }
}

// Multilevel node bisection of a connected graph.
static void ml_bisect(const Graph& G, vector<int>& where, std::mt19937& rng, int coarse_target) {
  vector<Graph> levels;
  vector<vector<int>> cmaps;
  const Graph* cur = &G;
  const int coarsest = std::max(100, std::min(coarse_target, G.n / 16));
  while (cur->n > coarsest) {
    vector<int> cmap;
    Graph C = coarsen(*cur, cmap, rng);
    if (C.n > 0.9 * cur->n) break;  // matching stalled
    cmaps.push_back(std::move(cmap));
    levels.push_back(std::move(C));
    cur = &levels.back();
  }
  vector<int> w;
  initial_separator(*cur, w, rng);
  for (int l = (int)levels.size() - 1; l >= 0; l--) {
    const Graph& fine = l > 0 ? levels[l - 1] : G;
    vector<int> wf(fine.n);
    for (int v = 0; v < fine.n; v++) wf[v] = w[cmaps[l][v]];
    flow_refine(fine, wf, kBand);
    node_fm(fine, wf, 6);
    w.swap(wf);
  }
  where.swap(w);
}

// Breadth-first level-set separator from a pseudo-peripheral vertex (the
// level that splits the weight most evenly among those of small size),
// refined by node FM.
static void bfs_bisect(const Graph& G, vector<int>& where) {
  vector<int> dist(G.n), order;
  peripheral(G, 0, dist, order);
  int nl = 0;
  for (int v : order) nl = std::max(nl, dist[v] + 1);
  vector<long long> cnt(nl, 0), cum(nl + 1, 0);
  for (int v : order) cnt[dist[v]] += G.vw[v];
  for (int l = 0; l < nl; l++) cum[l + 1] = cum[l] + cnt[l];
  int best = -1;
  double best_score = 1e300;
  for (int l = 1; l + 1 < nl; l++) {
    const long long below = cum[l], above = G.total - cum[l + 1];
    if (4 * std::min(below, above) < G.total - cnt[l]) continue;
    Sides s;
    s.w[0] = below;
    s.w[1] = above;
    s.w[2] = cnt[l];
    const double sc = sep_score(s);
    if (sc < best_score) {
      best_score = sc;
      best = l;
    }
  }
  if (best < 0) {
    best = std::max(1, std::min(nl - 2, nl / 2));
  }
  where.assign(G.n, 0);
  for (int v : order) where[v] = dist[v] < best ? 0 : (dist[v] > best ? 1 : 2);
  flow_refine(G, where, kBand);
  node_fm(G, where, 6);
}

// Subgraph induced by the vertices with where[v] == side (local ids in the
// order of increasing global id).
static Graph induced(const Graph& G, const vector<int>& where, int side, vector<int>& ids,
                     vector<int>& loc) {
  ids.clear();
  for (int v = 0; v < G.n; v++)
    if (where[v] == side) {
      loc[v] = (int)ids.size();
      ids.push_back(v);
    }
  Graph H;
  H.n = (int)ids.size();
  H.xadj.assign(H.n + 1, 0);
  H.vw.resize(H.n);
  for (int i = 0; i < H.n; i++) {
    const int v = ids[i];
    H.vw[i] = G.vw[v];
    for (int p = G.xadj[v]; p < G.xadj[v + 1]; p++)
      if (where[G.adj[p]] == side) {
        H.adj.push_back(loc[G.adj[p]]);
        H.ew.push_back(G.ew[p]);
      }
    H.xadj[i + 1] = (int)H.adj.size();
  }
  H.finish();
  return H;
}

// --- compression ------------------------------------------------------------

// Merges vertices with identical closed neighbourhoods.  Returns the
// compressed graph; members[group] lists the original vertices of a group
// in increasing order.
static Graph compress(int n, const int* rowp, const int* cols, vector<vector<int>>& members) {
  // closed neighbourhood of v: its row with the diagonal added, sorted
  vector<uint64_t> hash(n);
  vector<int> deg(n);
  auto row = [&](int v, vector<int>& out) {
    out.clear();
    bool diag = false;
    for (int p = rowp[v]; p < rowp[v + 1]; p++) {
      out.push_back(cols[p]);
      diag |= cols[p] == v;
    }
    if (!diag) {
      out.push_back(v);
      std::sort(out.begin(), out.end());
    }
  };
  vector<int> r, r2;
  for (int v = 0; v < n; v++) {
    row(v, r);
    uint64_t h = 1469598103934665603ull;
    for (int c : r) h = (h ^ (uint64_t)(uint32_t)c) * 1099511628211ull;
    hash[v] = h;
    deg[v] = (int)r.size();
  }
  vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) {
    return hash[a] != hash[b] ? hash[a] < hash[b] : (deg[a] != deg[b] ? deg[a] < deg[b] : a < b);
  });
  vector<int> group(n, -1);
  members.clear();
  for (size_t i = 0; i < idx.size();) {
    size_t j = i;
    while (j < idx.size() && hash[idx[j]] == hash[idx[i]] && deg[idx[j]] == deg[idx[i]]) j++;
    // exact comparison inside the hash bucket
    for (size_t a = i; a < j; a++) {
      const int v = idx[a];
      if (group[v] >= 0) continue;
      group[v] = (int)members.size();
      members.push_back({v});
      row(v, r);
      for (size_t b = a + 1; b < j; b++) {
        const int u = idx[b];
        if (group[u] >= 0) continue;
        row(u, r2);
        if (r2 == r) {
          group[u] = group[v];
          members.back().push_back(u);
        }
      }
    }
    i = j;
  }
  // groups numbered by their smallest member, members ascending
  for (auto& m : members) std::sort(m.begin(), m.end());
  vector<int> gorder(members.size());
  std::iota(gorder.begin(), gorder.end(), 0);
  std::sort(gorder.begin(), gorder.end(),
            [&](int a, int b) { return members[a][0] < members[b][0]; });
  vector<vector<int>> sorted(members.size());
  vector<int> renum(members.size());
  for (size_t g = 0; g < gorder.size(); g++) {
    renum[gorder[g]] = (int)g;
    sorted[g] = std::move(members[gorder[g]]);
  }
  members.swap(sorted);
  for (int v = 0; v < n; v++) group[v] = renum[group[v]];
  Graph G;
  G.n = (int)members.size();
  G.xadj.assign(G.n + 1, 0);
  G.vw.resize(G.n);
  vector<int> mark(G.n, -1);
  for (int g = 0; g < G.n; g++) {
    G.vw[g] = (int)members[g].size();
    const int v = members[g][0];
    for (int p = rowp[v]; p < rowp[v + 1]; p++) {
      const int h = group[cols[p]];
      if (h == g || mark[h] == g) continue;
      mark[h] = g;
      G.adj.push_back(h);
      G.ew.push_back(1);
    }
    G.xadj[g + 1] = (int)G.adj.size();
  }
  G.finish();
  return G;
}

}  // namespace

namespace {

// Orders a set of columns with identical structure below them so that
// columns that will probably need a 2x2 pivot together are adjacent: the
// large-front panels then meet them as a pivot and its next candidate.  The
// prediction uses the values of the pattern at analysis time: a column whose
// diagonal fails the threshold test against its largest entry inside the set
// (|a_vv| < ustab |a_vu|) is paired with that column u; columns are visited
// in increasing |a_vv| / max_u |a_vu| (zero diagonals first), each taking its
// largest not yet paired neighbour.  Units (pairs, or single columns) keep
// the order of their smallest original index; inside a pair the larger
// diagonal comes first.
static void pair_order(const int* rowp, const int* cols, const double* vals, double ustab,
                       vector<int>& seg, vector<int>& local) {
  const int ns = (int)seg.size();
  if (ns <= 2) return;
  for (int i = 0; i < ns; i++) local[seg[i]] = i;
  vector<double> dg(ns, 0.0);
  vector<vector<std::pair<double, int>>> cand(ns);
  for (int i = 0; i < ns; i++) {
    const int v = seg[i];
    for (int p = rowp[v]; p < rowp[v + 1]; p++) {
      const int c = cols[p];
      if (c == v) dg[i] = std::fabs(vals[p]);
      else if (local[c] >= 0) cand[i].push_back({std::fabs(vals[p]), local[c]});
    }
    std::sort(cand[i].begin(), cand[i].end(), [](const auto& x, const auto& y) {
      return x.first != y.first ? x.first > y.first : x.second < y.second;
    });
  }
  vector<int> order(ns);
  std::iota(order.begin(), order.end(), 0);
  auto need = [&](int i) {
    return cand[i].empty() ? std::numeric_limits<double>::infinity() : dg[i] / cand[i][0].first;
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return need(x) < need(y); });
  vector<int> unit_of(ns, -1);
  vector<std::pair<int, int>> units;  // (first, second or -1)
  for (int i : order) {
    if (unit_of[i] >= 0) continue;
    int partner = -1;
    for (const auto& cv : cand[i])
      if (unit_of[cv.second] < 0) {
        if (dg[i] < ustab * cv.first) partner = cv.second;
        break;
      }
    const int id = (int)units.size();
    unit_of[i] = id;
    if (partner < 0) {
      units.push_back({i, -1});
    } else {
      unit_of[partner] = id;
      units.push_back(dg[i] >= dg[partner] ? std::make_pair(i, partner)
                                           : std::make_pair(partner, i));
    }
  }
  auto key = [&](const std::pair<int, int>& u) {
    return std::min(seg[u.first], u.second >= 0 ? seg[u.second] : INT32_MAX);
  };
  std::stable_sort(units.begin(), units.end(),
                   [&](const auto& x, const auto& y) { return key(x) < key(y); });
  vector<int> out;
  out.reserve(ns);
  for (const auto& u : units) {
    out.push_back(seg[u.first]);
    if (u.second >= 0) out.push_back(seg[u.second]);
  }
  for (int v : seg) local[v] = -1;
  seg.swap(out);
}

// Recursive dissection of one part (a connected or disconnected subgraph of
// the compressed graph).  Appends the elimination order of the part's
// original vertices to `out`.  `seed` is derived from the part's position in
// the dissection tree, so the result does not depend on thread timing.
struct Dissector {
  int n;
  const int* rowp;
  const int* cols;
  const vector<vector<int>>* members;

  void leaf(const Graph& g, const vector<int>& gid, vector<int>& out) const {
    vector<int> verts;
    for (int v = 0; v < g.n; v++)
      for (int o : (*members)[gid[v]]) verts.push_back(o);
    std::sort(verts.begin(), verts.end());
    // local numbering by binary search in the sorted vertex list
    vector<int> lp(verts.size() + 1, 0), lc;
    for (size_t i = 0; i < verts.size(); i++) {
      const int v = verts[i];
      for (int p = rowp[v]; p < rowp[v + 1]; p++) {
        auto it = std::lower_bound(verts.begin(), verts.end(), cols[p]);
        if (it != verts.end() && *it == cols[p]) lc.push_back((int)(it - verts.begin()));
      }
      lp[i + 1] = (int)lc.size();
    }
    vector<int> lperm(verts.size());
    amd_ordering((int)verts.size(), lp.data(), lc.data(), lperm.data());
    for (size_t i = 0; i < verts.size(); i++) out.push_back(verts[lperm[i]]);
  }

  void emit(const vector<int>& gid, const vector<int>& local, bool sorted, vector<int>& out) const {
    const size_t b = out.size();
    for (int v : local)
      for (int o : (*members)[gid[v]]) out.push_back(o);
    if (sorted) std::sort(out.begin() + b, out.end());
  }

  static Graph sub(const Graph& g, const vector<int>& gid, const vector<int>& where, int side,
                   vector<int>& subgid) {
    vector<int> ids, loc(g.n, -1);
    Graph h = induced(g, where, side, ids, loc);
    subgid.resize(ids.size());
    for (size_t i = 0; i < ids.size(); i++) subgid[i] = gid[ids[i]];
    return h;
  }

  void run(const Graph& g, const vector<int>& gid, uint64_t seed, int depth,
           vector<int>& out) const {
    if (g.n == 0) return;
    if (g.total <= kLeaf || g.n <= 3) {
      leaf(g, gid, out);
      return;
    }
    std::mt19937 rng((uint32_t)(seed ^ (seed >> 32)));
    vector<int> dist(g.n, -1), order, where;
    bfs(g, 0, dist, order);
    if ((int)order.size() < g.n) {
      // disconnected: every component on its own, in the order of its
      // smallest local vertex
      vector<int> comp(g.n, -1);
      int nc = 0;
      vector<int> q;
      for (int s0 = 0; s0 < g.n; s0++) {
        if (comp[s0] >= 0) continue;
        q.assign(1, s0);
        comp[s0] = nc;
        for (size_t h = 0; h < q.size(); h++)
          for (int p = g.xadj[q[h]]; p < g.xadj[q[h] + 1]; p++)
            if (comp[g.adj[p]] < 0) {
              comp[g.adj[p]] = nc;
              q.push_back(g.adj[p]);
            }
        nc++;
      }
      // small components are gathered into one leaf-sized part each
      vector<long long> cw(nc, 0);
      for (int v = 0; v < g.n; v++) cw[comp[v]] += g.vw[v];
      vector<int> part(nc);
      int np = 0;
      long long acc = 0;
      for (int c = 0; c < nc; c++) {
        if (cw[c] > kLeaf || acc + cw[c] > kLeaf) {
          if (acc > 0) np++;
          acc = 0;
        }
        part[c] = np;
        acc += cw[c];
        if (cw[c] > kLeaf) {
          np++;
          acc = 0;
        }
      }
      if (acc > 0) np++;
      for (int pi = 0; pi < np; pi++) {
        vector<int> w(g.n, 0), sg;
        for (int v = 0; v < g.n; v++) w[v] = part[comp[v]] == pi ? 0 : 1;
        Graph h = sub(g, gid, w, 0, sg);
        uint64_t sp = seed * 2862933555777941757ull + 3037000493ull * (uint64_t)(pi + 1);
        run(h, sg, sp, depth + 1, out);
      }
      return;
    } else {
      // candidate separators: multilevel (several trials on large parts)
      // and breadth-first; the better balanced-weighted one is kept
      const int trials = g.total >= 20000 ? 3 : 1;
      const int coarse[3] = {1000, 300, 2000};
      vector<vector<int>> cand(trials);
      {
        vector<std::future<void>> jobs;
        for (int t = 0; t < trials; t++) {
          const uint32_t ts = (uint32_t)rng();
          auto job = [&, t, ts] {
            std::mt19937 r(ts);
            ml_bisect(g, cand[t], r, coarse[t]);
          };
          if (t + 1 < trials) jobs.push_back(std::async(std::launch::async, job));
          else job();
        }
        for (auto& j : jobs) j.get();
      }
      double best = 1e300;
      for (int t = 0; t < trials; t++) {
        const Sides sw = side_weights(g, cand[t]);
        const double sc = sep_score(sw) * (balanced(g, sw) ? 1.0 : 1e6);
        if (sc < best) {
          best = sc;
          where.swap(cand[t]);
        }
      }
      {
        vector<int> w;
        bfs_bisect(g, w);
        const Sides sw = side_weights(g, w);
        const double sc = sep_score(sw) * (balanced(g, sw) ? 1.0 : 1e6);
        if (sc < best) {
          best = sc;
          where.swap(w);
        }
      }
      const Sides ss = side_weights(g, where);
      if (ss.w[0] == 0 || ss.w[1] == 0) {
        // no useful separator (e.g. a clique): order the part directly
        vector<int> all(g.n);
        std::iota(all.begin(), all.end(), 0);
        emit(gid, all, false, out);
        return;
      }
    }
    vector<int> ga, gb;
    Graph A = sub(g, gid, where, 0, ga);
    Graph B = sub(g, gid, where, 1, gb);
    vector<int> sep;
    for (int v = 0; v < g.n; v++)
      if (where[v] == 2) sep.push_back(v);
    const uint64_t sa = seed * 6364136223846793005ull + 1442695040888963407ull;
    const uint64_t sb = sa * 6364136223846793005ull + 1442695040888963407ull;
    if (depth < 4 && A.total >= 20000 && B.total >= 20000) {
      vector<int> outa;
      auto fa = std::async(std::launch::async, [&] { run(A, ga, sa, depth + 1, outa); });
      vector<int> outb;
      run(B, gb, sb, depth + 1, outb);
      fa.get();
      out.insert(out.end(), outa.begin(), outa.end());
      out.insert(out.end(), outb.begin(), outb.end());
    } else {
      run(A, ga, sa, depth + 1, out);
      run(B, gb, sb, depth + 1, out);
    }
    // separator last, in increasing original index
    emit(gid, sep, true, out);
  }
};

}  // namespace

void nd_ordering(int n, const int* rowp, const int* cols, int* perm) {
  vector<vector<int>> members;
  Graph G0 = compress(n, rowp, cols, members);
  vector<int> gid(G0.n);
  std::iota(gid.begin(), gid.end(), 0);
  Dissector D{n, rowp, cols, &members};
  vector<int> out;
  out.reserve(n);
  D.run(G0, gid, 20240605u, 0, out);
  for (int i = 0; i < n; i++) perm[i] = out[i];
}

void pair_partners(int n, const int* rowp, const int* cols, const double* vals, double ustab,
                   const SymbolicResult& S, int* perm) {
  // elimination order = the permuted order; a run of consecutive columns j
  // with parent[j] = j + 1 and colcount[j + 1] = colcount[j] - 1 shares its
  // structure below the run (a fundamental supernode), so any order inside
  // the run gives the same fill and the same supernode partition
  vector<int> local(n, -1), seg;
  int j = 0;
  while (j < n) {
    int e = j + 1;
    while (e < n && S.parent[e - 1] == e && S.colcount[e] == S.colcount[e - 1] - 1) e++;
    seg.clear();
    for (int t = j; t < e; t++) seg.push_back(S.permuted ? S.perm[t] : t);
    if (e - j > 2) pair_order(rowp, cols, vals, ustab, seg, local);
    for (int t = j; t < e; t++) perm[t] = seg[t - j];
    j = e;
  }
}

}  // namespace kktldl
