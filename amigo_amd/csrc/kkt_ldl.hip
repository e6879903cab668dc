// MI355X (gfx950) supernodal LDL^T for interior-point KKT systems.
//
// Design (one GPU, HBM-resident):
//  * Host: ordering + symbolic analysis (symbolic.cpp), then a device plan:
//    every supernode ("front") owns a dense column-major slot in one HBM
//    arena that it keeps for the whole factor/solve lifetime.  Nothing is
//    ever copied onto a stack: the factor L of a node is the leading
//    npiv columns of its front, and its contribution (Schur) block is the
//    trailing block of the same front, read in place by the parent.
//  * Fronts are processed level by level (leaves first).  Per level:
//      k_assemble  -- pull-based extend-add: each front entry is written by
//                     exactly one thread that sums the children's
//                     contributions in child order (deterministic, no
//                     atomics), then the node's original K entries and the
//                     barrier diagonal are added.
//      k_front_lds -- fronts up to kLdsFront: assembled, factored and
//                     written back from LDS by one workgroup each, with the
//                     reference's blocked threshold 1x1/2x2 pivoting and
//                     delayed pivots (include/sparse_ldl.h:1316-1553), or
//                     Bunch-Kaufman for roots (the reference calls LAPACK
//                     dsytrf there, :1226-1248).
//      k_panel_rr + k_update (every larger front) -- the same pivoting in
//                     16-column panels on one workgroup with the panel's L
//                     in LDS, and the rank-16 trailing update tiled over
//                     many workgroups with MFMA.
//      k_factor    -- Cholesky mode fronts above kLdsFront.
//  * Delayed pivots make front sizes data dependent: each front slot has
//    slack; a front that overflows marks itself and its ancestors failed,
//    the host enlarges the slack from the observed counts and refactors.
//  * Solve: multifrontal forward pass (front vectors pulled from the
//    children's update vectors, same maps as the assembly) and a backward
//    pass that reads ancestors' solution entries; level scheduled.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kkt_ldl.h"
#include "kkt_common.hpp"
#include "refine_policy.hpp"
#include "symbolic.hpp"

namespace kktldl {

static thread_local std::string g_last_error;

// --------------------------------------------------------------------------
// Per-level preparation: delayed pivots received from the children, their
// offsets inside this front, the actual front dimension, overflow check.
// --------------------------------------------------------------------------
__global__ void k_prepare(DevPlan P, const int* nodes, int count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int p = nodes[t];
  int din = 0, fail = 0;
  for (int q = P.chptr[p]; q < P.chptr[p + 1]; q++) {
    const int c = P.ch[q];
    if (P.stat[c]) {
      fail = 1;
      continue;
    }
    P.doff[c] = din;
    din += P.ndel[c];
  }
  const int m = P.ns[p] + din + P.cs[p];
  P.din[p] = din;
  P.m[p] = m;
  // defined until the node's factor kernel writes them (a node that fails
  // below is skipped by those kernels)
  P.npiv[p] = 0;
  P.ndel[p] = 0;
  if (fail || m > P.cap[p]) {
    P.stat[p] = 1;
    P.need[p] = fail ? -1 : m;
  } else {
    P.stat[p] = 0;
  }
  PanelState st;
  st.k = 0;
  st.nc = P.mode[p] == 1 ? m : P.ns[p] + din;
  st.kstart = 0;
  st.kw = 0;
  st.npos = st.nneg = st.n2 = st.pad = 0;
  P.ps[p] = st;
}

// --------------------------------------------------------------------------
// Assembly: one workgroup per (node, block of kAsmCols columns).  Each
// front entry (i >= j) is produced by one thread as the sum of the
// children's contribution entries that land on it, in child order.
// --------------------------------------------------------------------------
constexpr int kChildCache = 64;

// prep: the node's preparation (k_prepare's work) is done here instead, by
// every workgroup of the node for itself (delay offsets of the children in
// LDS) and written out by the workgroup of chunk 0; only for nodes with at
// most kChildCache children.
template <class T>
__global__ void __launch_bounds__(kTPB) k_assemble(DevPlan P, const int* wnode,
                                                   const int* wchunk, int prep) {
  const int p = wnode[blockIdx.x];
  const int chunk = wchunk[blockIdx.x];
  const int tid = threadIdx.x;
  const int c0 = P.chptr[p], nch = P.chptr[p + 1] - c0;
  const int ns = P.ns[p], cs = P.cs[p];
  __shared__ ChildView<T> s_cv[kChildCache];
  __shared__ int s_list[kChildCache];
  __shared__ int s_rj[kChildCache];
  __shared__ int s_doff[kChildCache];
  __shared__ int s_nl, s_fail, s_din;
  if (prep) {
    if (tid == 0) {
      int din = 0, cfail = 0;
      for (int q = 0; q < nch; q++) {
        const int c = P.ch[c0 + q];
        if (P.stat[c]) {
          cfail = 1;
          continue;
        }
        s_doff[q] = din;
        if (chunk == 0) P.doff[c] = din;
        din += P.ndel[c];
      }
      const int m = ns + din + cs;
      const int fail = cfail || m > P.cap[p];
      if (chunk == 0) {
        P.din[p] = din;
        P.m[p] = m;
        P.npiv[p] = 0;
        P.ndel[p] = 0;
        P.stat[p] = fail;
        if (fail) P.need[p] = cfail ? -1 : m;
        PanelState st;
        st.k = 0;
        st.nc = P.mode[p] == 1 ? m : ns + din;
        st.kstart = 0;
        st.kw = 0;
        st.npos = st.nneg = st.n2 = st.pad = 0;
        P.ps[p] = st;
      }
      s_fail = fail;
      s_din = din;
    }
    __syncthreads();
    if (s_fail) return;
  } else if (P.stat[p]) {
    return;
  }
  const int din = prep ? s_din : P.din[p];
  const int m = ns + din + cs;
  const int j0 = chunk * kAsmCols;
  if (j0 >= m) return;
  const int j1 = min(m, j0 + kAsmCols);
  const int ncache = min(nch, kChildCache);
  for (int t = tid; t < ncache; t += kTPB) {
    s_cv[t] = load_child<T>(P, P.ch[c0 + t]);
    if (prep) s_cv[t].doff = s_doff[t];
  }
  int* vars = P.vars + P.ioff[p];
  int* pre = P.pre + P.ioff[p];
  if (chunk == 0) {
    const int vp = P.vptr[p], cp = P.cptr[p];
    for (int i = tid; i < ns; i += kTPB) vars[i] = P.node_to_var[vp + i];
    for (int i = tid; i < cs; i += kTPB) vars[ns + din + i] = P.contrib_rows[cp + i];
    for (int i = tid; i < m; i += kTPB) pre[i] = i;
    for (int q = 0; q < nch; q++) {
      const int c = P.ch[c0 + q];
      const int nd = P.ndel[c], off = prep ? s_doff[q] : P.doff[c];
      const int* cv = P.vars + P.ioff[c] + P.npiv[c];
      for (int t = tid; t < nd; t += kTPB) vars[ns + off + t] = cv[t];
    }
  }
  __syncthreads();
  T* F = front_ptr<T>(P, p);
  const bool cached = nch <= kChildCache;
  for (int j = j0; j < j1; j++) {
    T* Fj = F + (size_t)j * m;
    if (cached) {
      // children that own position j, compacted in child order (wave 0)
      if (tid < kWave) {
        const int rj = tid < nch ? child_row(P, s_cv[tid], j, ns, din) : -1;
        const unsigned long long mask = __ballot(rj >= 0);
        const int pos = __popcll(mask & ((1ull << tid) - 1ull));
        if (rj >= 0) {
          s_list[pos] = tid;
          s_rj[pos] = rj;
        }
        if (tid == 0) s_nl = __popcll(mask);
      }
      __syncthreads();
      const int nl = s_nl;
      for (int i = j + tid; i < m; i += kTPB) {
        double v = 0.0;
        // four children at a time: the map lookups, then the entries, all
        // in flight together (a child without row i reads its entry at row
        // 0 and does not add it); the sum stays in child order
        for (int t0 = 0; t0 < nl; t0 += 4) {
          int ri[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int r = child_row_nb(P, s_cv[s_list[min(t0 + u, nl - 1)]], i, ns, din);
            ri[u] = t0 + u < nl ? r : -1;
          }
          T x[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int t = min(t0 + u, nl - 1);
            const ChildView<T>& cv = s_cv[s_list[t]];
            const int a = max(ri[u], 0), rj = s_rj[t];
            x[u] = cv.F[max(a, rj) + (size_t)min(a, rj) * cv.m];
          }
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (ri[u] >= 0) v += (double)x[u];
        }
        Fj[i] = (T)v;
      }
      __syncthreads();
    } else {
      for (int i = j + tid; i < m; i += kTPB) {
        double v = 0.0;
        for (int q = 0; q < nch; q++) {
          ChildView<T> cv = load_child<T>(P, P.ch[c0 + q]);
          const int rj = child_row(P, cv, j, ns, din);
          if (rj < 0) continue;
          const int ri = child_row(P, cv, i, ns, din);
          if (ri >= 0) v += (double)cv.F[max(ri, rj) + (size_t)min(ri, rj) * cv.m];
        }
        Fj[i] = (T)v;
      }
    }
  }
  __syncthreads();
  // original entries of the own columns in [j0, j1): distinct positions
  const int vp = P.vptr[p];
  if (j0 < ns) {
    const int e0 = P.am_ptr[vp + j0], e1 = P.am_ptr[vp + min(j1, ns)];
    for (int e = e0 + tid; e < e1; e += kTPB) {
      const int u = P.am_row[e], j = P.am_col[e];
      const size_t r = (u < ns ? u : u + din) + (size_t)j * m;
      F[r] = (T)((double)F[r] + P.vals[P.am_src[e]]);
    }
  }
  __syncthreads();
  if (P.diag)
    for (int j = j0 + tid; j < min(j1, ns); j += kTPB)
      F[(size_t)j * m + j] = (T)((double)F[(size_t)j * m + j] + P.diag[P.node_to_var[vp + j]]);
}

// Assembly with the workgroup's kAsmLdsCols columns, rows [r0, r0 +
// kAsmRows), staged in LDS (front element type): the gather writes LDS only, so
// no global store sits between one row's loads and the next row's (vector
// memory operations complete in order), the original entries and Sigma are
// added there with k_assemble's roundings, and the columns go to the front
// in one pass.  The node is prepared here (the children's delay offsets by
// a prefix sum over one wave; the workgroup of chunk 0 writes the node
// state).  For nodes with at most kChildCache children and fronts of at
// most kAsmLdsRows rows; bitwise equal to k_assemble.  A work item is
// (column chunk c, row chunk r), encoded c * kAsmRowChunks + r: the rows of
// a wide front spread over several workgroups, so that its assembly costs
// one round of map and entry loads per column instead of one per kTPB rows.
constexpr int kAsmLdsRows = 2048;
constexpr int kAsmLdsCols = 4;  // columns per workgroup of k_assemble_lds
constexpr int kAsmRows = kTPB;  // rows per workgroup of k_assemble_lds
constexpr int kAsmRowChunks = kAsmLdsRows / kAsmRows;
template <class T>
__global__ void __launch_bounds__(kTPB) k_assemble_lds(DevPlan P, const int* wnode,
                                                       const int* wchunk) {
  __shared__ T colv[kAsmLdsCols * kAsmRows];
  const int p = wnode[blockIdx.x];
  const int chunk = wchunk[blockIdx.x];
  const int r0 = (chunk % kAsmRowChunks) * kAsmRows;
  const int tid = threadIdx.x, wv = tid / kWave, ln = tid % kWave;
  const int c0 = P.chptr[p], nch = P.chptr[p + 1] - c0;
  const int ns = P.ns[p], cs = P.cs[p];
  const int vp = P.vptr[p];
  const int j0 = (chunk / kAsmRowChunks) * kAsmLdsCols;
  __shared__ ChildView<T> s_cv[kChildCache];
  __shared__ int s_list[kAsmLdsCols][kChildCache];
  __shared__ int s_rj[kAsmLdsCols][kChildCache];
  __shared__ int s_nl[kAsmLdsCols];
  __shared__ int s_fail, s_din;
  // the original entries of the own columns in [j0, j0 + kAsmLdsCols) and
  // their Sigma terms: loaded now, in flight during the preparation
  int e0 = 0, e1 = 0;
  if (j0 < ns) {
    e0 = P.am_ptr[vp + j0];
    e1 = P.am_ptr[vp + min(j0 + kAsmLdsCols, ns)];
  }
  int oe_u = -1, oe_j = 0;
  double oe_v = 0.0;
  if (e0 + tid < e1) {
    const int e = e0 + tid;
    oe_u = P.am_row[e];
    oe_j = P.am_col[e];
    oe_v = P.vals[P.am_src[e]];
  }
  double dg = 0.0;
  if (P.diag && tid < kAsmLdsCols && j0 + tid < ns) dg = P.diag[P.node_to_var[vp + j0 + tid]];
  if (wv == 0) {
    // children's delay counts in child order: offsets by an exclusive prefix
    // sum over the wave (a failed child adds nothing and fails the node)
    int c = 0, st = 0, nd = 0;
    if (ln < nch) {
      c = P.ch[c0 + ln];
      st = P.stat[c];
      nd = P.ndel[c];
    }
    const int nda = st ? 0 : nd;
    int incl = nda;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (ln >= off) incl += y;
    }
    const int doff = incl - nda;
    const int din = __shfl(incl, kWave - 1);
    const bool cfail = __ballot(ln < nch && st != 0) != 0ull;
    if (ln < nch) {
      ChildView<T> v;
      v.F = front_ptr<T>(P, c);
      v.m = P.m[c];
      v.fs = P.ns[c] + P.din[c];
      v.npiv = P.npiv[c];
      v.ndel = nd;
      v.doff = doff;
      v.inv = P.invoff[c];
      s_cv[ln] = v;
      if (chunk == 0 && !st) P.doff[c] = doff;
    }
    if (ln == 0) {
      const int m = ns + din + cs;
      const int fail = cfail || m > P.cap[p];
      if (chunk == 0) {
        P.din[p] = din;
        P.m[p] = m;
        P.npiv[p] = 0;
        P.ndel[p] = 0;
        P.stat[p] = fail;
        if (fail) P.need[p] = cfail ? -1 : m;
        PanelState ps;
        ps.k = 0;
        ps.nc = P.mode[p] == 1 ? m : ns + din;
        ps.kstart = 0;
        ps.kw = 0;
        ps.npos = ps.nneg = ps.n2 = ps.pad = 0;
        P.ps[p] = ps;
      }
      s_fail = fail;
      s_din = din;
    }
  }
  __syncthreads();
  if (s_fail) return;
  const int din = s_din;
  const int m = ns + din + cs;
  if (j0 >= m || r0 >= m) return;
  const int j1 = min(m, j0 + kAsmLdsCols);
  const int r1 = min(m, r0 + kAsmRows);
  if (chunk == 0) {
    int* vars = P.vars + P.ioff[p];
    int* pre = P.pre + P.ioff[p];
    const int cp = P.cptr[p];
    for (int i = tid; i < ns; i += kTPB) vars[i] = P.node_to_var[vp + i];
    for (int i = tid; i < cs; i += kTPB) vars[ns + din + i] = P.contrib_rows[cp + i];
    for (int i = tid; i < m; i += kTPB) pre[i] = i;
    for (int q = 0; q < nch; q++) {
      const ChildView<T>& cv = s_cv[q];
      const int* cvv = P.vars + P.ioff[P.ch[c0 + q]] + cv.npiv;
      for (int t = tid; t < cv.ndel; t += kTPB) vars[ns + cv.doff + t] = cvv[t];
    }
  }
  // the children that own each column, compacted in child order: wave w for
  // column j0 + w
  for (int cc = wv; cc < kAsmLdsCols; cc += kTPB / kWave) {
    const int j = j0 + cc;
    const int rj = (ln < nch && j < j1) ? child_row(P, s_cv[ln], j, ns, din) : -1;
    const unsigned long long mask = __ballot(rj >= 0);
    const int pos = __popcll(mask & ((1ull << ln) - 1ull));
    if (rj >= 0) {
      s_list[cc][pos] = ln;
      s_rj[cc][pos] = rj;
    }
    if (ln == 0) s_nl[cc] = __popcll(mask);
  }
  __syncthreads();
  for (int jj = 0; jj < j1 - j0; jj++) {
    const int j = j0 + jj;
    const int nl = s_nl[jj];
    T* cj = colv + (size_t)jj * kAsmRows;
    for (int i = max(j, r0) + tid; i < r1; i += kTPB) {
      double v = 0.0;
      // four children at a time: the map lookups, then the entries, all in
      // flight together; the sum stays in child order
      for (int t0 = 0; t0 < nl; t0 += 4) {
        int ri[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int r = child_row_nb(P, s_cv[s_list[jj][min(t0 + u, nl - 1)]], i, ns, din);
          ri[u] = t0 + u < nl ? r : -1;
        }
        T x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int t = min(t0 + u, nl - 1);
          const ChildView<T>& cv = s_cv[s_list[jj][t]];
          const int a = max(ri[u], 0), rj = s_rj[jj][t];
          x[u] = cv.F[max(a, rj) + (size_t)min(a, rj) * cv.m];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (ri[u] >= 0) v += (double)x[u];
      }
      cj[i - r0] = (T)v;
    }
  }
  __syncthreads();
  // original entries (distinct positions), then Sigma, as k_assemble adds them
  // (only those in the workgroup's rows)
  if (oe_u >= 0) {
    const int i = oe_u < ns ? oe_u : oe_u + din;
    if (i >= r0 && i < r1) {
      T* c = colv + (size_t)(oe_j - j0) * kAsmRows + (i - r0);
      *c = (T)((double)*c + oe_v);
    }
  }
  for (int e = e0 + kTPB + tid; e < e1; e += kTPB) {
    const int u = P.am_row[e], j = P.am_col[e];
    const int i = u < ns ? u : u + din;
    if (i >= r0 && i < r1) {
      T* c = colv + (size_t)(j - j0) * kAsmRows + (i - r0);
      *c = (T)((double)*c + P.vals[P.am_src[e]]);
    }
  }
  __syncthreads();
  if (P.diag && tid < kAsmLdsCols && j0 + tid < min(j1, ns) && j0 + tid >= r0 &&
      j0 + tid < r1) {
    T* c = colv + (size_t)tid * kAsmRows + (j0 + tid - r0);
    *c = (T)((double)*c + dg);
  }
  __syncthreads();
  T* F = front_ptr<T>(P, p);
  for (int jj = 0; jj < j1 - j0; jj++) {
    const int j = j0 + jj;
    for (int i = max(j, r0) + tid; i < r1; i += kTPB)
      F[(size_t)j * m + i] = colv[(size_t)jj * kAsmRows + (i - r0)];
  }
}

template <class T>
__global__ void __launch_bounds__(kTPB) k_factor(DevPlan P, const int* nodes) {
  constexpr int NT = kTPB;
  const int p = nodes[blockIdx.x];
  if (P.stat[p]) return;
  __shared__ FrontShared<NT> sh;
  const int tid = threadIdx.x;
  const int m = P.m[p];
  const int fs = P.ns[p] + P.din[p];
  const int mode = P.mode[p];
  T* F = front_ptr<T>(P, p);
  int* vars = P.vars + P.ioff[p];
  int* pre = P.pre + P.ioff[p];
  int* ptype = P.ptype + P.ioff[p];
  double* dv = P.dv + P.ioff[p];
  double* dv2 = P.dv2 + P.ioff[p];
  double* W = P.W + P.woff[p];
  const int ldw = P.cap[p];
  if (tid == 0) {
    sh.npos = 0;
    sh.nneg = 0;
    sh.n2 = 0;
  }
  __syncthreads();
  if (mode == 2) {
    cholesky_front<NT, T>(F, m, m, fs);
    for (int j = tid; j < fs; j += NT) {
      ptype[j] = 1;
      dv[j] = (double)F[j + (size_t)j * m];
    }
    if (tid == 0) {
      P.npiv[p] = fs;
      P.ndel[p] = 0;
      P.cnt[3 * p] = 0;
      P.cnt[3 * p + 1] = 0;
      P.cnt[3 * p + 2] = 0;
    }
    return;
  }
  // panel width as the reference (min(block_size, front_size)); at least 2
  // so that a 1 x 1 root front still takes its single pivot
  const int nb = max(2, min(kNB, m));
  int k = 0, nc = (mode == 1) ? m : fs;
  while (k < nc) {
    const int kstart = k;
    int kw = 0;
    panel_factor<NT, T>(F, m, m, vars, pre, ptype, dv, dv2, W, ldw, nb, mode, P.ustab,
                        P.ptol, k, nc, kw, sh);
    // rows of W below k hold L*D for the panel columns
    trailing_update_wg<NT, T>(F, m, m, W, ldw, kstart, k, kw);
  }
  if (tid == 0) {
    P.npiv[p] = nc;
    P.ndel[p] = fs - nc;
    P.cnt[3 * p] = sh.npos;
    P.cnt[3 * p + 1] = sh.nneg;
    P.cnt[3 * p + 2] = sh.n2;
  }
}

// --------------------------------------------------------------------------
// Small fronts: assembled straight into LDS, factored there, written back
// once.  LDS layout: F (m x m, ld m) then W (m x kNB, ld m).
// --------------------------------------------------------------------------
template <int NT, class T>
__global__ void __launch_bounds__(NT, 4) k_front_lds(DevPlan P, const int* nodes) {
  extern __shared__ double lds[];
  __shared__ FrontShared<NT> sh;
  __shared__ ChildView<T> s_cv[kChildCache];
  const int p = nodes[blockIdx.x];
  if (P.stat[p]) return;
  const int tid = threadIdx.x;
  const int m = P.m[p], ns = P.ns[p], cs = P.cs[p], din = P.din[p];
  const int fs = ns + din, mode = P.mode[p];
  const int c0 = P.chptr[p], nch = P.chptr[p + 1] - c0;
  const bool cached = nch <= kChildCache;
  if (cached)
    for (int t = tid; t < nch; t += NT) s_cv[t] = load_child<T>(P, P.ch[c0 + t]);
  // LDS: W (m x kNB, FP64) | dv | dv2 | F (m x m, T) | vars | pre | ptype
  double* W = lds;
  double* dv = W + (size_t)m * kNB;
  double* dv2 = dv + m;
  T* F = (T*)(dv2 + m);
  int* vars = (int*)(F + (size_t)m * m);
  int* pre = vars + m;
  int* ptype = pre + m;
  {
    const int vp = P.vptr[p], cp = P.cptr[p];
    for (int i = tid; i < ns; i += NT) vars[i] = P.node_to_var[vp + i];
    for (int i = tid; i < cs; i += NT) vars[fs + i] = P.contrib_rows[cp + i];
    for (int i = tid; i < m; i += NT) pre[i] = i;
    for (int q = 0; q < nch; q++) {
      const int c = P.ch[c0 + q];
      const int nd = P.ndel[c], off = P.doff[c];
      const int* cv = P.vars + P.ioff[c] + P.npiv[c];
      for (int t = tid; t < nd; t += NT) vars[ns + off + t] = cv[t];
    }
  }
  for (int i = tid; i < m; i += NT) ptype[i] = 0;
  if (tid == 0) {
    sh.npos = 0;
    sh.nneg = 0;
    sh.n2 = 0;
  }
  __syncthreads();
  // pull the children's contributions into the lower triangle
  for (int e = tid; e < m * m; e += NT) {
    const int j = e / m, i = e - j * m;
    double v = 0.0;
    if (i >= j) {
      for (int q = 0; q < nch; q++) {
        ChildView<T> cv = cached ? s_cv[q] : load_child<T>(P, P.ch[c0 + q]);
        const int rj = child_row(P, cv, j, ns, din);
        if (rj < 0) continue;
        const int ri = child_row(P, cv, i, ns, din);
        if (ri >= 0) v += (double)cv.F[max(ri, rj) + (size_t)min(ri, rj) * cv.m];
      }
    }
    F[e] = (T)v;
  }
  __syncthreads();
  const int vp = P.vptr[p];
  // every original entry of the own columns lands on a distinct position
  for (int e = P.am_ptr[vp] + tid; e < P.am_ptr[vp + ns]; e += NT) {
    const int u = P.am_row[e], j = P.am_col[e];
    const size_t r = (u < ns ? u : u + din) + (size_t)j * m;
    F[r] = (T)((double)F[r] + P.vals[P.am_src[e]]);
  }
  __syncthreads();
  if (P.diag)
    for (int j = tid; j < ns; j += NT)
      F[(size_t)j * m + j] = (T)((double)F[(size_t)j * m + j] + P.diag[P.node_to_var[vp + j]]);
  __syncthreads();
  if (mode == 2) {
    cholesky_front<NT, T>(F, m, m, fs);
    for (int j = tid; j < fs; j += NT) {
      ptype[j] = 1;
      dv[j] = (double)F[j + (size_t)j * m];
    }
    if (tid == 0) {
      P.npiv[p] = fs;
      P.ndel[p] = 0;
      P.cnt[3 * p] = P.cnt[3 * p + 1] = P.cnt[3 * p + 2] = 0;
    }
  } else {
    const int nb = max(2, min(kNB, m));
    int k = 0, nc = (mode == 1) ? m : fs;
    while (k < nc) {
      const int kstart = k;
      int kw = 0;
      panel_factor<NT, T>(F, m, m, vars, pre, ptype, dv, dv2, W, m, nb, mode, P.ustab,
                          P.ptol, k, nc, kw, sh);
      trailing_update_wg<NT, T>(F, m, m, W, m, kstart, k, kw);
    }
    if (tid == 0) {
      P.npiv[p] = nc;
      P.ndel[p] = fs - nc;
      P.cnt[3 * p] = sh.npos;
      P.cnt[3 * p + 1] = sh.nneg;
      P.cnt[3 * p + 2] = sh.n2;
    }
  }
  __syncthreads();
  T* Fg = front_ptr<T>(P, p);
  for (int e = tid; e < m * m; e += NT) Fg[e] = F[e];
  const long long io = P.ioff[p];
  for (int i = tid; i < m; i += NT) {
    P.vars[io + i] = vars[i];
    P.pre[io + i] = pre[i];
    P.ptype[io + i] = ptype[i];
    P.dv[io + i] = dv[i];
    P.dv2[io + i] = dv2[i];
  }
}

// FP32 fronts use v_mfma_f32_16x16x4f32 (FP32 accumulate), whose output is
// D[4(l/16) + r][l%16] (also checked on gfx950); FP64 fronts
// v_mfma_f64_16x16x4f64, whose output is D[l/16 + 4r][l%16].
template <class T>
__global__ void __launch_bounds__(kTPB) k_update(DevPlan P, const int* unode,
                                                 const int* ubase, const int* upref,
                                                 int nfront) {
  typedef T vec4 __attribute__((ext_vector_type(4)));
  // the workgroup's front: the last f whose first tile upref[f] is at most
  // blockIdx.x (the prefixes increase); every wave finds it by a ballot over
  // 64 fronts at a time, so the result is uniform without a barrier
  const int g = blockIdx.x;
  int p = 0, base = 0, t = 0;
  {
    const int ln = threadIdx.x % kWave;
    for (int f0 = 0; f0 < nfront; f0 += kWave) {
      const int q = min(f0 + ln, nfront - 1);
      const int pr = upref[q], nd = unode[q], bs = ubase[q];
      const unsigned long long b = __ballot(f0 + ln < nfront && pr <= g);
      if (b == 0ull) break;
      const int src = 63 - __clzll((long long)b);
      p = __shfl(nd, src);
      base = __shfl(bs, src);
      t = g - __shfl(pr, src);
      if (b != ~0ull) break;
    }
    // uniform values: the node's metadata then load through the scalar cache
    p = __builtin_amdgcn_readfirstlane(p);
    base = __builtin_amdgcn_readfirstlane(base);
    t = __builtin_amdgcn_readfirstlane(t);
  }
  // everything per node depends on p alone: one round of loads
  const int stat = P.stat[p];
  const PanelState st = P.ps[p];
  const int m = P.m[p];
  T* F = front_ptr<T>(P, p);
  const double* W = P.W + P.woff[p];
  const int ldw = P.cap[p];
  const int kw = st.kw;
  if (stat || kw == 0) return;
  const int k = st.k, kstart = st.kstart;
  int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > t) bi--;
  while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
  const int bj = t - bi * (bi + 1) / 2;
  const int i0 = base + 64 * bi, j0 = base + 64 * bj;
  if (i0 >= m || j0 >= m || i0 + 64 <= k || j0 + 64 <= k) return;
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  const int ri = i0 + 32 * (w >> 1), cj = j0 + 32 * (w & 1);
  if (ri + 32 <= cj) return;  // strictly above the diagonal
  const int lr = l & 15, lk = l >> 4;
  constexpr bool f64 = sizeof(T) == 8;
  // Register r of lane (lk, lr) holds the tile entry at row 4 lk + r,
  // column lr of its 16x16 subtile, so a lane's four rows are consecutive
  // in the column: one 16-byte (FP32) or two (FP64) accesses per subtile.
  // FP32 MFMA output row 4 (l/16) + r is that row already.  The FP64 MFMA
  // outputs row l/16 + 4 r; its A operand therefore supplies, in MFMA row
  // x = q + 4 s, the L row 4 q + s, which moves output row l/16 + 4 r to
  // row 4 (l/16) + r.
  const int arow = f64 ? 4 * (lr & 3) + (lr >> 2) : lr;
  // Accumulator starts from the F tile; the A operand is -L, so the MFMA
  // chain produces F - L W^T directly.  All loads are issued up front, from
  // addresses clamped into the front (and into W's kNBLarge columns), and
  // masked afterwards: a load under a condition becomes a branch, and the
  // compiler then waits for each such load before issuing the next.  The
  // tile reads can pass the front's last entry by up to 3 entries, inside
  // the arena's padding.
  typedef T vec4u __attribute__((ext_vector_type(4), aligned(sizeof(T))));
  vec4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = min(cj + 16 * b + lr, m - 1);
      const int i = min(ri + 16 * a + 4 * lk, m - 1);
      const vec4u v = *reinterpret_cast<const vec4u*>(F + i + (size_t)j * m);
#pragma unroll
      for (int r = 0; r < 4; r++) acc[a][b][r] = v[r];
    }
  constexpr int KS = kNBLarge / 4;  // k-steps of 4 (kw <= kNBLarge)
  T av[KS][2], bv[KS][2];
  double wv[KS][2];
#pragma unroll
  for (int q = 0; q < KS; q++) {
    const int c = 4 * q + lk;
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const int row = min(ri + 16 * a + arow, m - 1);
      av[q][a] = F[row + (size_t)min(kstart + c, m - 1) * m];
      const int col = min(cj + 16 * a + lr, m - 1);
      wv[q][a] = W[col + (size_t)c * ldw];
    }
  }
  // rows and columns of the tile beyond m only reach outputs that are not
  // stored; the k columns beyond kw must be zero
#pragma unroll
  for (int q = 0; q < KS; q++) {
    const bool ok = 4 * q + lk < kw;
#pragma unroll
    for (int a = 0; a < 2; a++) {
      av[q][a] = ok ? -av[q][a] : (T)0;
      bv[q][a] = ok ? (T)wv[q][a] : (T)0;
    }
  }
#pragma unroll
  for (int q = 0; q < KS; q++)
    if (4 * q < kw) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
          if constexpr (f64)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q][a], bv[q][b], acc[a][b], 0, 0, 0);
          else
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q][a], bv[q][b], acc[a][b], 0, 0, 0);
        }
    }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = cj + 16 * b + lr;
      const int i = ri + 16 * a + 4 * lk;
      if (j < m && j >= k) {
        if (i >= j && i + 3 < m) {
          vec4u v;
#pragma unroll
          for (int r = 0; r < 4; r++) v[r] = acc[a][b][r];
          *reinterpret_cast<vec4u*>(F + i + (size_t)j * m) = v;
        } else {
#pragma unroll
          for (int r = 0; r < 4; r++)
            if (i + r < m && i + r >= j) F[i + r + (size_t)j * m] = acc[a][b][r];
        }
      }
    }
}

// --------------------------------------------------------------------------
// Solve.  Forward: front vector in LDS (post-pivot order), pulled from the
// right-hand side and the children's update vectors; the unit-lower L11 is
// applied in blocks of kSB columns: the diagonal block by one wave with the
// block held in registers, the rows below by the whole workgroup.
// --------------------------------------------------------------------------
constexpr int kSB = 64;
// small-front solve launches of at most this many workgroups take the
// register-resident diagonal block (REG)
constexpr int kRegLaunch = 256;
constexpr int kTinyFront = 64;  // fronts up to this dimension: one wave each
constexpr int kTinyJoin = 64;   // at most this many tiny fronts join a level's larger ones
// a level's small fronts share a launch with its large nodes' first launch
// (k_fwd_level / k_bwd_level) when they are at most this many and the rows
// staged by the backward products fit in kMergeLds bytes
constexpr int kMergeSmall = 512;
constexpr size_t kMergeLds = 64 * 1024;

// End of the solve block starting at jb: kSB columns, shortened so that a
// 2x2 pivot pair is never split.
__device__ inline int block_end(const int* ptype, int jb, int npiv, int chol) {
  int je = min(npiv, jb + kSB);
  if (!chol && je < npiv && ptype[je - 1] == 2) je--;
  return je;
}

// BIG nodes (the large fronts, few per level) only gather here: the
// pivot rows of the front vector go to Y and the rows below to U, and the
// panel steps (k_fwd_step) solve them over many workgroups.
// The diagonal block L[jb:jb+bs, jb:jb+bs] into LDS (column major, leading
// dimension ldb) as doubles, eight loads of a thread in flight together.
template <class T, int NTH>
__device__ __attribute__((always_inline)) inline void stage_block(const T* L, int m, int jb,
                                                                  int bs, double* Lb, int ldb) {
  const int n = bs * bs;
  for (int e0 = threadIdx.x; e0 < n; e0 += 8 * NTH) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = min(e0 + u * NTH, n - 1);
      const int c = e / bs, rr = e - c * bs;
      v[u] = (double)L[jb + rr + (size_t)(jb + c) * m];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * NTH;
      if (e < n) {
        const int c = e / bs, rr = e - c * bs;
        Lb[rr + c * ldb] = v[u];
      }
    }
  }
}

// t[j] -= sum_{i0 <= i < i1} L[i, j] t[i] for the columns j0 <= j < j1:
// one wave per column, lanes over i in steps of kWave (strided_dot's order),
// then a butterfly.  With at most two terms per lane, CB columns of a wave
// have their loads in flight together.
// One node p for right-hand side r: front vector t (LDS, P.lds_t doubles,
// then the diagonal block), child views s_cv / s_uo (LDS, kChildCache).
// Ends with a barrier, so that a workgroup can go on to another node.
template <class T, int NTH, bool REG>
__device__ __attribute__((always_inline)) inline void forward_node(
    const DevPlan& P, int p, int r, double* x, int ldx, long long ustride, int chol, double* t,
    ChildView<T>* s_cv, long long* s_uo) {
  const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
  const int m = P.m[p], npiv = P.npiv[p];
  const int ns = P.ns[p], din = P.din[p];
  const int fs = ns + din;
  const T* L = front_ptr<T>(P, p);
  const int* vars = P.vars + P.ioff[p];
  const int* pre = P.pre + P.ioff[p];
  const int* ptype = P.ptype + P.ioff[p];
  const double* dv = P.dv + P.ioff[p];
  const double* dv2 = P.dv2 + P.ioff[p];
  const int c0 = P.chptr[p], c1 = P.chptr[p + 1];
  const int vp = P.vptr[p];
  const int nch = c1 - c0;
  const bool cached = nch <= kChildCache;
  if (cached)
    for (int q = tid; q < nch; q += NTH) {
      s_cv[q] = load_child<T>(P, P.ch[c0 + q]);
      s_uo[q] = P.uoff[P.ch[c0 + q]];
    }
  __syncthreads();
  {
    double* xr = x + (size_t)r * ldx;
    for (int i = tid; i < m; i += NTH) {
      const int pi = i < fs ? pre[i] : i;
      // loads issued unconditionally (indices clamped), then selected
      const double xv = xr[P.node_to_var[vp + min(pi, ns - 1)]];
      double v = pi < ns ? xv : 0.0;
      if (cached) {
        // four children at a time: map lookups, then update-vector entries,
        // all in flight together; the sum stays in child order
        for (int q0 = 0; q0 < nch; q0 += 4) {
          int rc[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int rq = child_row_nb(P, s_cv[min(q0 + u, nch - 1)], pi, ns, din);
            rc[u] = q0 + u < nch ? rq : -1;
          }
          double uv[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int q = min(q0 + u, nch - 1);
            uv[u] = P.U[s_uo[q] + (long long)r * ustride + max(rc[u] - s_cv[q].npiv, 0)];
          }
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (rc[u] >= 0) v += uv[u];
        }
      } else {
        for (int q = 0; q < nch; q++) {
          const ChildView<T> cv = load_child<T>(P, P.ch[c0 + q]);
          const long long uo = P.uoff[P.ch[c0 + q]];
          const int rc = child_row(P, cv, pi, ns, din);
          if (rc >= 0) v += P.U[uo + (long long)r * ustride + (rc - cv.npiv)];
        }
      }
      t[i] = v;
    }
    __syncthreads();
    for (int jb = 0; jb < npiv;) {
      const int je = block_end(ptype, jb, npiv, chol);
      const int bs = je - jb;
      const int ldb = P.lds_b;
      double* Lb = t + P.lds_t;  // bs x bs block, column major, leading dimension ldb
      if (!REG) {
        stage_block<T, NTH>(L, m, jb, bs, Lb, ldb);
        __syncthreads();
      }
      if (w == 0) {
        // one wave, lane = row of the block; one short step per column, the
        // block's entries eight columns at a time (REG: the whole row in
        // registers, one round of loads from HBM; else from LDS).  Column c
        // updates the rows below it; the first column of a 2x2 pair skips
        // the pair's second row (that slot is D's coupling, not L), which
        // takes its own column's step next with its value unchanged.
        const int row = jb + lane;
        double tr = lane < bs ? t[row] : 0.0;
        const double dr = (chol && lane < bs) ? dv[row] : 1.0;
        const int pt = (!chol && lane < bs) ? ptype[row] : 1;
        auto steps8 = [&](int c0, const double(&lq)[8]) {
          if (chol) {  // L scaled by the diagonal
#pragma unroll
            for (int u = 0; u < 8; u++) {
              const int c = c0 + u;
              if (c >= bs) continue;
              if (lane == c) tr = tr / dr;
              const double tc = lane_d(tr, c);
              if (lane > c && lane < bs) tr -= lq[u] * tc;
            }
          } else {
#pragma unroll
            for (int u = 0; u < 8; u++) {
              const int c = c0 + u;
              if (c >= bs) continue;
              const int ty = __builtin_amdgcn_readlane(pt, c);
              const double tc = lane_d(tr, c);
              if (lane > c + (ty == 2) && lane < bs) tr -= lq[u] * tc;
            }
          }
        };
        if constexpr (REG) {
          const size_t ro = (size_t)min(row, m - 1);
          double lr[kSB];
#pragma unroll
          for (int c = 0; c < kSB; c++) lr[c] = (double)L[ro + (size_t)(jb + min(c, bs - 1)) * m];
#pragma unroll
          for (int c0 = 0; c0 < kSB; c0 += 8) {
            if (c0 >= bs) continue;
            double lq[8];
#pragma unroll
            for (int u = 0; u < 8; u++) lq[u] = lr[c0 + u];
            steps8(c0, lq);
          }
        } else {
          const double* Lr = Lb + min(lane, bs - 1);
          for (int c0 = 0; c0 < bs; c0 += 8) {
            double lq[8];
#pragma unroll
            for (int u = 0; u < 8; u++) lq[u] = Lr[min(c0 + u, bs - 1) * ldb];
            steps8(c0, lq);
          }
        }
        if (lane < bs) t[row] = tr;
      }
      __syncthreads();
      for (int i = je + tid; i < m; i += NTH)
        t[i] = sub_dot(t[i], L + i + (size_t)jb * m, m, t + jb, je - jb);
      __syncthreads();
      jb = je;
    }
    double* U = P.U + P.uoff[p] + (long long)r * ustride;
    for (int i = npiv + tid; i < m; i += NTH) U[i - npiv] = t[i];
    if (!chol) {
      // D solve, one thread per 1x1 pivot or 2x2 block
      for (int j = tid; j < npiv; j += NTH) {
        const int ty = ptype[j];
        if (ty == 1) {
          t[j] = t[j] / dv[j];
        } else if (ty == 2) {
          const double d11 = dv[j], d21 = dv2[j], d22 = dv[j + 1];
          const double inv = 1.0 / (d11 * d22 - d21 * d21);
          const double a = t[j], b = t[j + 1];
          t[j] = inv * (d22 * a - d21 * b);
          t[j + 1] = inv * (d11 * b - d21 * a);
        }
      }
      __syncthreads();
    }
    for (int j = tid; j < npiv; j += NTH) xr[unpiv(vars[j])] = t[j];
    __syncthreads();
  }
}

template <class T, int NTH, bool REG>
__global__ void __launch_bounds__(NTH) k_forward(DevPlan P, const int* nodes,
                                                 double* x, int ldx, int nrhs,
                                                 long long ustride, int chol) {
  extern __shared__ double t[];
  __shared__ ChildView<T> s_cv[kChildCache];
  __shared__ long long s_uo[kChildCache];
  // one node per workgroup, one right-hand side per grid row
  forward_node<T, NTH, REG>(P, nodes[blockIdx.x], blockIdx.y, x, ldx, ustride, chol, t, s_cv,
                            s_uo);
}

// A run of levels holding one small node each (a chain: each node's
// children are at earlier levels): one workgroup per right-hand side walks
// the nodes in level order, one launch for the run.
template <class T, int NTH>
__global__ void __launch_bounds__(NTH) k_forward_chain(DevPlan P, const int* nodes, int count,
                                                       double* x, int ldx, long long ustride,
                                                       int chol) {
  extern __shared__ double t[];
  __shared__ ChildView<T> s_cv[kChildCache];
  __shared__ long long s_uo[kChildCache];
  for (int q = 0; q < count; q++)
    forward_node<T, NTH, true>(P, nodes[q], blockIdx.y, x, ldx, ustride, chol, t, s_cv, s_uo);
}

// Backward: t1 = x[pivots] - L21^T x[rows below], then the L11^T back solve
// in blocks from the bottom (dot products over the rows below a block by
// the whole workgroup, the diagonal block by one wave in registers).
// BIG nodes are solved by k_bwd_gemv and the panel steps (k_bwd_step).
template <class T, int NTH>
__device__ __attribute__((always_inline)) inline void col_dots(const T* L, int m, double* t,
                                                               int j0, int j1, int i0, int i1) {
  constexpr int NW = NTH / kWave;
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (i1 - i0 <= 2 * kWave) {
    constexpr int CB = 8;
    const int ia = i0 + lane, ib = ia + kWave;
    const int ra = min(ia, max(i1 - 1, 0)), rb = min(ib, max(i1 - 1, 0));
    const double t0 = ia < i1 ? t[ia] : 0.0, t1 = ib < i1 ? t[ib] : 0.0;
    for (int jc = j0 + w; jc < j1; jc += CB * NW) {
      double a0[CB], a1[CB];
#pragma unroll
      for (int u = 0; u < CB; u++) {
        const size_t co = (size_t)min(jc + u * NW, j1 - 1) * m;
        a0[u] = (double)L[ra + co];
        a1[u] = (double)L[rb + co];
      }
#pragma unroll
      for (int u = 0; u < CB; u++) {
        double s = 0.0;
        if (ia < i1) s += a0[u] * t0;
        if (ib < i1) s += a1[u] * t1;
        for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
        const int j = jc + u * NW;
        if (lane == 0 && j < j1) t[j] -= s;
      }
    }
  } else {
    for (int j = j0 + w; j < j1; j += NW) {
      double s = strided_dot(L + (size_t)j * m, t, i0 + lane, i1, kWave);
      for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) t[j] -= s;
    }
  }
}

template <class T, int NTH, bool REG>
__device__ __attribute__((always_inline)) inline void backward_node(const DevPlan& P, int p, int r,
                                                                    double* x, int ldx, int chol,
                                                                    double* t) {
  const int tid = threadIdx.x;
  const int m = P.m[p], npiv = P.npiv[p];
  const int fs = P.ns[p] + P.din[p];
  const T* L = front_ptr<T>(P, p);
  const int* vars = P.vars + P.ioff[p];
  const int* ptype = P.ptype + P.ioff[p];
  const double* dv = P.dv + P.ioff[p];
  const int cp = P.cptr[p];
  const int w = tid / kWave, lane = tid % kWave;
  {
    double* xr = x + (size_t)r * ldx;
    for (int i = tid; i < m; i += NTH) {
      int v = i < npiv ? unpiv(vars[i]) : (i < fs ? vars[i] : P.contrib_rows[cp + i - fs]);
      t[i] = xr[v];
    }
    __syncthreads();
    // t[j] -= sum_{i >= npiv} L[i,j] t[i]
    col_dots<T, NTH>(L, m, t, 0, npiv, npiv, m);
    __syncthreads();
    for (int je = npiv; je > 0;) {
      int jb = max(0, je - kSB);
      if (!chol && jb > 0 && ptype[jb] == 3) jb++;
      const int bs = je - jb;
      // rows of L11 below the block are final: dot products per column
      if (je < npiv) {
        col_dots<T, NTH>(L, m, t, jb, je, je, npiv);
        __syncthreads();
      }
      const int ldb = P.lds_b;
      double* Lb = t + P.lds_t;
      if (!REG) {
        stage_block<T, NTH>(L, m, jb, bs, Lb, ldb);
        __syncthreads();
      }
      if (w == 0) {
        // lane c owns column jb + c of the block; one short step per row,
        // from the bottom, the block's entries eight rows at a time (REG:
        // the whole column in registers, one round of loads from HBM; else
        // from LDS).  Row i is final once the rows after it are in acc (a
        // 2x2 pair's two rows both at its second row, whose step then skips
        // the pair's first column: that slot is D's coupling, not L); its
        // value then enters acc.
        const int col = jb + lane;
        const double dc = (chol && lane < bs) ? dv[col] : 1.0;
        double tv = lane < bs ? t[col] : 0.0;
        double acc = 0.0;
        const int pt = (!chol && lane < bs) ? ptype[col] : 1;
        auto steps8 = [&](int i0, const double(&lq)[8]) {
          if (chol) {
#pragma unroll
            for (int u = 0; u < 8; u++) {
              const int i = i0 - u;
              if (i < 0 || i >= bs) continue;
              if (lane == i) tv = (tv - acc) / dc;
              const double ti = lane_d(tv, i);
              if (lane < i) acc += lq[u] * ti;
            }
          } else {
#pragma unroll
            for (int u = 0; u < 8; u++) {
              const int i = i0 - u;
              if (i < 0 || i >= bs) continue;
              const int ty = __builtin_amdgcn_readlane(pt, i);
              if ((ty != 2 && lane == i) || (ty == 3 && lane == i - 1)) tv = tv - acc;
              const double ti = lane_d(tv, i);
              if (lane < i - (ty == 3)) acc += lq[u] * ti;
            }
          }
        };
        if constexpr (REG) {
          const T* Lc = L + (size_t)min(col, je - 1) * m + jb;
          double lc[kSB];
#pragma unroll
          for (int i = 0; i < kSB; i++) lc[i] = (double)Lc[min(i, bs - 1)];
#pragma unroll
          for (int i0 = kSB - 1; i0 >= 0; i0 -= 8) {
            if (i0 - 7 >= bs) continue;
            double lq[8];
#pragma unroll
            for (int u = 0; u < 8; u++) lq[u] = lc[i0 - u];
            steps8(i0, lq);
          }
        } else {
          const double* Lc = Lb + (size_t)min(lane, bs - 1) * ldb;
          for (int i0 = bs - 1; i0 >= 0; i0 -= 8) {
            double lq[8];
#pragma unroll
            for (int u = 0; u < 8; u++) lq[u] = Lc[max(i0 - u, 0)];
            steps8(i0, lq);
          }
        }
        if (lane < bs) t[col] = tv;
      }
      __syncthreads();
      je = jb;
    }
    for (int j = tid; j < npiv; j += NTH) xr[unpiv(vars[j])] = t[j];
    __syncthreads();
  }
}

template <class T, int NTH, bool REG>
__global__ void __launch_bounds__(NTH) k_backward(DevPlan P, const int* nodes,
                                                  double* x, int ldx, int nrhs, int chol) {
  extern __shared__ double t[];
  // one node per workgroup, one right-hand side per grid row
  backward_node<T, NTH, REG>(P, nodes[blockIdx.x], blockIdx.y, x, ldx, chol, t);
}

// The backward sweep of a chain (k_forward_chain): the nodes in reverse
// level order, parents first.
template <class T, int NTH>
__global__ void __launch_bounds__(NTH) k_backward_chain(DevPlan P, const int* nodes, int count,
                                                        double* x, int ldx, int chol) {
  extern __shared__ double t[];
  for (int q = count - 1; q >= 0; q--)
    backward_node<T, NTH, true>(P, nodes[q], blockIdx.y, x, ldx, chol, t);
}


// --------------------------------------------------------------------------
// Large-node solves by partitioned inverse.  The pivot columns of a large
// node are cut into panels of kPB columns.  After the factorization the
// diagonal block B_s of every panel is inverted once by k_panel_inv.  B_s is
// unit lower triangular, with the slot L[j + 1, j] of every 2x2 pivot (j,
// j + 1) taken as zero: that slot does not belong to L (D holds the
// coupling), also where the pair is split by a panel boundary.  k_panel_inv2
// then composes the inverse of every pair of consecutive panels.  A solve
// walks the pivots in steps of two panels (kPS columns) with one launch per
// step for all large nodes of a level (and one grid row per right-hand
// side): every workgroup forms x_s = B_s^{-1} y_s (forward) or B_s^{-T} y_s
// (backward) itself, from the kPS entries of y_s and the composed inverse,
// and applies x_s to its own tile of the rows below the step (forward) or of
// the columns before it (backward).  A step therefore costs one memory round
// trip spread over many CUs, instead of kPS dependent column steps inside
// one workgroup.
// --------------------------------------------------------------------------
constexpr int kPB = 64;        // panel width of the partitioned inverse
constexpr int kStepRows = 64;  // forward step: rows per workgroup (one per lane)
constexpr int kStepCols = 32;  // backward step: columns per workgroup (8 per wave)
constexpr int kStepRhs = 4;    // right-hand sides per workgroup of a step (operands reused)
constexpr int kGatherRhs = 4;  // right-hand sides per workgroup of a large node's gather
constexpr long long kPinvPanel = 2LL * kPB * kPB;  // doubles per panel: B^{-1}, B^{-T}
constexpr int kPS = 2 * kPB;   // solve step width: two panels
// doubles per solve step: both panels' B^{-1}, B^{-T}, then X and X^T (below)
constexpr long long kPinvPair = 3 * kPinvPanel;

// One wave per (node, panel s): lane c computes column c of B_s^{-1} by
// forward substitution against B_s held in LDS (row major, read as
// broadcasts), with the column in registers; the inverse is written
// transposed (lanes = columns) and, through LDS, column major (lanes =
// rows), so that both solve directions read it coalesced.  Nodes that failed
// (stat != 0) are refactored on a new plan and skipped here.
template <class T>
__global__ void __launch_bounds__(kWave) k_panel_inv(DevPlan P, const int* inode,
                                                     const int* ipanel) {
  constexpr int LB = kPB + 1;         // padded row of Bs
  __shared__ double Bs[kPB * LB];     // B[i][j] at i * LB + j; later Z[i][c] at i * LB + c
  const int p = inode[blockIdx.x], s = ipanel[blockIdx.x];
  if (P.stat[p] != 0) return;
  const int m = P.m[p], npiv = P.npiv[p];
  if (m > P.cap[p] || npiv > m) return;
  const int b0 = s * kPB;
  if (b0 >= npiv) return;
  const int w = min(kPB, npiv - b0);
  const T* L = front_ptr<T>(P, p);
  const int* ptype = P.ptype + P.ioff[p];
  const int lane = threadIdx.x;
  // 16 columns at a time, every load issued before the first use (from
  // addresses clamped into the front, masked afterwards)
  const size_t ri = (size_t)min(b0 + lane, m - 1);
  for (int j0 = 0; j0 < kPB; j0 += 16) {
    T raw[16];
    int pt[16];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      raw[jj] = L[ri + (size_t)min(b0 + j0 + jj, m - 1) * m];
      pt[jj] = ptype[min(b0 + j0 + jj, npiv - 1)];
    }
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int i = lane, j = j0 + jj;
      double v = 0.0;
      if (i < w && j < w) {
        if (i == j) v = 1.0;
        else if (i > j && !(i == j + 1 && pt[jj] == 2)) v = (double)raw[jj];
      }
      Bs[i * LB + j] = v;
    }
  }
  __syncthreads();
  // z_i = -sum_{j < i} B[i][j] z_j for i > c (z_c = 1, z_j = 0 for j < c)
  double z[kPB];
#pragma unroll
  for (int i = 0; i < kPB; i++) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int j = 0; j < i; j++) {
      const double bij = Bs[i * LB + j];
      if (j & 1) a1 = __builtin_fma(bij, z[j], a1);
      else a0 = __builtin_fma(bij, z[j], a0);
    }
    z[i] = i == lane ? 1.0 : (i > lane ? -(a0 + a1) : 0.0);
  }
  double* inv = P.pinv + P.pinvoff[p] + (long long)(s / 2) * kPinvPair + (s % 2) * kPinvPanel;
  // transposed copy: row c of B^{-T} is column c of B^{-1}, held by lane c
#pragma unroll
  for (int i = 0; i < kPB; i++) inv[kPB * kPB + lane + i * kPB] = z[i];
  __syncthreads();  // every lane is done with B
#pragma unroll
  for (int i = 0; i < kPB; i++) Bs[i * LB + lane] = z[i];
  __syncthreads();
  for (int c = 0; c < kPB; c++) inv[lane + c * kPB] = Bs[lane * LB + c];
}

// Composition of two consecutive 64-column blocks into one 128-column solve
// step: for B = [B1 0; C B2] (C = L rows of block 2, columns of block 1, with
// the slot of a 2x2 pivot split by the block boundary taken as zero),
// B^{-1} = [B1^{-1} 0; X B2^{-1}] with X = -B2^{-1} C B1^{-1}.  One wave per
// (node, block pair) after k_panel_inv: lane c forms column c of C B1^{-1}
// from column c of B1^{-1} (registers) and C (LDS broadcasts), then column
// c of X against B2^{-1} (LDS broadcasts), and writes X and X^T.
template <class T>
__global__ void __launch_bounds__(kWave) k_panel_inv2(DevPlan P, const int* inode,
                                                      const int* ipair) {
  constexpr int LB = kPB + 1;
  __shared__ double S[kPB * LB];  // C[r][j] at r * LB + j; then B2^{-1}; then X
  const int p = inode[blockIdx.x], q = ipair[blockIdx.x];
  if (P.stat[p] != 0) return;
  const int m = P.m[p], npiv = P.npiv[p];
  if (m > P.cap[p] || npiv > m) return;
  const int ba = 2 * q * kPB, bb = ba + kPB;
  if (bb >= npiv) return;
  const int wb = min(kPB, npiv - bb);
  const T* L = front_ptr<T>(P, p);
  const int* ptype = P.ptype + P.ioff[p];
  double* base = P.pinv + P.pinvoff[p] + (long long)q * kPinvPair;
  const double* invaT = base + kPB * kPB;          // row c = column c of B1^{-1}
  const double* invb = base + kPinvPanel;          // column major
  double* X = base + 2 * kPinvPanel;
  const bool cut = ptype[bb - 1] == 2;
  const int lane = threadIdx.x;
  // 16 columns at a time, every load issued before the first use
  const size_t rb = (size_t)min(bb + lane, m - 1);
  for (int j0 = 0; j0 < kPB; j0 += 16) {
    T raw[16];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) raw[jj] = L[rb + (size_t)(ba + j0 + jj) * m];
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int j = j0 + jj;
      S[lane * LB + j] =
          (lane < wb && !(cut && lane == 0 && j == kPB - 1)) ? (double)raw[jj] : 0.0;
    }
  }
  // column c of B1^{-1} (zero above c)
  double t[kPB];
#pragma unroll
  for (int j = 0; j < kPB; j++) t[j] = invaT[lane + j * kPB];
  __syncthreads();
  // u = C z (u overwrites z from the top: u[r] needs all of z)
  double u[kPB];
#pragma unroll
  for (int r = 0; r < kPB; r++) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int j = 0; j < kPB; j++) {
      const double crj = S[r * LB + j];
      if (j & 1) a1 = __builtin_fma(crj, t[j], a1);
      else a0 = __builtin_fma(crj, t[j], a0);
    }
    u[r] = a0 + a1;
  }
  __syncthreads();  // C no longer needed
  for (int i0 = 0; i0 < kPB; i0 += 16) {  // B2^{-1}[i][lane], 16 loads in flight
    double raw[16];
#pragma unroll
    for (int ii = 0; ii < 16; ii++) raw[ii] = invb[i0 + ii + lane * kPB];
#pragma unroll
    for (int ii = 0; ii < 16; ii++) S[(i0 + ii) * LB + lane] = raw[ii];
  }
  __syncthreads();
  // column c of X = -B2^{-1} u (B2^{-1} lower triangular)
#pragma unroll
  for (int i = 0; i < kPB; i++) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int r = 0; r <= i; r++) {
      const double bir = S[i * LB + r];
      if (r & 1) a1 = __builtin_fma(bir, u[r], a1);
      else a0 = __builtin_fma(bir, u[r], a0);
    }
    t[i] = -(a0 + a1);
  }
  // X^T (row c = column c of X, held by lane c), then X column major through LDS
#pragma unroll
  for (int i = 0; i < kPB; i++) X[kPB * kPB + lane + i * kPB] = t[i];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPB; i++) S[i * LB + lane] = t[i];
  __syncthreads();
  for (int c = 0; c < kPB; c++) X[lane + c * kPB] = S[lane * LB + c];
}

// Entry i (post-pivot order) of a BIG node's front vector for the
// right-hand sides r0 .. r0 + nb - 1 (RB slots, the last repeated): x of the
// row's variable (zero for delayed and contribution rows) plus the
// children's update-vector entries in child order, four children's lookups
// and loads in flight at a time (the order of k_forward's gather).
template <class T, int RB>
__device__ __attribute__((always_inline)) inline void gather_big_row(
    const DevPlan& P, int i, int fs, int ns, int din, const int* pre, int vp, int c0, int nch,
    const ChildView<T>* s_cv, const long long* s_uo, const double* x, int ldx, int r0, int nb,
    long long ustride, double (&v)[RB]) {
  const int pi = i < fs ? pre[i] : i;
  const int var = P.node_to_var[vp + min(pi, ns - 1)];
#pragma unroll
  for (int rr = 0; rr < RB; rr++) {
    const double xv = x[(size_t)(r0 + min(rr, nb - 1)) * ldx + var];
    v[rr] = pi < ns ? xv : 0.0;
  }
  for (int q0 = 0; q0 < nch; q0 += 4) {
    int rc[4];
    ChildView<T> cvs[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int qq = min(q0 + u, nch - 1);
      cvs[u] = qq < kChildCache ? s_cv[qq] : load_child<T>(P, P.ch[c0 + qq]);
      const int rq = child_row_nb(P, cvs[u], pi, ns, din);
      rc[u] = q0 + u < nch ? rq : -1;
    }
#pragma unroll
    for (int rr = 0; rr < RB; rr++) {
      double uv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int qq = min(q0 + u, nch - 1);
        const long long uo = qq < kChildCache ? s_uo[qq] : P.uoff[P.ch[c0 + qq]];
        uv[u] = P.U[uo + (long long)(r0 + min(rr, nb - 1)) * ustride +
                    max(rc[u] - cvs[u].npiv, 0)];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (rc[u] >= 0) v[rr] += uv[u];
    }
  }
}

// Forward step s of the large nodes of a level: the 128 columns
// [b0, b0 + 128) as two 64-blocks a, b.  Y holds the pivot rows of each
// front vector (gathered by k_forward and updated by the earlier steps), U
// the rows below the pivots.  Workgroup (node, tile):
//   x_a = B_a^{-1} y_a,  x_b = X y_a + B_b^{-1} y_b          (lanes = rows)
//   y/U[i] -= L[i, b0:b1] x  for its kStepRows rows i >= b1
// Tile 0 also applies D^{-1} to x and writes the step's entries of x.  A
// 2x2 pivot (b1 - 1, b1) split by the step boundary is finished by tile 0 of
// the same step: the next step's block is unit lower triangular, so the
// solved entry of row b1 is its y after this step's update, which tile 0
// computes.
// G (step 0 only): the workgroup gathers the front vector entries it needs
// itself (the step's rows y_s and its own tile's rows, as k_gather_big would
// write them), instead of a gather launch before the step; its rows are
// written once, after the update.
template <class T, int RB, bool G>
__device__ __attribute__((always_inline)) inline void fwd_step_wg(
    const DevPlan& P, int bx, int by, const int* wnode, const int* wtile, int s, double* x,
    int ldx, int nrhs, long long ustride, double* Y, long long ystride) {
  constexpr int NW = kTPB / kWave;
  constexpr int CQ = kPB / NW;  // block columns per wave (inverse products)
  constexpr int CR = kPS / NW;  // step columns per wave (row update)
  __shared__ double ys[RB][kPS];
  __shared__ double xp[RB][kPS];
  __shared__ double part[2][RB][NW][kWave];
  __shared__ double gown[G ? RB : 1][kWave];
  __shared__ ChildView<T> s_cv[G ? kChildCache : 1];
  __shared__ long long s_uo[G ? kChildCache : 1];
  const int p = wnode[bx], tile = wtile[bx];
  // everything per node depends on p alone: one round of loads
  const int m = P.m[p], npiv = P.npiv[p];
  const T* L = front_ptr<T>(P, p);
  const long long pio = P.pinvoff[p], io = P.ioff[p];
  const int b0 = s * kPS;
  if (b0 >= npiv) return;
  const int b1 = min(npiv, b0 + kPS), w = b1 - b0;
  const int wa = min(w, kPB), wb = w - wa;
  const int i0 = b1 + tile * kStepRows;
  if (tile > 0 && i0 >= m) return;
  const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave;
  const double* base = P.pinv + pio + (long long)s * kPinvPair;
  const double* inva = base;
  const double* invb = base + kPinvPanel;
  const double* X = base + 2 * kPinvPanel;
  const int* ptype = P.ptype + io;
  // a 2x2 pivot split by the step boundary: its coupling slot L[b1, b1 - 1]
  // is not part of L and must not enter the update of row b1
  const bool cut = b1 < npiv && ptype[b1 - 1] == 2;
  const int i = i0 + lane;
  const size_t irow = (size_t)min(i, m - 1);
  // right-hand-side independent operands, loaded once from clamped (always
  // valid) addresses
  double lrow[CR], aa[CQ], ax[CQ], ab[CQ];
#pragma unroll
  for (int q = 0; q < CR; q++) {
    const int c = wv * CR + q;
    const double lv = (double)L[irow + (size_t)(b0 + min(c, w - 1)) * m];
    lrow[q] = (i < m && c < w && !(cut && i == b1 && c == w - 1)) ? lv : 0.0;
  }
#pragma unroll
  for (int q = 0; q < CQ; q++) {
    const int c = wv * CQ + q;
    aa[q] = c < wa ? inva[lane + c * kPB] : 0.0;
    ax[q] = (c < wa && wb > 0) ? X[lane + c * kPB] : 0.0;
    ab[q] = c < wb ? invb[lane + c * kPB] : 0.0;
  }
  const int* vars = P.vars + io;
  const double* dv = P.dv + io;
  const double* dv2 = P.dv2 + io;
  // RB right-hand sides per grid row, on the operands loaded once, all
  // with one set of barriers: y_s of every right-hand side staged in LDS
  // first (each sum keeps the order of the one-right-hand-side form)
  const int r0 = by * RB;
  const int nr = min(nrhs - r0, RB);
  if constexpr (G) {
    // the node's children, then the lanes of wave 0 gather the tile's rows
    // and the next kPS threads the step's rows y_s
    const int ns = P.ns[p], din = P.din[p], fs = ns + din;
    const int* pre = P.pre + io;
    const int c0 = P.chptr[p], nch = P.chptr[p + 1] - c0;
    const int vp = P.vptr[p];
    for (int q = tid; q < min(nch, kChildCache); q += kTPB) {
      s_cv[q] = load_child<T>(P, P.ch[c0 + q]);
      s_uo[q] = P.uoff[P.ch[c0 + q]];
    }
    __syncthreads();
    if (tid < kWave) {
      if (i < m) {
        double v[RB];
        gather_big_row<T, RB>(P, i, fs, ns, din, pre, vp, c0, nch, s_cv, s_uo, x, ldx, r0, nr,
                              ustride, v);
#pragma unroll
        for (int rr = 0; rr < RB; rr++) gown[rr][tid] = v[rr];
      }
    } else if (tid - kWave < w) {
      const int c = tid - kWave;
      double v[RB];
      gather_big_row<T, RB>(P, b0 + c, fs, ns, din, pre, vp, c0, nch, s_cv, s_uo, x, ldx, r0,
                            nr, ustride, v);
#pragma unroll
      for (int rr = 0; rr < RB; rr++)
        if (rr < nr) ys[rr][c] = v[rr];
    }
    for (int e = tid; e < RB * kPS; e += kTPB) {
      const int rr = e / kPS, c = e % kPS;
      if (rr >= nr || c >= w) ys[rr][c] = 0.0;
    }
  } else {
    for (int e = tid; e < RB * kPS; e += kTPB) {
      const int rr = e / kPS, c = e % kPS;
      ys[rr][c] = (rr < nr && c < w) ? Y[(long long)(r0 + rr) * ystride + io + b0 + c] : 0.0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < RB; rr++) {
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int q = 0; q < CQ; q++) {
      const int c = wv * CQ + q;
      const double ya = ys[rr][c];
      const double yb = ys[rr][kPB + c];
      sa += aa[q] * ya;
      sb += ax[q] * ya + ab[q] * yb;
    }
    part[0][rr][wv][lane] = sa;
    part[1][rr][wv][lane] = sb;
  }
  __syncthreads();
  // x_s per right-hand side: wave wv sums right-hand sides wv, wv + NW
  for (int rr = wv; rr < RB; rr += NW) {
    xp[rr][lane] = lane < wa ? ((part[0][rr][0][lane] + part[0][rr][1][lane]) +
                                part[0][rr][2][lane]) + part[0][rr][3][lane]
                             : 0.0;
    xp[rr][kPB + lane] = lane < wb ? ((part[1][rr][0][lane] + part[1][rr][1][lane]) +
                                      part[1][rr][2][lane]) + part[1][rr][3][lane]
                                   : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < RB; rr++) {
    double g = 0.0;
#pragma unroll
    for (int q = 0; q < CR; q++) g += lrow[q] * xp[rr][wv * CR + q];
    part[0][rr][wv][lane] = g;
  }
  __syncthreads();
  for (int rr = wv; rr < nr; rr += NW) {
    const int r = r0 + rr;
    double* y = Y + (long long)r * ystride + io;
    double* xr = x + (size_t)r * ldx;
    if (i < m) {
      const double sum = ((part[0][rr][0][lane] + part[0][rr][1][lane]) + part[0][rr][2][lane]) +
                         part[0][rr][3][lane];
      if (i < npiv) {
        const double yn = (G ? gown[rr][lane] : y[i]) - sum;
        y[i] = yn;
        if (cut && tile == 0 && lane == 0) {  // i == b1: the pair (b1 - 1, b1)
          const int j = b1 - 1;
          const double d11 = dv[j], d21 = dv2[j], d22 = dv[j + 1];
          const double dinv = 1.0 / (d11 * d22 - d21 * d21);
          const double ta = xp[rr][w - 1], tb = yn;
          // (G: the step's own pivot results go to Y, see below)
          (G ? y[j] : xr[unpiv(vars[j])]) = dinv * (d22 * ta - d21 * tb);
          xr[unpiv(vars[j + 1])] = dinv * (d11 * tb - d21 * ta);
        }
      } else {
        double& u = P.U[P.uoff[p] + (long long)r * ustride + (i - npiv)];
        u = (G ? gown[rr][lane] : u) - sum;
      }
    }
  }
  if (tile == 0) {
    // 1x1 pivots and 2x2 pivots inside the step (also across its 64-column
    // blocks); a pair split by the step boundary is written above (second
    // pivot first in the step: by the step before).  G: the other
    // workgroups of the node read x of these pivots' variables (their
    // right-hand side) while gathering, so the results go to Y (rows
    // 0 .. b1 - 1, which nothing else reads) and the backward sweep takes
    // them from there.
    for (int e = tid; e < nr * kPS; e += kTPB) {
      const int rr = e / kPS, jj = e % kPS;
      if (jj >= w) continue;
      double* xr = x + (size_t)(r0 + rr) * ldx;
      double* yr = Y + (long long)(r0 + rr) * ystride + io;
      const int j = b0 + jj;
      const int ty = ptype[j];
      if (ty == 1) {
        (G ? yr[j] : xr[unpiv(vars[j])]) = xp[rr][jj] / dv[j];
      } else if (ty == 2 && jj + 1 < w) {
        const double d11 = dv[j], d21 = dv2[j], d22 = dv[j + 1];
        const double dinv = 1.0 / (d11 * d22 - d21 * d21);
        const double ta = xp[rr][jj], tb = xp[rr][jj + 1];
        (G ? yr[j] : xr[unpiv(vars[j])]) = dinv * (d22 * ta - d21 * tb);
        (G ? yr[j + 1] : xr[unpiv(vars[j + 1])]) = dinv * (d11 * tb - d21 * ta);
      }
    }
  }
}

template <class T, int RB, bool G>
__global__ void __launch_bounds__(kTPB) k_fwd_step(DevPlan P, const int* wnode,
                                                   const int* wtile, int s, double* x,
                                                   int ldx, int nrhs, long long ustride,
                                                   double* Y, long long ystride) {
  fwd_step_wg<T, RB, G>(P, blockIdx.x, blockIdx.y, wnode, wtile, s, x, ldx, nrhs, ustride, Y,
                        ystride);
}

// Backward step s: step q = (number of steps) - 1 - s, from the bottom.  Y
// holds the right-hand side of L11^T (k_bwd_gemv), updated by the earlier
// steps.  Workgroup (node, tile):
//   x_b = B_b^{-T} y_b,  x_a = B_a^{-T} y_a + X^T y_b        (lanes = rows)
//   y[c] -= L[b0:b1, c]^T x  for its kStepCols columns c < b0
// (one wave per column, lanes over the step's rows, two per lane).  Tile 0
// writes x.
template <class T, int RB>
__global__ void __launch_bounds__(kTPB) k_bwd_step(DevPlan P, const int* wnode,
                                                   const int* wtile, int s, double* x,
                                                   int ldx, int nrhs, double* Y,
                                                   long long ystride) {
  constexpr int NW = kTPB / kWave;
  constexpr int CQ = kPB / NW;         // block rows per wave (inverse products)
  constexpr int CPW = kStepCols / NW;  // tile columns per wave
  __shared__ double ys[RB][kPS];
  __shared__ double xp[RB][kPS];
  __shared__ double part[2][RB][NW][kWave];
  const int p = wnode[blockIdx.x], tile = wtile[blockIdx.x];
  // everything per node depends on p alone: one round of loads
  const int m = P.m[p], npiv = P.npiv[p];
  const T* L = front_ptr<T>(P, p);
  const long long pio = P.pinvoff[p], io = P.ioff[p];
  const int q = (npiv + kPS - 1) / kPS - 1 - s;
  if (q < 0) return;
  const int b0 = q * kPS, w = min(kPS, npiv - b0);
  const int wa = min(w, kPB), wb = w - wa;
  const int c0 = tile * kStepCols;
  if (tile > 0 && c0 >= b0) return;
  const int tid = threadIdx.x, wv = tid / kWave, lane = tid % kWave;
  const double* base = P.pinv + pio + (long long)q * kPinvPair;
  const double* invaT = base + kPB * kPB;
  const double* invbT = base + kPinvPanel + kPB * kPB;
  const double* XT = base + 2 * kPinvPanel + kPB * kPB;
  const int cmax = max(b0 - 1, 0);
  // a 2x2 pivot split by the step boundary: L[b0, b0 - 1] is not part of L
  const bool cut = b0 > 0 && P.ptype[io + b0 - 1] == 2;
  double lcol[CPW][2], ta[CQ], tx[CQ], tb[CQ];
#pragma unroll
  for (int k = 0; k < CPW; k++) {
    const int c = c0 + wv * CPW + k;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int rr = h * kPB + lane;
      const double lv = (double)L[(size_t)(b0 + min(rr, w - 1)) + (size_t)min(c, cmax) * m];
      lcol[k][h] = (rr < w && c < b0 && !(cut && rr == 0 && c == b0 - 1)) ? lv : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < CQ; k++) {
    const int i = wv * CQ + k;
    ta[k] = i < wa ? invaT[lane + i * kPB] : 0.0;
    tx[k] = i < wb ? XT[lane + i * kPB] : 0.0;
    tb[k] = i < wb ? invbT[lane + i * kPB] : 0.0;
  }
  const int* vars = P.vars + io;
  // RB right-hand sides per grid row, on the operands loaded once, with one
  // set of barriers (each sum in the order of the one-right-hand-side form)
  const int r0 = blockIdx.y * RB;
  const int nr = min(nrhs - r0, RB);
  for (int e = tid; e < RB * kPS; e += kTPB) {
    const int rr = e / kPS, c = e % kPS;
    ys[rr][c] = (rr < nr && c < w) ? Y[(long long)(r0 + rr) * ystride + io + b0 + c] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < RB; rr++) {
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int k = 0; k < CQ; k++) {
      const int i = wv * CQ + k;
      const double ya = ys[rr][i];
      const double yb = ys[rr][kPB + i];
      sa += ta[k] * ya + tx[k] * yb;
      sb += tb[k] * yb;
    }
    part[0][rr][wv][lane] = sa;
    part[1][rr][wv][lane] = sb;
  }
  __syncthreads();
  for (int rr = wv; rr < RB; rr += NW) {
    xp[rr][lane] = lane < wa ? ((part[0][rr][0][lane] + part[0][rr][1][lane]) +
                                part[0][rr][2][lane]) + part[0][rr][3][lane]
                             : 0.0;
    xp[rr][kPB + lane] = lane < wb ? ((part[1][rr][0][lane] + part[1][rr][1][lane]) +
                                      part[1][rr][2][lane]) + part[1][rr][3][lane]
                                   : 0.0;
  }
  __syncthreads();
  // the wave's CPW columns for every right-hand side: full sums in every
  // lane, then lane rr * CPW + k updates column k of right-hand side rr (one
  // read-modify-write per lane, all in flight together)
  static_assert(RB * CPW <= kWave, "one lane per (right-hand side, column)");
  double gsel = 0.0;
#pragma unroll
  for (int rr = 0; rr < RB; rr++) {
    const double xl0 = xp[rr][lane], xl1 = xp[rr][kPB + lane];
#pragma unroll
    for (int k = 0; k < CPW; k++) {
      double g = lcol[k][0] * xl0 + lcol[k][1] * xl1;
      for (int off = kWave / 2; off > 0; off >>= 1) g += __shfl_xor(g, off);
      if (lane == rr * CPW + k) gsel = g;
    }
  }
  {
    const int rr = lane / CPW, k = lane % CPW;
    const int c = c0 + wv * CPW + k;
    if (rr < nr && c < b0) Y[(long long)(r0 + rr) * ystride + io + c] -= gsel;
  }
  if (tile == 0)
    for (int e = tid; e < nr * kPS; e += kTPB) {
      const int rr = e / kPS, jj = e % kPS;
      if (jj < w) x[(size_t)(r0 + rr) * ldx + unpiv(vars[b0 + jj])] = xp[rr][jj];
    }
}

// Pivot columns of a BIG node, backward: Y[j] = x[pivot j] - sum_{i >= npiv}
// L[i, j] x[i], the right-hand side of L11^T, for 16 columns per workgroup
// (4 per wave, lanes over the rows, coalesced), with the node's x rows
// staged in LDS.
constexpr int kGemvCols = 16;
// RB right-hand sides per workgroup with their x rows staged in LDS; RB = 0:
// one right-hand side, x read directly (rows below the pivots beyond LDS).
// yfw: the level's forward sweep left the results of each node's first
// kPS pivots in Y (the first step gathered its own front vector), so they
// are read there instead of from x.
template <class T, int RB>
__device__ __attribute__((always_inline)) inline void bwd_gemv_wg(
    const DevPlan& P, int bx, int by, const int* wnode, const int* wtile, const double* x,
    int ldx, int nrhs, double* Y, long long ystride, int ldxs, int yfw,
    double* xs) {  // xs: LDS for x of the rows below the pivots, RB right-hand sides
  const int p = wnode[bx];
  const int m = P.m[p], npiv = P.npiv[p];
  const int j0 = kGemvCols * wtile[bx];
  if (j0 >= npiv) return;
  const int fs = P.ns[p] + P.din[p];
  const int cp = P.cptr[p];
  const int* vars = P.vars + P.ioff[p];
  const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
  const T* L = front_ptr<T>(P, p);
  const int nr = m - npiv;
  const int yk = yfw ? min(npiv, kPS) : 0;
  constexpr int RR = RB > 0 ? RB : 1;
  const int r0 = by * RR;
  const int nb = min(nrhs - r0, RR);
  // both index loads issued (clamped into their arrays; a node without
  // contribution rows reads vars twice), then the gather
  const int cs = P.cs[p];
  const int* crow = cs > 0 ? P.contrib_rows + cp : vars;
  auto xrow = [&](int q, int rr) {  // x of row npiv + q, right-hand side r0 + rr
    const int i = npiv + q;
    const int iv = vars[min(i, fs - 1)];
    const int ic = crow[min(max(i - fs, 0), max(cs - 1, 0))];
    return x[(size_t)(r0 + rr) * ldx + (i < fs ? iv : ic)];
  };
  if constexpr (RB > 0) {
    for (int q = tid; q < nr; q += kTPB)
#pragma unroll
      for (int rr = 0; rr < RB; rr++)
        if (rr < nb) xs[rr * ldxs + q] = xrow(q, rr);
    __syncthreads();
  }
  // the wave's four columns together, so that their loads are in flight
  // at once (columns past npiv read column npiv - 1 and are not written);
  // each column accumulates over q = lane, lane + 64, ... in order, for
  // every right-hand side on the same L loads
  constexpr int CW = kGemvCols / 4;
  const int jb = j0 + 4 * w;
  const T* Lc[CW];
#pragma unroll
  for (int jj = 0; jj < CW; jj++) Lc[jj] = L + (size_t)min(jb + jj, npiv - 1) * m + npiv;
  double sc[CW][RR];
#pragma unroll
  for (int jj = 0; jj < CW; jj++)
#pragma unroll
    for (int rr = 0; rr < RR; rr++) sc[jj][rr] = 0.0;
  auto xv = [&](int q, int rr) { return RB > 0 ? xs[rr * ldxs + q] : xrow(q, rr); };
  int q = lane;
  for (; q + kWave < nr; q += 2 * kWave) {
    T a[CW][2];
#pragma unroll
    for (int jj = 0; jj < CW; jj++) {
      a[jj][0] = Lc[jj][q];
      a[jj][1] = Lc[jj][q + kWave];
    }
#pragma unroll
    for (int rr = 0; rr < RR; rr++) {
      const double x0 = xv(q, rr), x1 = xv(q + kWave, rr);
#pragma unroll
      for (int jj = 0; jj < CW; jj++) {
        sc[jj][rr] = __builtin_fma((double)a[jj][0], x0, sc[jj][rr]);
        sc[jj][rr] = __builtin_fma((double)a[jj][1], x1, sc[jj][rr]);
      }
    }
  }
  if (q < nr) {
#pragma unroll
    for (int rr = 0; rr < RR; rr++) {
      const double x0 = xv(q, rr);
#pragma unroll
      for (int jj = 0; jj < CW; jj++)
        sc[jj][rr] = __builtin_fma((double)Lc[jj][q], x0, sc[jj][rr]);
    }
  }
#pragma unroll
  for (int jj = 0; jj < CW; jj++)
#pragma unroll
    for (int rr = 0; rr < RR; rr++)
      for (int off = kWave / 2; off > 0; off >>= 1) sc[jj][rr] += __shfl_xor(sc[jj][rr], off);
  // every lane holds the sums; lane rr * CW + jj writes column jj of
  // right-hand side rr
  static_assert(RR * CW <= kWave, "one lane per (right-hand side, column)");
  double v = 0.0;
#pragma unroll
  for (int rr = 0; rr < RR; rr++)
#pragma unroll
    for (int jj = 0; jj < CW; jj++)
      if (lane == rr * CW + jj) v = sc[jj][rr];
  {
    const int rr = lane / CW, j = jb + lane % CW;
    if (rr < nb && j < npiv) {
      const double* xr = x + (size_t)(r0 + rr) * ldx;
      double* yj = Y + (long long)(r0 + rr) * ystride + P.ioff[p] + j;
      *yj = (j < yk ? *yj : xr[unpiv(vars[j])]) - v;
    }
  }
}

template <class T, int RB>
__global__ void __launch_bounds__(kTPB) k_bwd_gemv(DevPlan P, const int* wnode,
                                                   const int* wtile, const double* x,
                                                   int ldx, int nrhs, double* Y,
                                                   long long ystride, int ldxs, int yfw) {
  extern __shared__ double xs[];
  bwd_gemv_wg<T, RB>(P, blockIdx.x, blockIdx.y, wnode, wtile, x, ldx, nrhs, Y, ystride, ldxs,
                     yfw, xs);
}

// Forward gather of a BIG node for RB right-hand sides per grid row: the
// pivot rows of the front vector to Y, the rows below to U (the panel steps
// solve them).  Each row's child lookups are done once for all the
// workgroup's right-hand sides; each entry sums its children in child order
// (the order of k_forward's gather).
template <class T, int RB>
__global__ void __launch_bounds__(1024) k_gather_big(DevPlan P, const int* nodes, double* x,
                                                     int ldx, int nrhs, long long ustride,
                                                     double* Y, long long ystride) {
  constexpr int NTH = 1024;
  const int p = nodes[blockIdx.x];
  const int tid = threadIdx.x;
  const int m = P.m[p], npiv = P.npiv[p];
  const int ns = P.ns[p], din = P.din[p];
  const int fs = ns + din;
  const long long io = P.ioff[p];
  const int* pre = P.pre + io;
  const int c0 = P.chptr[p], c1 = P.chptr[p + 1];
  const int vp = P.vptr[p];
  __shared__ ChildView<T> s_cv[kChildCache];
  __shared__ long long s_uo[kChildCache];
  const int nch = c1 - c0;
  for (int q = tid; q < min(nch, kChildCache); q += NTH) {
    s_cv[q] = load_child<T>(P, P.ch[c0 + q]);
    s_uo[q] = P.uoff[P.ch[c0 + q]];
  }
  __syncthreads();
  const int r0 = blockIdx.y * RB;
  const int nb = min(nrhs - r0, RB);
  double* U = P.U + P.uoff[p];
  for (int i = tid; i < m; i += NTH) {
    double v[RB];
    gather_big_row<T, RB>(P, i, fs, ns, din, pre, vp, c0, nch, s_cv, s_uo, x, ldx, r0, nb,
                          ustride, v);
#pragma unroll
    for (int rr = 0; rr < RB; rr++) {
      if (rr >= nb) break;
      const long long r = r0 + rr;
      if (i < npiv) Y[r * ystride + io + i] = v[rr];
      else U[r * ustride + (i - npiv)] = v[rr];
    }
  }
}

// One launch for a level's small fronts and the first forward step of its
// large nodes, which gathers their front vectors itself (one right-hand
// side): the two are independent (every child of either is at an earlier
// level).  Workgroups 0 .. nsmall - 1 take the small fronts (LDS shape of P),
// the rest the step's tiles.
template <class T, bool REG>
__global__ void __launch_bounds__(kTPB) k_fwd_level(DevPlan P, const int* snodes, int nsmall,
                                                    double* x, int ldx, long long ustride,
                                                    int chol, const int* wnode, const int* wtile,
                                                    double* Y, long long ystride) {
  if ((int)blockIdx.x < nsmall) {
    extern __shared__ double t[];
    __shared__ ChildView<T> s_cv[kChildCache];
    __shared__ long long s_uo[kChildCache];
    forward_node<T, kTPB, REG>(P, snodes[blockIdx.x], 0, x, ldx, ustride, chol, t, s_cv, s_uo);
  } else {
    fwd_step_wg<T, 1, true>(P, blockIdx.x - nsmall, 0, wnode, wtile, 0, x, ldx, 1, ustride, Y,
                            ystride);
  }
}

// The backward counterpart: a level's small fronts (their parents are at
// later levels, done) and the L21^T products of its large nodes in one
// launch (one right-hand side; the products stage the rows below in LDS).
template <class T, bool REG>
__global__ void __launch_bounds__(kTPB) k_bwd_level(DevPlan P, const int* snodes, int nsmall,
                                                    double* x, int ldx, int chol,
                                                    const int* bnode, const int* btile, double* Y,
                                                    long long ystride, int ldxs, int yfw) {
  extern __shared__ double lds[];
  if ((int)blockIdx.x < nsmall)
    backward_node<T, kTPB, REG>(P, snodes[blockIdx.x], 0, x, ldx, chol, lds);
  else
    bwd_gemv_wg<T, 1>(P, blockIdx.x - nsmall, 0, bnode, btile, x, ldx, 1, Y, ystride, ldxs, yfw,
                      lds);
}

// r = b - (K + diag(d)) x for nrhs columns; 16 lanes per row.
__global__ void k_residual(int n, const int* rowp, const int* cols,
                           const double* vals, const double* diag,
                           const double* x, const double* b, double* r,
                           int ld, int nrhs) {
  const int lane16 = threadIdx.x & 15;
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  if (row >= n) return;
  for (int c = 0; c < nrhs; c++) {
    const double* xc = x + (size_t)c * ld;
    double s = 0.0;
    for (int p = rowp[row] + lane16; p < rowp[row + 1]; p += 16) s += vals[p] * xc[cols[p]];
    for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
    if (lane16 == 0) {
      double dv = diag ? diag[row] * xc[row] : 0.0;
      r[row + (size_t)c * ld] = b[row + (size_t)c * ld] - (s + dv);
    }
  }
}

// Sum of squares of each column: partial sums per (column, block), then a
// fixed-order final sum per column (deterministic, no atomics).
constexpr int kNormBlocks = 128;
constexpr int kMaxNormRhs = 64;  // right-hand sides whose norms use fixed host buffers

__global__ void k_colnorm2_part(int n, const double* x, int ld, double* part) {
  __shared__ double red[kTPB / kWave];
  const int c = blockIdx.y;
  const double* xc = x + (size_t)c * ld;
  double s = 0.0;
  for (int i = blockIdx.x * kTPB + threadIdx.x; i < n; i += kNormBlocks * kTPB)
    s += xc[i] * xc[i];
  for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kTPB / kWave; i++) t += red[i];
    part[(size_t)c * kNormBlocks + blockIdx.x] = t;
  }
}

__global__ void k_colnorm2_final(int nrhs, const double* part, double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  double t = 0.0;
  for (int b = 0; b < kNormBlocks; b++) t += part[(size_t)c * kNormBlocks + b];
  out[c] = t;
}

// x += dx over nrhs columns, keeping the previous x in xb (the refinement
// returns to it if the step raised the residual)
__global__ void k_axpy_cols(int n, double* x, const double* dx, double* xb, int ld, int nrhs) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t tot = (size_t)n * nrhs;
  if (i >= tot) return;
  size_t c = i / n, r = i % n;
  const double v = x[r + c * ld];
  xb[r + c * ld] = v;
  x[r + c * ld] = v + dx[r + c * ld];
}

// Zero and non-finite pivots of the last factorization, one wave per node:
// out[2p] = exact zero 1x1 pivots (Cholesky: zero diagonal of L) plus 2x2
// blocks with det = 0, out[2p+1] = pivots whose FP64 D entries are NaN or
// infinite.  Reads the FP64 D that the inertia counts come from.
__global__ void __launch_bounds__(kWave) k_pivot_check(DevPlan P, int* out) {
  const int p = blockIdx.x;
  // pivots lie in the node's slot; a node that failed holds none
  const int npiv = P.stat[p] ? 0 : min(max(P.npiv[p], 0), P.cap[p]);
  const long long io = P.ioff[p];
  int nz = 0, nf = 0;
  for (int j = threadIdx.x; j < npiv; j += kWave) {
    const int t = P.ptype[io + j];
    const double d = P.dv[io + j];
    if (t == 1) {
      nz += d == 0.0;
      nf += !isfinite(d);
    } else if (t == 2 && j + 1 < npiv) {
      const double e = P.dv[io + j + 1], c = P.dv2[io + j];
      nz += (d * e - c * c) == 0.0;
      nf += !(isfinite(d) && isfinite(e) && isfinite(c));
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    nz += __shfl_xor(nz, o);
    nf += __shfl_xor(nf, o);
  }
  if (threadIdx.x == 0) {
    out[2 * p] = nz;
    out[2 * p + 1] = nf;
  }
}

// --------------------------------------------------------------------------
// Host side.
// --------------------------------------------------------------------------
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    HIPCHECK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    n = std::max<size_t>(count, 1);
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

enum KernelClass {
  KC_PREPARE, KC_ASSEMBLE, KC_FRONT_LDS, KC_FACTOR, KC_PANEL, KC_UPDATE,
  KC_FORWARD, KC_BACKWARD, KC_REFINE, KC_INVERT, KC_COUNT
};

struct Handle {
  kkt_ldl_options opt;
  // large-front panels choose pivots speculatively and verify them; the
  // environment variable KKT_LDL_PANEL_SPECULATION=0 selects the sequential
  // panel everywhere (same factors, bit for bit)
  int panel_spec = 1;
  // factor schedule: by default a node starts at the tick after its last
  // child's last panel; the environment variable KKT_LDL_LEVEL_SCHEDULE=1
  // selects the level schedule (every node of a tree level starts together
  // and the level lasts as long as its longest panel chain).  Same factors,
  // bit for bit.
  bool level_sched = false;
  // optional per-kernel-class device timing (events around every launch)
  bool profiling = false;
  std::vector<hipEvent_t> prof_ev;   // pairs
  std::vector<int> prof_cls;
  double prof_ms[KC_COUNT] = {0};
  int prof_n[KC_COUNT] = {0};
  int n = 0, nnz = 0;
  std::vector<int> rowp, cols;
  SymbolicResult S;
  int N = 0;
  // plan (host)
  std::vector<int> ns, cs, cap, mode, level, lvl_ptr, lvl_nodes;
  std::vector<long long> foff, ioff, woff, uoff, invoff;
  std::vector<int> invnom, fwdnom, am_ptr, am_src, am_row, am_col;
  std::vector<int> asm_node, asm_chunk;
  long long f_total = 0, i_total = 0, w_total = 0, u_total = 0;
  int max_cap = 0;
  int max_below = 0;  // largest cap - npiv over the large fronts (last factorization)
  // device
  DBuf<int> d_ns, d_cs, d_cap, d_vptr, d_cptr, d_chptr, d_ch, d_mode, d_n2v,
      d_crows, d_invnom, d_fwdnom, d_amptr, d_amsrc, d_amrow, d_amcol, d_asm_node,
      d_asm_chunk, d_rowp, d_cols;
  DBuf<long long> d_invoff, d_foff, d_ioff, d_woff, d_uoff;
  DBuf<int> d_din, d_m, d_npiv, d_ndel, d_doff, d_stat, d_need, d_cnt;
  DBuf<PanelState> d_ps;
  // per level (the solve's schedule): small fronts (one workgroup each),
  // large fronts, LDS fronts, and the largest panel count of a level
  std::vector<int> small_ptr, small_nodes, large_ptr, large_nodes, large_panels;
  std::vector<int> lds_ptr, lds_nodes;
  int lds_front = kLdsFront;  // largest front dimension kept in LDS
  DBuf<int> d_large_nodes;
  // The factorization's schedule, in ticks.  Tick t runs k_prepare, the LDS
  // fronts, the assembly and the small fronts of the nodes that start at t
  // (tk_grp: all of them; tk_lds, tk_small, tk_asm), then one panel launch
  // over the large fronts active at t (tk_act, with the launch's largest
  // slot and whether it holds a root front) and one update launch over
  // their trailing tiles (upd_ptr per tick; upd_base: the first column of
  // the tile grid, from the front's panel index at t).
  int ntick = 0;
  std::vector<int> tk_start;  // per node: start tick
  // side-stream event slot per tick (-1: no LDS fronts), and per tick the
  // latest slot its starting nodes wait for (-1: none)
  std::vector<int> tk_lds_ev, tk_wait;
  std::vector<int> tk_chain_end;
  std::vector<int> tk_prep;  // per tick: k_prepare launched (else prepared inline)
  std::vector<int> tk_asm_cap;  // per tick: largest slot assembled in HBM
  std::vector<int> tk_asm_lds;  // per tick: assembled by k_assemble_lds
  // LDS fronts take the right-looking kernels of kkt_front.hip (every mode
  // but Cholesky, which keeps k_front_lds)
  bool rl_fronts = true;
  // large root fronts keep Bunch-Kaufman pivoting throughout (mode 1)
  bool root_bk = false;
  int n_lds_ev = 0;
  std::vector<hipEvent_t> lds_ev;
  std::vector<int> tk_grp_ptr, tk_grp_nodes, tk_lds_ptr, tk_lds_nodes, tk_small_ptr,
      tk_small_nodes, tk_asm_ptr, tk_act_ptr, tk_act_nodes, tk_act_cap, tk_act_root;
  // per tick t, fronts upd_ptr[t] .. upd_ptr[t + 1] - 1 of the update launch:
  // node, first column of its tile grid, its first workgroup (exclusive
  // prefix of the tiles); tk_upd_tiles[t] workgroups in all
  std::vector<int> upd_ptr, upd_node, upd_pref, upd_base, tk_upd_tiles;
  DBuf<int> d_tk_grp_nodes, d_tk_lds_nodes, d_tk_small_nodes, d_tk_act_nodes, d_upd_node,
      d_upd_pref, d_upd_base;
  // solve: per level, the nodes solved by one workgroup each (sv), the
  // backward L21^T GEMV tiles of the large nodes (bg, 16 columns each) and
  // their panel steps (fs forward, bs backward: fs_lvl[l] is the first step
  // slot of level l, a slot's work items are [fs_ptr[slot], fs_ptr[slot+1]));
  // Y holds a large node's pivot vector
  std::vector<int> sv_ptr, sv_nodes, bg_ptr, bg_node, bg_tile;
  // sv nodes of a level: the first sv_tiny[l] have cap <= kTinyFront and are
  // solved by one wave each; LDS shape per (level, group): sv_lt, sv_lb
  std::vector<int> sv_tiny, sv_lt, sv_lb;
  // runs of levels holding one small node each and no large node (chains,
  // e.g. banded KKTs under AMD): the solve walks a run in one launch
  // (k_forward_chain, k_backward_chain).  sv_run_end[l] / sv_run_start[e]:
  // the run [l, e] (-1 elsewhere), its LDS shape sv_run_lt / sv_run_lb.
  // KKT_LDL_SOLVE_CHAINS=0 keeps one launch per level (same results, bit
  // for bit).
  std::vector<int> sv_run_end, sv_run_start, sv_run_lt, sv_run_lb;
  bool solve_chains = true;
  std::vector<int> fs_lvl, fs_ptr, fs_node, fs_tile, bs_lvl, bs_ptr, bs_node, bs_tile;
  DBuf<int> d_sv_nodes, d_bg_node, d_bg_tile, d_fs_node, d_fs_tile, d_bs_node, d_bs_tile;
  DBuf<double> d_Y;
  // panel inverses of the large nodes: one (node, panel) item per block
  std::vector<long long> pinvoff;
  std::vector<int> inv_node, inv_panel, inv2_node, inv2_pair;
  long long pinv_total = 0;
  std::vector<int> lvl_steps;  // per level: panel steps needed by the last factorization
  // per level: every large node has pivots (last factorization), so the
  // forward gather is done by the first step's workgroups (k_fwd_step<...,
  // true>) instead of its own launch; KKT_LDL_FUSE_GATHER=0 keeps the launch
  // (same results, bit for bit)
  std::vector<char> lvl_gfuse;
  bool fuse_gather = true;
  // KKT_LDL_MERGE_LEVELS=0 keeps a level's small fronts in launches of their
  // own (k_fwd_level / k_bwd_level otherwise; same results, bit for bit)
  bool merge_levels = true;
  // KKT_LDL_RL_INSTANCES=0: every LDS front launch takes the kernel
  // instance for the largest fronts (else the one for its front bound:
  // fewer registers, more workgroups per CU; the same results)
  bool rl_sized = true;
  // KKT_LDL_SIDE_STREAM=1: the LDS fronts of a tick that has other work run
  // on the handle's second stream beside it.  Off by default: with it, some
  // repeated FP32 factorizations of one matrix differed in their last bits
  // (an ordering hazard between the two streams whose cause is not found;
  // none in 180 factorizations on one stream)
  bool side_lds = false;
  DBuf<long long> d_pinvoff;
  DBuf<int> d_inv_node, d_inv_panel, d_inv2_node, d_inv2_pair;
  DBuf<double> d_pinv;
  int y_nrhs = 0;
  DBuf<double> d_F, d_W, d_U, d_vals, d_diag, d_x, d_r, d_b, d_xb, d_norm, d_dv, d_dv2;
  int esize = 8;  // bytes per front entry: 8 (FP64) or 4 (FP32 factor)
  DBuf<int> d_vars, d_pre, d_ptype;
  bool has_diag = false;
  bool vals_loaded = false;  // d_vals holds the values of the last kkt_ldl_factor
  bool factored = false;
  bool uploaded = false;
  int u_nrhs = 0;
  kkt_ldl_stats stats;
  // host copies of the factorization's per-node results, in one pinned
  // buffer (stat | npiv | ndel | cnt: 8 N ints) so that they arrive with one
  // synchronisation
  int* h_pin = nullptr;
  int *h_stat = nullptr, *h_npiv = nullptr, *h_ndel = nullptr, *h_cnt = nullptr;
  // pinned host copies of the refined solve's squared norms (right-hand
  // sides, residuals: kMaxNormRhs each), so that their reads stay
  // asynchronous
  double* h_norm = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  float last_factor_ms = 0.f, last_solve_ms = 0.f;
  // second stream (default priority) for the LDS fronts when side_lds is
  // set, which run beside the panel chain: ev_fork orders them after their
  // k_prepare, lds_ev[] (one per tick with LDS fronts) orders their parents
  // after them
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr;
};

// LDS bytes of k_front_lds for a front of dimension c.
static size_t lds_front_bytes(int c, size_t esize) {
  return sizeof(double) * ((size_t)c * kNB + 2 * (size_t)c) + esize * (size_t)c * c +
         3 * sizeof(int) * (size_t)c + 16;
}

static void build_plan(Handle& H, const std::vector<int>& slack) {
  const SymbolicResult& S = H.S;
  const int N = S.num_nodes;
  H.N = N;
  H.ns.resize(N);
  H.cs.resize(N);
  H.cap.resize(N);
  H.mode.resize(N);
  for (int k = 0; k < N; k++) {
    H.ns[k] = S.node_ptr[k + 1] - S.node_ptr[k];
    H.cs[k] = S.contrib_ptr[k + 1] - S.contrib_ptr[k];
    H.cap[k] = H.ns[k] + H.cs[k] + slack[k];
    bool root = S.node_parent[k] < 0;
    H.mode[k] = H.opt.solver_type == KKT_LDL_CHOLESKY ? 2 : (root ? 1 : 0);
    // a root front too large for LDS: threshold pivoting with delays first
    // (the tests the speculative panel accepts), Bunch-Kaufman on what it
    // delays (mode 3); KKT_LDL_ROOT_BK=1 keeps Bunch-Kaufman throughout
    if (H.mode[k] == 1 && !H.root_bk && H.cap[k] > H.lds_front) H.mode[k] = 3;
  }
  // levels (leaves = 0)
  H.level.assign(N, 0);
  int nlev = 0;
  for (int k = 0; k < N; k++) {
    for (int q = S.children_ptr[k]; q < S.children_ptr[k + 1]; q++)
      H.level[k] = std::max(H.level[k], H.level[S.children[q]] + 1);
    nlev = std::max(nlev, H.level[k] + 1);
  }
  H.lvl_ptr.assign(nlev + 1, 0);
  for (int k = 0; k < N; k++) H.lvl_ptr[H.level[k] + 1]++;
  for (int l = 0; l < nlev; l++) H.lvl_ptr[l + 1] += H.lvl_ptr[l];
  H.lvl_nodes.assign(N, 0);
  {
    std::vector<int> fill(H.lvl_ptr.begin(), H.lvl_ptr.end() - 1);
    for (int k = 0; k < N; k++) H.lvl_nodes[fill[H.level[k]]++] = k;
  }
  // arenas
  H.foff.resize(N);
  H.ioff.resize(N);
  H.woff.resize(N);
  H.uoff.resize(N);
  long long f = 0, iv = 0, u = 0;
  H.max_cap = 0;
  for (int k = 0; k < N; k++) {
    H.foff[k] = f;
    f += (long long)H.cap[k] * H.cap[k];
    f = (f + 31) & ~31LL;
    H.ioff[k] = iv;
    iv += H.cap[k];
    H.uoff[k] = u;
    u += H.cap[k];
    H.max_cap = std::max(H.max_cap, H.cap[k]);
  }
  H.f_total = f;
  H.i_total = iv;
  H.u_total = u;
  // Per level (the solve's schedule): LDS fronts sorted by cap, small
  // (Cholesky) fronts, large fronts and the level's largest panel count.
  // Large fronts get the multi-workgroup panel/update path.  A panel ends
  // when kw >= kNBLarge-1 or the candidates run out, so every panel but a
  // front's last takes at least kNBLarge-1 pivots: ceil(pivots / 15) panel
  // launches suffice, and the update tiles of a front's panel t may start
  // at the smaller bound t*(kNBLarge-2).
  auto is_large = [&](int k) { return H.cap[k] > H.lds_front && H.mode[k] != 2; };
  // (a mode-3 root may end its threshold phase with a short panel: one more)
  auto panels_of = [&](int k) {
    return (H.cap[k] - H.cs[k] + kNBLarge - 2) / (kNBLarge - 1) + (H.mode[k] == 3 ? 1 : 0);
  };
  H.small_ptr.assign(nlev + 1, 0);
  H.large_ptr.assign(nlev + 1, 0);
  H.lds_ptr.assign(nlev + 1, 0);
  H.small_nodes.clear();
  H.large_nodes.clear();
  H.lds_nodes.clear();
  H.large_panels.assign(nlev, 0);
  for (int l = 0; l < nlev; l++) {
    for (int q = H.lvl_ptr[l]; q < H.lvl_ptr[l + 1]; q++) {
      int k = H.lvl_nodes[q];
      if (H.cap[k] <= H.lds_front) {
        H.lds_nodes.push_back(k);
      } else if (is_large(k)) {
        H.large_nodes.push_back(k);
        H.large_panels[l] = std::max(H.large_panels[l], panels_of(k));
      } else {
        H.small_nodes.push_back(k);
      }
    }
    std::sort(H.lds_nodes.begin() + H.lds_ptr[l], H.lds_nodes.end(),
              [&](int a, int b) { return H.cap[a] < H.cap[b]; });
    H.lds_ptr[l + 1] = (int)H.lds_nodes.size();
    H.small_ptr[l + 1] = (int)H.small_nodes.size();
    H.large_ptr[l + 1] = (int)H.large_nodes.size();
  }
  // The factorization's schedule.  Node k starts at tick tk_start[k] and,
  // if large, runs one panel per tick until tick end[k]; an LDS or small
  // front is factored by one launch at its start tick.  Children are
  // numbered before their parents.
  std::vector<int>& start = H.tk_start;
  start.assign(N, 0);
  std::vector<int> end(N, 0);
  if (H.level_sched) {
    // every node of level l starts at the level's first tick; the level's
    // large fronts stay active for its largest panel count
    std::vector<int> t0(nlev + 1, 0);
    for (int l = 0; l < nlev; l++) t0[l + 1] = t0[l] + std::max(1, H.large_panels[l]);
    for (int k = 0; k < N; k++) {
      const int l = H.level[k];
      start[k] = t0[l];
      end[k] = is_large(k) ? t0[l] + H.large_panels[l] - 1 : t0[l];
    }
  } else {
    // a node starts at the tick after its last child ends; a large front
    // ends after its panels, an LDS or small front at its start tick (an
    // LDS front runs on the side stream; its parent's k_prepare waits for it
    // by an event).  Starting LDS-front parents later, by a cost model of
    // the side stream, was measured slower.
    for (int k = 0; k < N; k++) {
      int t = 0;
      for (int q = S.children_ptr[k]; q < S.children_ptr[k + 1]; q++)
        t = std::max(t, end[S.children[q]] + 1);
      start[k] = t;
      end[k] = is_large(k) ? t + panels_of(k) - 1 : t;
    }
  }
  int T = 0;
  for (int k = 0; k < N; k++) T = std::max(T, end[k] + 1);
  H.ntick = T;
  // nodes by start tick (node order inside a tick); LDS fronts sorted by cap
  auto bucket = [&](std::vector<int>& ptr, std::vector<int>& list, auto pick) {
    ptr.assign(T + 1, 0);
    for (int k = 0; k < N; k++)
      if (pick(k)) ptr[start[k] + 1]++;
    for (int t = 0; t < T; t++) ptr[t + 1] += ptr[t];
    list.assign(ptr[T], 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int k = 0; k < N; k++)
      if (pick(k)) list[fill[start[k]]++] = k;
  };
  bucket(H.tk_grp_ptr, H.tk_grp_nodes, [](int) { return true; });
  bucket(H.tk_lds_ptr, H.tk_lds_nodes, [&](int k) { return H.cap[k] <= H.lds_front; });
  bucket(H.tk_small_ptr, H.tk_small_nodes,
         [&](int k) { return H.cap[k] > H.lds_front && !is_large(k); });
  for (int t = 0; t < T; t++)
    std::sort(H.tk_lds_nodes.begin() + H.tk_lds_ptr[t], H.tk_lds_nodes.begin() + H.tk_lds_ptr[t + 1],
              [&](int a, int b) { return H.cap[a] < H.cap[b]; });
  // k_prepare runs at a tick only where a node needs it before its factor
  // kernel: Cholesky small fronts, LDS fronts of the blocked kernel, fronts
  // with more children than k_assemble's LDS cache.  Otherwise k_assemble
  // and the right-looking LDS kernels prepare their nodes themselves.
  H.tk_prep.assign(T, 0);
  for (int t = 0; t < T; t++) {
    int need = !H.rl_fronts || H.tk_small_ptr[t + 1] > H.tk_small_ptr[t];
    for (int q = H.tk_grp_ptr[t]; q < H.tk_grp_ptr[t + 1]; q++) {
      const int k = H.tk_grp_nodes[q];
      need |= S.children_ptr[k + 1] - S.children_ptr[k] > kChildCache;
    }
    H.tk_prep[t] = need;
  }
  // assembly work items of the fronts assembled in HBM, by start tick:
  // kAsmLdsCols columns per item where the tick's fronts are assembled
  // through LDS (k_assemble_lds), kAsmCols where k_assemble runs
  H.asm_node.clear();
  H.asm_chunk.clear();
  H.tk_asm_ptr.assign(T + 1, 0);
  H.tk_asm_cap.assign(T, 0);
  H.tk_asm_lds.assign(T, 0);
  for (int t = 0; t < T; t++) {
    for (int q = H.tk_grp_ptr[t]; q < H.tk_grp_ptr[t + 1]; q++) {
      const int k = H.tk_grp_nodes[q];
      if (H.cap[k] > H.lds_front) H.tk_asm_cap[t] = std::max(H.tk_asm_cap[t], H.cap[k]);
    }
    H.tk_asm_lds[t] = !H.tk_prep[t] && H.tk_asm_cap[t] <= kAsmLdsRows;
    const int width = H.tk_asm_lds[t] ? kAsmLdsCols : kAsmCols;
    for (int q = H.tk_grp_ptr[t]; q < H.tk_grp_ptr[t + 1]; q++) {
      const int k = H.tk_grp_nodes[q];
      if (H.cap[k] <= H.lds_front) continue;  // assembled inside the LDS front kernels
      const int nch = (H.cap[k] + width - 1) / width;
      // k_assemble_lds: the row chunks of each column chunk from the one
      // holding its first column (rows above a column are not assembled)
      const int nrc = H.tk_asm_lds[t] ? (H.cap[k] + kAsmRows - 1) / kAsmRows : 1;
      for (int c = 0; c < nch; c++) {
        const int rc0 = H.tk_asm_lds[t] ? c * width / kAsmRows : 0;
        for (int r = rc0; r < nrc; r++) {
          H.asm_node.push_back(k);
          H.asm_chunk.push_back(H.tk_asm_lds[t] ? c * kAsmRowChunks + r : c);
        }
      }
    }
    H.tk_asm_ptr[t + 1] = (int)H.asm_node.size();
  }
  // active large fronts per tick and their update tiles
  {
    std::vector<std::vector<int>> act(T);
    for (int k = 0; k < N; k++)
      if (is_large(k))
        for (int t = start[k]; t <= end[k]; t++) act[t].push_back(k);
    H.tk_act_ptr.assign(T + 1, 0);
    H.tk_act_nodes.clear();
    H.tk_act_cap.assign(T, 0);
    H.tk_act_root.assign(T, 0);
    H.upd_ptr.assign(T + 1, 0);
    H.upd_node.clear();
    H.upd_pref.clear();
    H.upd_base.clear();
    H.tk_upd_tiles.assign(T, 0);
    for (int t = 0; t < T; t++) {
      int tiles = 0;
      for (int k : act[t]) {
        H.tk_act_nodes.push_back(k);
        H.tk_act_cap[t] = std::max(H.tk_act_cap[t], H.cap[k]);
        H.tk_act_root[t] |= H.mode[k] == 1 || H.mode[k] == 3;
        // first column of the tile grid: a lower bound of the panel's start
        // (one short panel allowed for a mode-3 root)
        const int tp = t - start[k] - (H.mode[k] == 3 ? 1 : 0);
        const int base = std::max(tp, 0) * (kNBLarge - 2);
        const int nt = std::max((H.cap[k] - base + 63) / 64, 0);
        H.upd_node.push_back(k);
        H.upd_base.push_back(base);
        H.upd_pref.push_back(tiles);
        tiles += nt * (nt + 1) / 2;
      }
      H.tk_act_ptr[t + 1] = (int)H.tk_act_nodes.size();
      H.upd_ptr[t + 1] = (int)H.upd_node.size();
      H.tk_upd_tiles[t] = tiles;
    }
  }
  // Side-stream events: one per tick that launches LDS fronts, recorded
  // after them, in tick order; the nodes starting at tick t wait (before
  // their k_prepare, which reads the children's results) for the event of
  // the latest tick holding one of their LDS children.  The side stream runs
  // in order, so that event covers every earlier LDS front as well.
  // A tick whose LDS fronts are its only work runs them on the main stream
  // instead: a hop between streams costs more than it hides there (measured
  // about 18 us per dependent hop on a chain of LDS fronts).
  H.tk_lds_ev.assign(T, -1);
  H.n_lds_ev = 0;
  for (int t = 0; t < T; t++) {
    const bool lds = H.tk_lds_ptr[t + 1] > H.tk_lds_ptr[t];
    const bool other = H.tk_small_ptr[t + 1] > H.tk_small_ptr[t] ||
                       H.tk_lds_ptr[t + 1] - H.tk_lds_ptr[t] < H.tk_grp_ptr[t + 1] - H.tk_grp_ptr[t] ||
                       H.tk_act_ptr[t + 1] > H.tk_act_ptr[t];
    if (lds && other) H.tk_lds_ev[t] = H.n_lds_ev++;
  }
  H.tk_wait.assign(T, -1);
  for (int t = 0; t < T; t++)
    for (int q = H.tk_grp_ptr[t]; q < H.tk_grp_ptr[t + 1]; q++) {
      const int k = H.tk_grp_nodes[q];
      for (int c = S.children_ptr[k]; c < S.children_ptr[k + 1]; c++) {
        const int ch = S.children[c];
        if (H.cap[ch] <= H.lds_front)  // -1: that tick ran its LDS fronts on this stream
          H.tk_wait[t] = std::max(H.tk_wait[t], H.tk_lds_ev[start[ch]]);
      }
    }
  // Chains: runs of consecutive ticks whose only work is one LDS front
  // (threshold or Bunch-Kaufman mode) are factored by one workgroup in one
  // launch (k_front_chain); tk_chain_end[t] is the run's last tick at its
  // first tick t, else -1.  Every front of a run has its children in
  // earlier ticks, so the list order is an elimination order.
  H.tk_chain_end.assign(T, -1);
  if (H.rl_fronts) {
    auto chain_tick = [&](int t) {
      if (H.tk_grp_ptr[t + 1] - H.tk_grp_ptr[t] != 1) return false;
      if (H.tk_act_ptr[t + 1] > H.tk_act_ptr[t]) return false;
      const int k = H.tk_grp_nodes[H.tk_grp_ptr[t]];
      return H.cap[k] <= H.lds_front && H.mode[k] != 2;
    };
    for (int t = 0; t < T;) {
      if (!chain_tick(t)) {
        t++;
        continue;
      }
      int e = t;
      while (e + 1 < T && chain_tick(e + 1)) e++;
      H.tk_chain_end[t] = e;
      t = e + 1;
    }
  }
  // panel scratch W (cap x kNB per front factored in HBM), used only by the
  // factor kernels of the front's own ticks, all on the main stream: a
  // region is reused by a front that starts after the region's last front
  // ended (best fit by size, in start order)
  {
    struct Region {
      long long off, size;
      int until;
    };
    std::vector<Region> regions;
    long long w = 0;
    for (int q = 0; q < (int)H.tk_grp_nodes.size(); q++) {
      const int k = H.tk_grp_nodes[q];
      H.woff[k] = 0;
      if (H.cap[k] <= H.lds_front) continue;
      const long long need = (long long)H.cap[k] * kNB;
      int best = -1;
      for (int r = 0; r < (int)regions.size(); r++)
        if (regions[r].until < start[k] && regions[r].size >= need &&
            (best < 0 || regions[r].size < regions[best].size))
          best = r;
      if (best < 0) {
        regions.push_back({w, need, 0});
        w += need;
        best = (int)regions.size() - 1;
      }
      regions[best].until = end[k];
      H.woff[k] = regions[best].off;
    }
    H.w_total = std::max(w, 1LL);
  }
  H.sv_ptr.assign(nlev + 1, 0);
  H.sv_tiny.assign(nlev, 0);
  H.sv_lt.assign(2 * (size_t)nlev, 0);
  H.sv_lb.assign(2 * (size_t)nlev, 0);
  H.bg_ptr.assign(nlev + 1, 0);
  H.sv_nodes.clear();
  H.bg_node.clear();
  H.bg_tile.clear();
  H.fs_lvl.assign(nlev + 1, 0);
  H.bs_lvl.assign(nlev + 1, 0);
  H.fs_ptr.assign(1, 0);
  H.bs_ptr.assign(1, 0);
  H.fs_node.clear();
  H.fs_tile.clear();
  H.bs_node.clear();
  H.bs_tile.clear();
  H.pinvoff.assign(N, 0);
  H.inv_node.clear();
  H.inv_panel.clear();
  H.inv2_node.clear();
  H.inv2_pair.clear();
  long long pv = 0;
  // panels of a large node: its pivots never exceed cap - cs; a solve step
  // covers two panels
  auto npanels = [&](int k) { return (H.cap[k] - H.cs[k] + kPB - 1) / kPB; };
  auto nsteps = [&](int k) { return (npanels(k) + 1) / 2; };
  for (int l = 0; l < nlev; l++) {
    for (int q = H.lds_ptr[l]; q < H.lds_ptr[l + 1]; q++) H.sv_nodes.push_back(H.lds_nodes[q]);
    for (int q = H.small_ptr[l]; q < H.small_ptr[l + 1]; q++)
      H.sv_nodes.push_back(H.small_nodes[q]);
    int maxp = 0;
    for (int q = H.large_ptr[l]; q < H.large_ptr[l + 1]; q++) {
      const int k = H.large_nodes[q];
      for (int t = 0; t < (H.cap[k] + kGemvCols - 1) / kGemvCols; t++) {
        H.bg_node.push_back(k);
        H.bg_tile.push_back(t);
      }
      H.pinvoff[k] = pv;
      pv += nsteps(k) * kPinvPair;
      for (int t = 0; t < npanels(k); t++) {
        H.inv_node.push_back(k);
        H.inv_panel.push_back(t);
      }
      for (int t = 0; 2 * t + 1 < npanels(k); t++) {
        H.inv2_node.push_back(k);
        H.inv2_pair.push_back(t);
      }
      maxp = std::max(maxp, nsteps(k));
    }
    H.sv_ptr[l + 1] = (int)H.sv_nodes.size();
    H.bg_ptr[l + 1] = (int)H.bg_node.size();
    // tiny fronts (cap <= kTinyFront) lead the level's list (LDS fronts are
    // sorted by cap); each group gets LDS for its largest front vector and
    // its largest diagonal block
    for (int q = H.sv_ptr[l]; q < H.sv_ptr[l + 1]; q++) {
      const int k = H.sv_nodes[q];
      const bool tiny = H.cap[k] <= kTinyFront && q - H.sv_ptr[l] == H.sv_tiny[l];
      if (tiny) H.sv_tiny[l]++;
      const int g = 2 * l + (tiny ? 0 : 1);
      H.sv_lt[g] = std::max(H.sv_lt[g], (H.cap[k] + 7) & ~7);
      H.sv_lb[g] = std::max(H.sv_lb[g], std::min(kSB, H.cap[k] - H.cs[k]));
    }
    // forward step t: rows [b1, m) with b1 > t*kPS; tile 0 always exists
    H.fs_lvl[l] = (int)H.fs_ptr.size() - 1;
    for (int t = 0; t < maxp; t++) {
      for (int q = H.large_ptr[l]; q < H.large_ptr[l + 1]; q++) {
        const int k = H.large_nodes[q];
        if (t >= nsteps(k)) continue;
        const int nt = std::max(1, (H.cap[k] - t * kPS + kStepRows - 1) / kStepRows);
        for (int u = 0; u < nt; u++) {
          H.fs_node.push_back(k);
          H.fs_tile.push_back(u);
        }
      }
      H.fs_ptr.push_back((int)H.fs_node.size());
    }
    // backward step t: step (steps - 1 - t), columns before it
    H.bs_lvl[l] = (int)H.bs_ptr.size() - 1;
    for (int t = 0; t < maxp; t++) {
      for (int q = H.large_ptr[l]; q < H.large_ptr[l + 1]; q++) {
        const int k = H.large_nodes[q];
        if (t >= nsteps(k)) continue;
        const int cols = (nsteps(k) - 1 - t) * kPS;
        const int nt = std::max(1, (cols + kStepCols - 1) / kStepCols);
        for (int u = 0; u < nt; u++) {
          H.bs_node.push_back(k);
          H.bs_tile.push_back(u);
        }
      }
      H.bs_ptr.push_back((int)H.bs_node.size());
    }
  }
  H.fs_lvl[nlev] = (int)H.fs_ptr.size() - 1;
  H.bs_lvl[nlev] = (int)H.bs_ptr.size() - 1;
  H.sv_run_end.assign(nlev, -1);
  H.sv_run_start.assign(nlev, -1);
  H.sv_run_lt.assign(nlev, 0);
  H.sv_run_lb.assign(nlev, 0);
  auto single = [&](int l) {
    return H.sv_ptr[l + 1] - H.sv_ptr[l] == 1 && H.large_ptr[l + 1] == H.large_ptr[l];
  };
  for (int l = 0; l < nlev;) {
    if (!single(l)) {
      l++;
      continue;
    }
    int e = l;
    while (e + 1 < nlev && single(e + 1)) e++;
    if (e > l) {
      H.sv_run_end[l] = e;
      H.sv_run_start[e] = l;
      int lt = 0, lb = 1;
      for (int q = H.sv_ptr[l]; q < H.sv_ptr[e + 1]; q++) {
        const int k = H.sv_nodes[q];
        lt = std::max(lt, (H.cap[k] + 7) & ~7);
        lb = std::max(lb, std::min(kSB, H.cap[k] - H.cs[k]));
      }
      H.sv_run_lt[l] = lt;
      H.sv_run_lb[l] = lb;
    }
    l = e + 1;
  }
  H.pinv_total = pv;
}

// Static maps that do not depend on the slack.
static void build_maps(Handle& H) {
  const SymbolicResult& S = H.S;
  const int N = S.num_nodes, n = H.n;
  std::vector<int> node_of_var(n), pos_in_node(n);
  for (int k = 0; k < N; k++)
    for (int q = S.node_ptr[k]; q < S.node_ptr[k + 1]; q++) {
      node_of_var[S.node_to_var[q]] = k;
      pos_in_node[S.node_to_var[q]] = q - S.node_ptr[k];
    }
  // inverse nominal maps: for child c of parent p, for every nominal
  // position u of p (own columns then contribution rows), the index of
  // that variable in c's contribution rows, or -1.
  H.invoff.assign(N, -1);
  H.invnom.clear();
  std::vector<int> mark(n, -1);
  for (int c = 0; c < N; c++) {
    int p = S.node_parent[c];
    if (p < 0) continue;
    for (int q = S.contrib_ptr[c]; q < S.contrib_ptr[c + 1]; q++)
      mark[S.contrib_rows[q]] = q - S.contrib_ptr[c];
    H.invoff[c] = (long long)H.invnom.size();
    for (int q = S.node_ptr[p]; q < S.node_ptr[p + 1]; q++)
      H.invnom.push_back(mark[S.node_to_var[q]]);
    for (int q = S.contrib_ptr[p]; q < S.contrib_ptr[p + 1]; q++)
      H.invnom.push_back(mark[S.contrib_rows[q]]);
    for (int q = S.contrib_ptr[c]; q < S.contrib_ptr[c + 1]; q++)
      mark[S.contrib_rows[q]] = -1;
  }
  // forward maps: for every contribution row of a child, the parent's
  // nominal position of that variable (own columns, then contribution rows)
  H.fwdnom.assign((size_t)S.contrib_ptr[N] + 1, -1);
  for (int c = 0; c < N; c++) {
    int p = S.node_parent[c];
    if (p < 0) continue;
    const int nsp = S.node_ptr[p + 1] - S.node_ptr[p];
    for (int q = S.node_ptr[p]; q < S.node_ptr[p + 1]; q++) mark[S.node_to_var[q]] = q - S.node_ptr[p];
    for (int q = S.contrib_ptr[p]; q < S.contrib_ptr[p + 1]; q++)
      mark[S.contrib_rows[q]] = nsp + (q - S.contrib_ptr[p]);
    for (int q = S.contrib_ptr[c]; q < S.contrib_ptr[c + 1]; q++) {
      H.fwdnom[q] = mark[S.contrib_rows[q]];
      if (H.fwdnom[q] < 0) throw Error(-2, "contribution row outside the parent's structure");
    }
    for (int q = S.node_ptr[p]; q < S.node_ptr[p + 1]; q++) mark[S.node_to_var[q]] = -1;
    for (int q = S.contrib_ptr[p]; q < S.contrib_ptr[p + 1]; q++) mark[S.contrib_rows[q]] = -1;
  }
  // original entries per own column: entries (i, var) with permuted index
  // of i >= permuted index of var, placed at the front's nominal row.
  H.am_ptr.assign(n + 1, 0);
  H.am_src.clear();
  H.am_row.clear();
  H.am_col.clear();
  std::vector<int> nominal(n, -1);
  for (int k = 0; k < N; k++) {
    const int ns = S.node_ptr[k + 1] - S.node_ptr[k];
    for (int q = S.node_ptr[k]; q < S.node_ptr[k + 1]; q++)
      nominal[S.node_to_var[q]] = q - S.node_ptr[k];
    for (int q = S.contrib_ptr[k]; q < S.contrib_ptr[k + 1]; q++)
      nominal[S.contrib_rows[q]] = ns + (q - S.contrib_ptr[k]);
    for (int q = S.node_ptr[k]; q < S.node_ptr[k + 1]; q++) {
      const int v = S.node_to_var[q];
      const int pv = S.permuted ? S.iperm[v] : v;
      for (int e = H.rowp[v]; e < H.rowp[v + 1]; e++) {
        const int i = H.cols[e];
        const int pi = S.permuted ? S.iperm[i] : i;
        if (pi >= pv && nominal[i] >= 0) {
          H.am_src.push_back(e);
          H.am_row.push_back(nominal[i]);
          H.am_col.push_back(q - S.node_ptr[k]);
        }
      }
      H.am_ptr[q + 1] = (int)H.am_src.size();
    }
    for (int q = S.node_ptr[k]; q < S.node_ptr[k + 1]; q++) nominal[S.node_to_var[q]] = -1;
    for (int q = S.contrib_ptr[k]; q < S.contrib_ptr[k + 1]; q++)
      nominal[S.contrib_rows[q]] = -1;
  }
}

static void upload_static(Handle& H) {
  const SymbolicResult& S = H.S;
  H.d_ns.upload(H.ns);
  H.d_cs.upload(H.cs);
  H.d_cap.upload(H.cap);
  H.d_mode.upload(H.mode);
  H.d_vptr.upload(S.node_ptr);
  H.d_cptr.upload(S.contrib_ptr);
  H.d_chptr.upload(S.children_ptr);
  H.d_ch.upload(S.children);
  H.d_n2v.upload(S.node_to_var);
  H.d_crows.upload(S.contrib_rows);
  H.d_invnom.upload(H.invnom);
  H.d_fwdnom.upload(H.fwdnom);
  H.d_invoff.upload(H.invoff);
  H.d_amptr.upload(H.am_ptr);
  H.d_amsrc.upload(H.am_src);
  H.d_amrow.upload(H.am_row);
  H.d_amcol.upload(H.am_col);
  H.d_foff.upload(H.foff);
  H.d_ioff.upload(H.ioff);
  H.d_woff.upload(H.woff);
  H.d_uoff.upload(H.uoff);
  H.d_tk_grp_nodes.upload(H.tk_grp_nodes);
  H.d_asm_node.upload(H.asm_node);
  H.d_asm_chunk.upload(H.asm_chunk);
  H.d_rowp.upload(H.rowp);
  H.d_cols.upload(H.cols);
  const int N = H.N;
  H.d_din.alloc(N);
  H.d_m.alloc(N);
  H.d_npiv.alloc(N);
  H.d_ndel.alloc(N);
  H.d_doff.alloc(N);
  H.d_stat.alloc(N);
  H.d_need.alloc(N);
  H.d_cnt.alloc(5 * (size_t)N);  // [3N] inertia counters, [2N] zero / non-finite pivots
  H.d_ps.alloc(N);
  H.d_tk_small_nodes.upload(H.tk_small_nodes);
  H.d_tk_lds_nodes.upload(H.tk_lds_nodes);
  H.d_tk_act_nodes.upload(H.tk_act_nodes);
  H.d_large_nodes.upload(H.large_nodes);
  H.d_upd_node.upload(H.upd_node);
  H.d_upd_pref.upload(H.upd_pref);
  H.d_upd_base.upload(H.upd_base);
  H.d_sv_nodes.upload(H.sv_nodes);
  H.d_bg_node.upload(H.bg_node);
  H.d_bg_tile.upload(H.bg_tile);
  H.d_fs_node.upload(H.fs_node);
  H.d_fs_tile.upload(H.fs_tile);
  H.d_bs_node.upload(H.bs_node);
  H.d_bs_tile.upload(H.bs_tile);
  H.d_pinvoff.upload(H.pinvoff);
  H.d_inv_node.upload(H.inv_node);
  H.d_inv_panel.upload(H.inv_panel);
  H.d_inv2_node.upload(H.inv2_node);
  H.d_inv2_pair.upload(H.inv2_pair);
  H.d_pinv.alloc((size_t)H.pinv_total);
  H.y_nrhs = 0;
  // 64 bytes of padding: the trailing update's 16-byte tile reads can pass
  // the last front's last entry by up to 12 bytes
  H.d_F.alloc(((size_t)H.f_total * H.esize + 7) / 8 + 8);
  H.d_dv.alloc((size_t)H.i_total);
  H.d_dv2.alloc((size_t)H.i_total);
  H.d_vars.alloc((size_t)H.i_total);
  H.d_pre.alloc((size_t)H.i_total);
  H.d_ptype.alloc((size_t)H.i_total);
  H.d_W.alloc((size_t)H.w_total);
  H.u_nrhs = 0;
}

static DevPlan dev_plan(Handle& H) {
  DevPlan P;
  P.ns = H.d_ns.p;
  P.cs = H.d_cs.p;
  P.cap = H.d_cap.p;
  P.vptr = H.d_vptr.p;
  P.cptr = H.d_cptr.p;
  P.chptr = H.d_chptr.p;
  P.ch = H.d_ch.p;
  P.invoff = H.d_invoff.p;
  P.foff = H.d_foff.p;
  P.ioff = H.d_ioff.p;
  P.woff = H.d_woff.p;
  P.uoff = H.d_uoff.p;
  P.mode = H.d_mode.p;
  P.node_to_var = H.d_n2v.p;
  P.contrib_rows = H.d_crows.p;
  P.invnom = H.d_invnom.p;
  P.fwdnom = H.d_fwdnom.p;
  P.am_ptr = H.d_amptr.p;
  P.am_src = H.d_amsrc.p;
  P.am_row = H.d_amrow.p;
  P.am_col = H.d_amcol.p;
  P.din = H.d_din.p;
  P.m = H.d_m.p;
  P.npiv = H.d_npiv.p;
  P.ndel = H.d_ndel.p;
  P.doff = H.d_doff.p;
  P.stat = H.d_stat.p;
  P.need = H.d_need.p;
  P.cnt = H.d_cnt.p;
  P.ps = H.d_ps.p;
  P.F = (void*)H.d_F.p;
  P.dv = H.d_dv.p;
  P.dv2 = H.d_dv2.p;
  P.vars = H.d_vars.p;
  P.pre = H.d_pre.p;
  P.ptype = H.d_ptype.p;
  P.W = H.d_W.p;
  P.U = H.d_U.p;
  P.pinv = H.d_pinv.p;
  P.pinvoff = H.d_pinvoff.p;
  P.vals = H.d_vals.p;
  P.diag = H.has_diag ? H.d_diag.p : nullptr;
  P.ustab = H.opt.ustab;
  P.ptol = H.opt.pivot_tol;
  P.spec = H.panel_spec;
  return P;
}

static std::vector<int> initial_slack(const Handle& H) {
  const SymbolicResult& S = H.S;
  const int N = S.num_nodes;
  std::vector<int> slack(N, 0);
  if (H.opt.solver_type == KKT_LDL_CHOLESKY) return slack;
  const double g = std::max(1.0, H.opt.delay_growth);
  for (int k = 0; k < N; k++) {
    long long below = 0;
    for (int q = S.children_ptr[k]; q < S.children_ptr[k + 1]; q++) {
      int c = S.children[q];
      below += (S.node_ptr[c + 1] - S.node_ptr[c]) + slack[c];
    }
    int nominal = (S.node_ptr[k + 1] - S.node_ptr[k]) + (S.contrib_ptr[k + 1] - S.contrib_ptr[k]);
    long long want = (long long)std::ceil((g - 1.0) * 0.0625 * nominal) + 4;
    slack[k] = (int)std::min<long long>(below, want);
  }
  return slack;
}

// An event pair per launch (or group of launches) when profiling is on.
// By default the pair brackets the launches on the stream, which includes
// their dispatch.  With kernel = true the pair is handed to one
// hipExtLaunchKernelGGL (start(), stop()), which stamps the kernel's own
// start and end, as a kernel trace measures it.
struct ProfScope {
  Handle& H;
  hipStream_t st;
  bool on, kernel;
  size_t slot = 0;
  ProfScope(Handle& h, hipStream_t s, int cls, bool kernel_only = false)
      : H(h), st(s), on(h.profiling), kernel(kernel_only) {
    if (!on) return;
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    H.prof_ev.push_back(a);
    H.prof_ev.push_back(b);
    H.prof_cls.push_back(cls);
    slot = H.prof_ev.size() - 2;
    if (!kernel) HIPCHECK(hipEventRecord(a, st));
  }
  hipEvent_t start() const { return on ? H.prof_ev[slot] : nullptr; }
  hipEvent_t stop() const { return on ? H.prof_ev[slot + 1] : nullptr; }
  ~ProfScope() {
    if (on && !kernel) (void)hipEventRecord(H.prof_ev[slot + 1], st);
  }
};

static void prof_reset(Handle& H) {
  for (auto e : H.prof_ev) (void)hipEventDestroy(e);
  H.prof_ev.clear();
  H.prof_cls.clear();
  for (int c = 0; c < KC_COUNT; c++) {
    H.prof_ms[c] = 0.0;
    H.prof_n[c] = 0;
  }
}

static void prof_collect(Handle& H) {
  if (!H.profiling) return;
  for (size_t q = 0; q < H.prof_cls.size(); q++) {
    float ms = 0.f;
    HIPCHECK(hipEventSynchronize(H.prof_ev[2 * q + 1]));
    HIPCHECK(hipEventElapsedTime(&ms, H.prof_ev[2 * q], H.prof_ev[2 * q + 1]));
    H.prof_ms[H.prof_cls[q]] += ms;
    H.prof_n[H.prof_cls[q]]++;
  }
  for (auto e : H.prof_ev) (void)hipEventDestroy(e);
  H.prof_ev.clear();
  H.prof_cls.clear();
}

template <class T>
static void run_factor_levels_t(Handle& H, hipStream_t st) {
  DevPlan P = dev_plan(H);
  while ((int)H.lds_ev.size() < H.n_lds_ev) {
    hipEvent_t e;
    HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    H.lds_ev.push_back(e);
  }
  int waited = -1;  // latest side-stream event the main stream has waited for
  for (int t = 0; t < H.ntick; t++) {
    if (H.tk_chain_end[t] >= 0) {
      // a chain of single-front ticks: one workgroup, one launch
      const int e = H.tk_chain_end[t];
      int wait = -1, cap = 0;
      for (int u = t; u <= e; u++) {
        wait = std::max(wait, H.tk_wait[u]);
        cap = std::max(cap, H.cap[H.tk_grp_nodes[H.tk_grp_ptr[u]]]);
      }
      if (wait > waited) {
        waited = wait;
        HIPCHECK(hipStreamWaitEvent(st, H.lds_ev[waited], 0));
      }
      {
        ProfScope ps_(H, st, KC_FRONT_LDS);
        launch_front_rl((const T*)nullptr, cap, e - t + 1, true, true, st, P,
                        H.d_tk_grp_nodes.p + H.tk_grp_ptr[t]);
      }
      t = e;
      continue;
    }
    const int n0 = H.tk_grp_ptr[t], n1 = H.tk_grp_ptr[t + 1];
    const int d0 = H.tk_lds_ptr[t], d1 = H.tk_lds_ptr[t + 1];
    const int a0 = H.tk_asm_ptr[t], a1 = H.tk_asm_ptr[t + 1];
    const int s0 = H.tk_small_ptr[t], s1 = H.tk_small_ptr[t + 1];
    const int g0 = H.tk_act_ptr[t], g1 = H.tk_act_ptr[t + 1];
    if (n1 > n0) {
      // the nodes starting now: their large children ended at earlier ticks
      // on this stream; their LDS children ran on the side stream
      if (H.tk_wait[t] > waited) {
        waited = H.tk_wait[t];
        HIPCHECK(hipStreamWaitEvent(st, H.lds_ev[waited], 0));
      }
      if (H.tk_prep[t]) {
        ProfScope ps_(H, st, KC_PREPARE);
        hipLaunchKernelGGL(k_prepare, dim3((n1 - n0 + kTPB - 1) / kTPB), dim3(kTPB), 0, st,
                           P, H.d_tk_grp_nodes.p + n0, n1 - n0);
      }
      // The LDS fronts (assembled in LDS) depend only on k_prepare and on
      // their children; with side_lds they run on the side stream beside the
      // tick's other work (else on this stream), in size buckets, so that
      // small fronts get small LDS allocations and several workgroups per CU.
      if (d1 > d0) {
        const int ev = H.side_lds ? H.tk_lds_ev[t] : -1;
        hipStream_t ss = ev >= 0 ? H.side : st;
        if (ev >= 0) {
          HIPCHECK(hipEventRecord(H.ev_fork, st));
          HIPCHECK(hipStreamWaitEvent(H.side, H.ev_fork, 0));
        }
        for (int q0 = d0; q0 < d1;) {
          const int c0 = H.cap[H.tk_lds_nodes[q0]];
          int bucket = 32;
          while (bucket < c0) bucket += 16;
          int q1 = q0;
          while (q1 < d1 && H.cap[H.tk_lds_nodes[q1]] <= bucket) q1++;
          ProfScope ps_(H, ss, KC_FRONT_LDS);
          if (H.rl_fronts) {
            launch_front_rl((const T*)nullptr, std::min(bucket, H.lds_front), q1 - q0, false,
                            H.rl_sized, ss,
                            P, H.d_tk_lds_nodes.p + q0);
          } else {
            const size_t bytes = lds_front_bytes(bucket, sizeof(T));
            hipLaunchKernelGGL((k_front_lds<kTPB, T>), dim3(q1 - q0), dim3(kTPB), bytes, ss, P,
                               H.d_tk_lds_nodes.p + q0);
          }
          q0 = q1;
        }
        if (ev >= 0) HIPCHECK(hipEventRecord(H.lds_ev[ev], H.side));
      }
      if (a1 > a0) {
        ProfScope ps_(H, st, KC_ASSEMBLE);
        if (H.tk_asm_lds[t])
          hipLaunchKernelGGL(k_assemble_lds<T>, dim3(a1 - a0), dim3(kTPB), 0, st, P,
                             H.d_asm_node.p + a0, H.d_asm_chunk.p + a0);
        else
          hipLaunchKernelGGL(k_assemble<T>, dim3(a1 - a0), dim3(kTPB), 0, st, P,
                             H.d_asm_node.p + a0, H.d_asm_chunk.p + a0, H.tk_prep[t] ? 0 : 1);
      }
      if (s1 > s0) {
        ProfScope ps_(H, st, KC_FACTOR);
        hipLaunchKernelGGL(k_factor<T>, dim3(s1 - s0), dim3(kTPB), 0, st, P,
                           H.d_tk_small_nodes.p + s0);
      }
    }
    if (g1 > g0) {
      // one panel of every large front active at this tick, then the
      // trailing update of those panels
      {
        ProfScope ps_(H, st, KC_PANEL, true);
        launch_large_panel((const T*)nullptr, H.tk_act_root[t] != 0, H.tk_act_cap[t],
                           dim3(g1 - g0), st, ps_.start(), ps_.stop(), P,
                           H.d_tk_act_nodes.p + g0);
      }
      const int u0 = H.upd_ptr[t], u1 = H.upd_ptr[t + 1];
      if (H.tk_upd_tiles[t] > 0) {
        ProfScope ps_(H, st, KC_UPDATE, true);
        hipExtLaunchKernelGGL(k_update<T>, dim3(H.tk_upd_tiles[t]), dim3(kTPB), 0, st,
                              ps_.start(), ps_.stop(), 0, P, H.d_upd_node.p + u0,
                              H.d_upd_base.p + u0, H.d_upd_pref.p + u0, u1 - u0);
      }
    }
  }
  // every LDS front is done before the factorization's last kernels
  if (H.n_lds_ev > 0 && waited < H.n_lds_ev - 1)
    HIPCHECK(hipStreamWaitEvent(st, H.lds_ev[H.n_lds_ev - 1], 0));
  if (!H.inv_node.empty()) {
    ProfScope ps_(H, st, KC_INVERT);
    hipLaunchKernelGGL(k_panel_inv<T>, dim3((unsigned)H.inv_node.size()), dim3(kWave), 0, st, P,
                       H.d_inv_node.p, H.d_inv_panel.p);
    if (!H.inv2_node.empty())
      hipLaunchKernelGGL(k_panel_inv2<T>, dim3((unsigned)H.inv2_node.size()), dim3(kWave), 0, st,
                         P, H.d_inv2_node.p, H.d_inv2_pair.p);
  }
  HIPCHECK(hipGetLastError());
}

static void run_factor_levels(Handle& H, hipStream_t st) {
  if (H.esize == 4) run_factor_levels_t<float>(H, st);
  else run_factor_levels_t<double>(H, st);
}

// Device memory is claimed on the first factorization, so that the
// symbolic analysis (create, get_perm, get_supernodes) needs no GPU.
static void ensure_device(Handle& H) {
  if (H.uploaded) return;
  // Fit the LDS front bound to this device (m^2 + 64 m doubles + ~4 KB).
  int dev = 0, lim = 0;
  HIPCHECK(hipGetDevice(&dev));
  HIPCHECK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  int mlds = kLdsFront;
  while (mlds > 8 && (double)lds_front_bytes(mlds, 8) + 4096.0 > lim) mlds--;
  if (mlds != H.lds_front) {
    H.lds_front = mlds;
    std::vector<int> slack(H.N);
    for (int k = 0; k < H.N; k++) slack[k] = H.cap[k] - H.ns[k] - H.cs[k];
    build_plan(H, slack);
  }
  upload_static(H);
  HIPCHECK(hipEventCreate(&H.ev0));
  HIPCHECK(hipEventCreate(&H.ev1));
  HIPCHECK(hipEventCreate(&H.ev2));
  HIPCHECK(hipEventCreate(&H.ev3));
  HIPCHECK(hipStreamCreateWithFlags(&H.side, hipStreamNonBlocking));
  HIPCHECK(hipEventCreateWithFlags(&H.ev_fork, hipEventDisableTiming));
  HIPCHECK(hipHostMalloc((void**)&H.h_pin, sizeof(int) * 8 * (size_t)std::max(H.N, 1)));
  H.h_stat = H.h_pin;
  H.h_npiv = H.h_pin + H.N;
  H.h_ndel = H.h_pin + 2 * (size_t)H.N;
  H.h_cnt = H.h_pin + 3 * (size_t)H.N;
  HIPCHECK(hipHostMalloc((void**)&H.h_norm, sizeof(double) * 2 * kMaxNormRhs));
  H.uploaded = true;
}

static void ensure_u(Handle& H, int nrhs) {
  if (H.u_nrhs >= nrhs) return;
  H.d_U.alloc((size_t)H.u_total * nrhs);
  H.u_nrhs = nrhs;
}

// vals == nullptr: refactorization with the values of the last call, which
// stay on the device (only the diagonal is new).
static int factor_impl(Handle& H, const double* vals, const double* diag, int mem,
                       hipStream_t st) {
  auto t0 = std::chrono::steady_clock::now();
  if (mem != KKT_LDL_HOST && mem != KKT_LDL_DEVICE) throw Error(-1, "unknown memory space");
  if (!vals && !H.vals_loaded) throw Error(-3, "refactor called before factor");
  ensure_device(H);
  // the large-front panel (vars/pre) and the solve's front gather and
  // L21^T products stage a front's rows in LDS (two int or one double per row)
  // (fronts of any size: the kernels whose LDS staging grows with the front
  // have global-memory forms for the fronts beyond LDS)
  const hipMemcpyKind kind = mem == KKT_LDL_DEVICE ? hipMemcpyDeviceToDevice
                                                   : hipMemcpyHostToDevice;
  if (vals) {
    H.d_vals.alloc(H.nnz);
    HIPCHECK(hipMemcpyAsync(H.d_vals.p, vals, sizeof(double) * H.nnz, kind, st));
    H.vals_loaded = true;
  }
  H.has_diag = diag != nullptr;
  if (diag) {
    H.d_diag.alloc(H.n);
    HIPCHECK(hipMemcpyAsync(H.d_diag.p, diag, sizeof(double) * H.n, kind, st));
  }
  prof_reset(H);
  for (int attempt = 0;; attempt++) {
    if (attempt > 0) prof_reset(H);
    HIPCHECK(hipEventRecord(H.ev0, st));
    run_factor_levels(H, st);
    HIPCHECK(hipEventRecord(H.ev1, st));
    // zero and non-finite pivots (nodes that failed count none), then the
    // per-node results to the host, all behind one synchronisation
    {
      DevPlan P = dev_plan(H);
      hipLaunchKernelGGL(k_pivot_check, dim3((unsigned)H.N), dim3(kWave), 0, st, P,
                         H.d_cnt.p + 3 * (size_t)H.N);
      HIPCHECK(hipGetLastError());
    }
    const size_t nb = sizeof(int) * (size_t)H.N;
    HIPCHECK(hipMemcpyAsync(H.h_stat, H.d_stat.p, nb, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(H.h_npiv, H.d_npiv.p, nb, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(H.h_ndel, H.d_ndel.p, nb, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(H.h_cnt, H.d_cnt.p, 5 * nb, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    bool fail = false;
    for (int k = 0; k < H.N; k++) fail |= H.h_stat[k] != 0;
    if (!fail) break;
    if (attempt >= 20) throw Error(-5, "front re-planning did not converge");
    // Grow the slack of every failed node from the observed requirement.
    std::vector<int> need(H.N), m(H.N);
    HIPCHECK(hipMemcpy(need.data(), H.d_need.p, sizeof(int) * H.N, hipMemcpyDeviceToHost));
    std::vector<int> slack(H.N);
    for (int k = 0; k < H.N; k++) {
      int nominal = H.ns[k] + H.cs[k];
      slack[k] = H.cap[k] - nominal;
      if (H.h_stat[k]) {
        int req = need[k] > 0 ? need[k] - nominal : 2 * slack[k] + 8;
        slack[k] = std::max(slack[k], (int)std::ceil(1.25 * req) + 4);
      }
    }
    build_plan(H, slack);
    upload_static(H);
    H.stats.replans++;
  }
  // solve buffers for at least one right-hand side, so that a solve that
  // may not allocate (kkt_ldl_solve_refined_noalloc) finds them in place
  ensure_u(H, 1);
  if (H.y_nrhs < 1) {
    H.d_Y.alloc((size_t)H.i_total);
    H.y_nrhs = 1;
  }
  HIPCHECK(hipEventElapsedTime(&H.last_factor_ms, H.ev0, H.ev1));
  prof_collect(H);
  // rows below the pivots of the large fronts (the backward GEMV's x staging)
  H.max_below = 0;
  for (int k : H.large_nodes) H.max_below = std::max(H.max_below, H.cap[k] - H.h_npiv[k]);
  // the solve launches only the panel steps that hold pivots
  const int nlev = (int)H.lvl_ptr.size() - 1;
  H.lvl_steps.assign(nlev, 0);
  H.lvl_gfuse.assign(nlev, 1);
  for (int l = 0; l < nlev; l++)
    for (int q = H.large_ptr[l]; q < H.large_ptr[l + 1]; q++) {
      const int np = H.h_npiv[H.large_nodes[q]];
      H.lvl_steps[l] = std::max(H.lvl_steps[l], (np + kPS - 1) / kPS);
      if (np == 0) H.lvl_gfuse[l] = 0;
    }
  int nd = 0, n2 = 0, nz = 0, nf = 0;
  long long stored = 0;
  for (int k = 0; k < H.N; k++) {
    nd += H.h_ndel[k];
    n2 += H.h_cnt[3 * k + 2];
    nz += H.h_cnt[3 * (size_t)H.N + 2 * k];
    nf += H.h_cnt[3 * (size_t)H.N + 2 * k + 1];
  }
  H.stats.num_delayed = nd;
  H.stats.num_2x2 = n2;
  H.stats.num_zero_pivots = nz;
  H.stats.num_nonfinite_pivots = nf;
  H.stats.stored_factor = stored;
  H.factored = true;
  H.stats.t_factor_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (nf > 0) return KKT_LDL_WARN_NONFINITE;
  if (nz > 0) return KKT_LDL_WARN_SINGULAR;
  return 0;
}


template <class T>
static void run_solve_levels_t(Handle& H, double* dx, int ldx, int nrhs, hipStream_t st) {
  ensure_u(H, nrhs);
  if (H.y_nrhs < nrhs) {
    H.d_Y.alloc((size_t)H.i_total * nrhs);
    H.y_nrhs = nrhs;
  }
  DevPlan P = dev_plan(H);
  const int nlev = (int)H.lvl_ptr.size() - 1;
  const int chol = H.opt.solver_type == KKT_LDL_CHOLESKY;
  const int lds_t = (H.max_cap + 31) & ~31;
  P.lds_t = lds_t;
  P.lds_b = kSB;
  // per (level, group) plan copies with their LDS shape
  // (g = 2 l + 1 with tiny fronts joining: the shape covers both groups)
  auto sv_plan = [&](int g, size_t* bytes, bool join = false) {
    DevPlan Q = P;
    Q.lds_t = join ? std::max(H.sv_lt[g], H.sv_lt[g - 1]) : H.sv_lt[g];
    Q.lds_b = std::max(1, join ? std::max(H.sv_lb[g], H.sv_lb[g - 1]) : H.sv_lb[g]);
    *bytes = sizeof(double) * ((size_t)Q.lds_t + (size_t)Q.lds_b * Q.lds_b);
    if (*bytes > 160 * 1024) throw Error(-6, "front too large for the LDS solve kernels");
    return Q;
  };
  size_t lds = 0;
  const size_t lds_bg = sizeof(double) * (size_t)std::max(H.max_below, 1);
  const long long ys = H.i_total, us = H.u_total;
  double* Y = H.d_Y.p;
  constexpr int kGatherThreads = 1024;
  // the level's tiny fronts (one wave each) take their own launch, unless
  // they are few and the level has larger small fronts: then they join
  // those fronts' launch (same results)
  auto tiny_alone = [&](int l) {
    const int nt = H.sv_tiny[l], nall = H.sv_ptr[l + 1] - H.sv_ptr[l];
    return !(nt <= kTinyJoin && nall > nt);
  };
  // levels whose forward gather is done by the first step (see k_fwd_step);
  // one right-hand side only (with more, every step workgroup would gather
  // its rows for each group of right-hand sides again)
  auto gfuse = [&](int l) {
    return H.fuse_gather && nrhs == 1 && (int)H.lvl_gfuse.size() > l && H.lvl_gfuse[l] != 0;
  };
  // levels whose small fronts (one group: no tiny launch of their own) run
  // in the launch of the large nodes' fused first step (forward) and L21^T
  // products (backward)
  auto merge_small = [&](int l, int s0, int s1, int sm, int g0, int g1) {
    return H.merge_levels && gfuse(l) && g1 > g0 && sm == s0 && s1 > s0 &&
           s1 - s0 <= kMergeSmall;
  };
  // a chain's LDS shape and launch (one workgroup per right-hand side)
  auto chain_plan = [&](int l0, size_t* bytes) {
    DevPlan Q = P;
    Q.lds_t = H.sv_run_lt[l0];
    Q.lds_b = H.sv_run_lb[l0];
    *bytes = sizeof(double) * ((size_t)Q.lds_t + (size_t)Q.lds_b * Q.lds_b);
    if (*bytes > 160 * 1024) throw Error(-6, "front too large for the LDS solve kernels");
    return Q;
  };
  for (int l = 0; l < nlev; l++) {
    ProfScope ps_(H, st, KC_FORWARD);
    if (H.solve_chains && H.sv_run_end[l] > l) {
      const int e = H.sv_run_end[l];
      const DevPlan Q = chain_plan(l, &lds);
      hipLaunchKernelGGL((k_forward_chain<T, kTPB>), dim3(1, nrhs), dim3(kTPB), lds, st, Q,
                         H.d_sv_nodes.p + H.sv_ptr[l], H.sv_ptr[e + 1] - H.sv_ptr[l], dx, ldx,
                         us, chol);
      l = e;
      continue;
    }
    const int s0 = H.sv_ptr[l], s1 = H.sv_ptr[l + 1];
    // a few tiny fronts join the level's other small fronts in one launch
    const int sm = s0 + (tiny_alone(l) ? H.sv_tiny[l] : 0);
    const int g0 = H.large_ptr[l], g1 = H.large_ptr[l + 1];
    // few small fronts (all in one group) and a first step that gathers:
    // both in one launch
    const bool mrg = merge_small(l, s0, s1, sm, g0, g1);
    if (mrg) {
      const DevPlan Q = sv_plan(2 * l + 1, &lds, H.sv_tiny[l] > 0);
      const int slot = H.fs_lvl[l];
      const int f0 = H.fs_ptr[slot], f1 = H.fs_ptr[slot + 1];
      const dim3 grid((unsigned)(s1 - s0 + f1 - f0));
      if (s1 - s0 <= kRegLaunch)
        hipLaunchKernelGGL((k_fwd_level<T, true>), grid, dim3(kTPB), lds, st, Q,
                           H.d_sv_nodes.p + s0, s1 - s0, dx, ldx, us, chol, H.d_fs_node.p + f0,
                           H.d_fs_tile.p + f0, Y, ys);
      else
        hipLaunchKernelGGL((k_fwd_level<T, false>), grid, dim3(kTPB), lds, st, Q,
                           H.d_sv_nodes.p + s0, s1 - s0, dx, ldx, us, chol, H.d_fs_node.p + f0,
                           H.d_fs_tile.p + f0, Y, ys);
    }
    // launches of few workgroups are latency bound: their diagonal blocks
    // are read straight into registers (REG); larger ones stage them in LDS
    // and keep more workgroups per CU (same results either way)
    if (!mrg && sm > s0) {
      const DevPlan Q = sv_plan(2 * l, &lds);
      if ((sm - s0) * nrhs <= kRegLaunch)
        hipLaunchKernelGGL((k_forward<T, kWave, true>), dim3(sm - s0, nrhs), dim3(kWave), lds, st,
                           Q, H.d_sv_nodes.p + s0, dx, ldx, nrhs, us, chol);
      else
        hipLaunchKernelGGL((k_forward<T, kWave, false>), dim3(sm - s0, nrhs), dim3(kWave), lds, st,
                           Q, H.d_sv_nodes.p + s0, dx, ldx, nrhs, us, chol);
    }
    if (!mrg && s1 > sm) {
      const DevPlan Q = sv_plan(2 * l + 1, &lds, sm == s0 && H.sv_tiny[l] > 0);
      if ((s1 - sm) * nrhs <= kRegLaunch)
        hipLaunchKernelGGL((k_forward<T, kTPB, true>), dim3(s1 - sm, nrhs), dim3(kTPB), lds, st,
                           Q, H.d_sv_nodes.p + sm, dx, ldx, nrhs, us, chol);
      else
        hipLaunchKernelGGL((k_forward<T, kTPB, false>), dim3(s1 - sm, nrhs), dim3(kTPB), lds, st,
                           Q, H.d_sv_nodes.p + sm, dx, ldx, nrhs, us, chol);
    }
    if (g1 > g0) {
      // the gather done by the first step's workgroups where every large
      // node of the level has pivots (else a gather launch)
      const bool gf = gfuse(l);
      if (!gf) {
        if (nrhs == 1)
          hipLaunchKernelGGL((k_gather_big<T, 1>), dim3(g1 - g0, 1), dim3(kGatherThreads), 0, st,
                             P, H.d_large_nodes.p + g0, dx, ldx, nrhs, us, Y, ys);
        else
          hipLaunchKernelGGL((k_gather_big<T, kGatherRhs>),
                             dim3(g1 - g0, (nrhs + kGatherRhs - 1) / kGatherRhs),
                             dim3(kGatherThreads), 0, st, P, H.d_large_nodes.p + g0, dx, ldx, nrhs,
                             us, Y, ys);
      }
      const int nst = std::min(H.fs_lvl[l + 1] - H.fs_lvl[l], H.lvl_steps[l]);
      for (int t = mrg ? 1 : 0; t < nst; t++) {
        const int slot = H.fs_lvl[l] + t;
        const int f0 = H.fs_ptr[slot], f1 = H.fs_ptr[slot + 1];
        const dim3 g1r(f1 - f0, 1), g4r(f1 - f0, (nrhs + kStepRhs - 1) / kStepRhs);
        const int* fn = H.d_fs_node.p + f0;
        const int* ft = H.d_fs_tile.p + f0;
        if (gf && t == 0) {
          if (nrhs == 1)
            hipLaunchKernelGGL((k_fwd_step<T, 1, true>), g1r, dim3(kTPB), 0, st, P, fn, ft, t, dx,
                               ldx, nrhs, us, Y, ys);
          else
            hipLaunchKernelGGL((k_fwd_step<T, kStepRhs, true>), g4r, dim3(kTPB), 0, st, P, fn, ft,
                               t, dx, ldx, nrhs, us, Y, ys);
        } else {
          if (nrhs == 1)
            hipLaunchKernelGGL((k_fwd_step<T, 1, false>), g1r, dim3(kTPB), 0, st, P, fn, ft, t,
                               dx, ldx, nrhs, us, Y, ys);
          else
            hipLaunchKernelGGL((k_fwd_step<T, kStepRhs, false>), g4r, dim3(kTPB), 0, st, P, fn,
                               ft, t, dx, ldx, nrhs, us, Y, ys);
        }
      }
    }
  }
  for (int l = nlev - 1; l >= 0; l--) {
    ProfScope ps_(H, st, KC_BACKWARD);
    if (H.solve_chains && H.sv_run_start[l] >= 0 && H.sv_run_start[l] < l) {
      const int b = H.sv_run_start[l];
      const DevPlan Q = chain_plan(b, &lds);
      hipLaunchKernelGGL((k_backward_chain<T, kTPB>), dim3(1, nrhs), dim3(kTPB), lds, st, Q,
                         H.d_sv_nodes.p + H.sv_ptr[b], H.sv_ptr[l + 1] - H.sv_ptr[b], dx, ldx,
                         chol);
      l = b;
      continue;
    }
    const int g0 = H.large_ptr[l], g1 = H.large_ptr[l + 1];
    const int s0 = H.sv_ptr[l], s1 = H.sv_ptr[l + 1];
    const int sm = s0 + (tiny_alone(l) ? H.sv_tiny[l] : 0);
    // few small fronts (one group) with the L21^T products: one launch
    const bool mrg = merge_small(l, s0, s1, sm, g0, g1) && lds_bg <= kMergeLds;
    if (g1 > g0) {
      const int b0 = H.bg_ptr[l], b1 = H.bg_ptr[l + 1];
      const int yfw = gfuse(l) ? 1 : 0;
      // one right-hand side per workgroup (batching them here serialises a
      // workgroup's dot products and was measured slower); the x rows below
      // the pivots staged in LDS where they fit
      if (mrg) {
        const DevPlan Q = sv_plan(2 * l + 1, &lds, H.sv_tiny[l] > 0);
        const dim3 grid((unsigned)(s1 - s0 + b1 - b0));
        const size_t bytes = std::max(lds, lds_bg);
        if (s1 - s0 <= kRegLaunch)
          hipLaunchKernelGGL((k_bwd_level<T, true>), grid, dim3(kTPB), bytes, st, Q,
                             H.d_sv_nodes.p + s0, s1 - s0, dx, ldx, chol, H.d_bg_node.p + b0,
                             H.d_bg_tile.p + b0, Y, ys, H.max_below, yfw);
        else
          hipLaunchKernelGGL((k_bwd_level<T, false>), grid, dim3(kTPB), bytes, st, Q,
                             H.d_sv_nodes.p + s0, s1 - s0, dx, ldx, chol, H.d_bg_node.p + b0,
                             H.d_bg_tile.p + b0, Y, ys, H.max_below, yfw);
      } else if (lds_bg <= 128 * 1024)
        hipLaunchKernelGGL((k_bwd_gemv<T, 1>), dim3(b1 - b0, nrhs), dim3(kTPB), lds_bg, st, P,
                           H.d_bg_node.p + b0, H.d_bg_tile.p + b0, dx, ldx, nrhs, Y, ys,
                           H.max_below, yfw);
      else
        hipLaunchKernelGGL((k_bwd_gemv<T, 0>), dim3(b1 - b0, nrhs), dim3(kTPB), 0, st, P,
                           H.d_bg_node.p + b0, H.d_bg_tile.p + b0, dx, ldx, nrhs, Y, ys, 0, yfw);
      const int nst = std::min(H.bs_lvl[l + 1] - H.bs_lvl[l], H.lvl_steps[l]);
      for (int t = 0; t < nst; t++) {
        const int slot = H.bs_lvl[l] + t;
        const int f0 = H.bs_ptr[slot], f1 = H.bs_ptr[slot + 1];
        if (nrhs == 1)
          hipLaunchKernelGGL((k_bwd_step<T, 1>), dim3(f1 - f0, 1), dim3(kTPB), 0, st, P,
                             H.d_bs_node.p + f0, H.d_bs_tile.p + f0, t, dx, ldx, nrhs, Y, ys);
        else
          hipLaunchKernelGGL((k_bwd_step<T, kStepRhs>),
                             dim3(f1 - f0, (nrhs + kStepRhs - 1) / kStepRhs), dim3(kTPB), 0, st,
                             P, H.d_bs_node.p + f0, H.d_bs_tile.p + f0, t, dx, ldx, nrhs, Y, ys);
      }
    }
    if (!mrg && sm > s0) {
      const DevPlan Q = sv_plan(2 * l, &lds);
      if ((sm - s0) * nrhs <= kRegLaunch)
        hipLaunchKernelGGL((k_backward<T, kWave, true>), dim3(sm - s0, nrhs), dim3(kWave), lds, st,
                           Q, H.d_sv_nodes.p + s0, dx, ldx, nrhs, chol);
      else
        hipLaunchKernelGGL((k_backward<T, kWave, false>), dim3(sm - s0, nrhs), dim3(kWave), lds, st,
                           Q, H.d_sv_nodes.p + s0, dx, ldx, nrhs, chol);
    }
    if (!mrg && s1 > sm) {
      const DevPlan Q = sv_plan(2 * l + 1, &lds, sm == s0 && H.sv_tiny[l] > 0);
      if ((s1 - sm) * nrhs <= kRegLaunch)
        hipLaunchKernelGGL((k_backward<T, kTPB, true>), dim3(s1 - sm, nrhs), dim3(kTPB), lds, st,
                           Q, H.d_sv_nodes.p + sm, dx, ldx, nrhs, chol);
      else
        hipLaunchKernelGGL((k_backward<T, kTPB, false>), dim3(s1 - sm, nrhs), dim3(kTPB), lds, st,
                           Q, H.d_sv_nodes.p + sm, dx, ldx, nrhs, chol);
    }
  }
  HIPCHECK(hipGetLastError());
}

static void run_solve_levels(Handle& H, double* dx, int ldx, int nrhs, hipStream_t st) {
  if (H.esize == 4) run_solve_levels_t<float>(H, dx, ldx, nrhs, st);
  else run_solve_levels_t<double>(H, dx, ldx, nrhs, st);
}

// Buffers of the refinement: right-hand side copy, residual / correction,
// previous iterate (each ldx * nrhs) and norm partials ((kNormBlocks + 2) *
// nrhs).  Given by kkt_ldl_solve_refined_noalloc (its static scratch);
// otherwise the handle's own buffers, allocated on demand.
struct RefineWork {
  double *b, *r, *xb, *norm;
};

static int solve_impl(Handle& H, int nrhs, double* x, int ldx, int mem,
                      hipStream_t st, bool refine, const RefineWork* ws = nullptr) {
  if (!H.factored) throw Error(-3, "solve called before factor");
  if (nrhs <= 0) return 0;
  if (ldx < H.n) throw Error(-1, "ldx < n");
  const size_t tot = (size_t)ldx * nrhs;
  double* dx = x;
  if (!ws && (mem == KKT_LDL_HOST || refine)) {
    H.d_x.alloc(tot);
    HIPCHECK(hipMemcpyAsync(H.d_x.p, x, tot * sizeof(double),
                            mem == KKT_LDL_HOST ? hipMemcpyHostToDevice
                                                : hipMemcpyDeviceToDevice, st));
    dx = H.d_x.p;
  }
  int status = 0;
  if (!ws) prof_reset(H);  // (no event is destroyed by the scratch path)
  HIPCHECK(hipEventRecord(H.ev2, st));
  if (refine) {
    RefineWork w;
    if (ws) {
      w = *ws;
    } else {
      H.d_b.alloc(tot);
      H.d_r.alloc(tot);
      H.d_xb.alloc(tot);
      H.d_norm.alloc(2 * (size_t)nrhs + (size_t)kNormBlocks * nrhs);
      w = RefineWork{H.d_b.p, H.d_r.p, H.d_xb.p, H.d_norm.p};
    }
    HIPCHECK(hipMemcpyAsync(w.b, dx, tot * sizeof(double), hipMemcpyDeviceToDevice, st));
    // host copies of the squared norms (the handle's pinned buffers, set up
    // by the first factorization: no allocation here)
    std::vector<double> bnv, rnv;
    double* bnp = H.h_norm;
    double* rnp = H.h_norm + kMaxNormRhs;
    if (nrhs > kMaxNormRhs) {
      bnv.resize(nrhs);
      rnv.resize(nrhs);
      bnp = bnv.data();
      rnp = rnv.data();
    }
    double* part = w.norm + 2 * (size_t)nrhs;
    auto colnorm2 = [&](const double* v, double* out) {
      hipLaunchKernelGGL(k_colnorm2_part, dim3(kNormBlocks, nrhs), dim3(kTPB), 0, st, H.n, v,
                         ldx, part);
      hipLaunchKernelGGL(k_colnorm2_final, dim3((nrhs + 63) / 64), dim3(64), 0, st, nrhs,
                         (const double*)part, out);
    };
    colnorm2(w.b, w.norm);
    HIPCHECK(hipMemcpyAsync(bnp, w.norm, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
    run_solve_levels(H, dx, ldx, nrhs, st);
    int it = 0;
    double best = std::numeric_limits<double>::infinity();
    const dim3 rg((unsigned)((H.n * 16 + kTPB - 1) / kTPB));
    for (;; it++) {
      {
        ProfScope ps_(H, st, KC_REFINE);
        hipLaunchKernelGGL(k_residual, rg, dim3(kTPB), 0, st, H.n, H.d_rowp.p, H.d_cols.p,
                           H.d_vals.p, H.has_diag ? H.d_diag.p : nullptr, dx, w.b,
                           w.r, ldx, nrhs);
      }
      colnorm2(w.r, w.norm);
      HIPCHECK(hipMemcpyAsync(rnp, w.norm, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      const double wr = refine_worst(bnp, rnp, nrhs);
      const int act = refine_next(it, wr, best, H.opt.refine_tol, H.opt.max_refine);
      if (act == REFINE_UNDO_LAST) {
        // the last correction did not lower the residual: back to the
        // iterate before it, whose residual is `best`
        HIPCHECK(hipMemcpyAsync(dx, w.xb, tot * sizeof(double), hipMemcpyDeviceToDevice, st));
        it--;
        break;
      }
      best = wr;
      if (act == REFINE_DONE) break;
      run_solve_levels(H, w.r, ldx, nrhs, st);
      const size_t tb = ((size_t)H.n * nrhs + kTPB - 1) / kTPB;
      hipLaunchKernelGGL(k_axpy_cols, dim3((unsigned)tb), dim3(kTPB), 0, st, H.n, dx, w.r,
                         w.xb, ldx, nrhs);
    }
    H.stats.refine_steps = it;
    H.stats.residual = best;
    if (!(best <= H.opt.refine_tol)) status = KKT_LDL_WARN_NOT_CONVERGED;
  } else {
    run_solve_levels(H, dx, ldx, nrhs, st);
  }
  HIPCHECK(hipEventRecord(H.ev3, st));
  if (dx != x)
    HIPCHECK(hipMemcpyAsync(x, dx, tot * sizeof(double),
                            mem == KKT_LDL_HOST ? hipMemcpyDeviceToHost
                                                : hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
  HIPCHECK(hipEventElapsedTime(&H.last_solve_ms, H.ev2, H.ev3));
  if (!ws) prof_collect(H);
  return status;
}

static void create_impl(Handle& H, int n, const int* rowp, const int* cols,
                        const kkt_ldl_options* o) {
  auto t0 = std::chrono::steady_clock::now();
  if (o) H.opt = *o;
  else kkt_ldl_default_options(&H.opt);
  if (n <= 0) throw Error(-1, "n must be positive");
  if (H.opt.precision != KKT_LDL_FP64 && H.opt.precision != KKT_LDL_FP32_REFINE)
    throw Error(-1, "unknown precision");
  H.esize = (H.opt.precision == KKT_LDL_FP32_REFINE && H.opt.solver_type == KKT_LDL_LDL) ? 4 : 8;
  H.n = n;
  H.nnz = rowp[n];
  H.rowp.assign(rowp, rowp + n + 1);
  H.cols.assign(cols, cols + H.nnz);
  for (int i = 0; i < n; i++) {
    if (rowp[i + 1] < rowp[i]) throw Error(-1, "rowp not monotone");
    for (int p = rowp[i]; p < rowp[i + 1]; p++) {
      if (cols[p] < 0 || cols[p] >= n) throw Error(-1, "column index out of range");
      if (p > rowp[i] && cols[p] <= cols[p - 1]) throw Error(-1, "columns not sorted/unique");
    }
  }
  std::vector<int> perm;
  const int* pp = nullptr;
  if (H.opt.order == KKT_LDL_ORDER_MULTI_COLOR)
    throw Error(-1, "MULTI_COLOR ordering is not supported");
  if (H.opt.order != KKT_LDL_ORDER_ND && H.opt.order != KKT_LDL_ORDER_AMD &&
      H.opt.order != KKT_LDL_ORDER_NATURAL && H.opt.order != KKT_LDL_ORDER_GIVEN)
    throw Error(-1, "unknown ordering");
  if (H.opt.solver_type != KKT_LDL_LDL && H.opt.solver_type != KKT_LDL_CHOLESKY)
    throw Error(-1, "unknown solver type");
  if (H.opt.order == KKT_LDL_ORDER_AMD) {
    perm.resize(n);
    amd_ordering(n, rowp, cols, perm.data());
    pp = perm.data();
  } else if (H.opt.order == KKT_LDL_ORDER_ND) {
    perm.resize(n);
    nd_ordering(n, rowp, cols, perm.data());
    pp = perm.data();
  } else if (H.opt.order == KKT_LDL_ORDER_GIVEN) {
    if (!H.opt.perm) throw Error(-1, "ORDER_GIVEN without perm");
    perm.assign(H.opt.perm, H.opt.perm + n);
    std::vector<char> seen(n, 0);
    for (int v : perm) {
      if (v < 0 || v >= n || seen[v]) throw Error(-1, "perm is not a permutation");
      seen[v] = 1;
    }
    pp = perm.data();
  }
  analyze(n, rowp, cols, pp, H.S);
  if (H.opt.order == KKT_LDL_ORDER_ND && H.opt.vals_hint) {
    // predicted 2x2 partners adjacent inside every fundamental supernode
    // (same fill and supernodes, new order inside them)
    std::vector<int> paired(n);
    pair_partners(n, rowp, cols, H.opt.vals_hint, H.opt.ustab, H.S, paired.data());
    SymbolicResult S2;
    analyze(n, rowp, cols, paired.data(), S2);
    H.S = std::move(S2);
  }
  H.opt.perm = nullptr;
  H.opt.vals_hint = nullptr;
  build_maps(H);
  if (const char* e = std::getenv("KKT_LDL_LEVEL_SCHEDULE")) H.level_sched = std::atoi(e) != 0;
  // KKT_LDL_RL_FRONTS=0 keeps the blocked LDS front kernel (k_front_lds)
  // and one launch per tick for every mode
  H.rl_fronts = H.opt.solver_type != KKT_LDL_CHOLESKY;
  if (const char* e = std::getenv("KKT_LDL_RL_FRONTS")) H.rl_fronts &= std::atoi(e) != 0;
  H.lds_front = std::min(H.lds_front, rl_max_front());
  if (const char* e = std::getenv("KKT_LDL_ROOT_BK")) H.root_bk = std::atoi(e) != 0;
  if (const char* e = std::getenv("KKT_LDL_FUSE_GATHER")) H.fuse_gather = std::atoi(e) != 0;
  if (const char* e = std::getenv("KKT_LDL_MERGE_LEVELS")) H.merge_levels = std::atoi(e) != 0;
  if (const char* e = std::getenv("KKT_LDL_RL_INSTANCES")) H.rl_sized = std::atoi(e) != 0;
  if (const char* e = std::getenv("KKT_LDL_SIDE_STREAM")) H.side_lds = std::atoi(e) != 0;
  if (const char* e = std::getenv("KKT_LDL_SOLVE_CHAINS")) H.solve_chains = std::atoi(e) != 0;
  build_plan(H, initial_slack(H));
  std::memset(&H.stats, 0, sizeof(H.stats));
  H.stats.n = n;
  H.stats.num_nodes = H.N;
  H.stats.num_snodes = H.S.num_snodes;
  H.stats.num_levels = (int)H.lvl_ptr.size() - 1;
  int mf = 0;
  for (int k = 0; k < H.N; k++) mf = std::max(mf, H.ns[k] + H.cs[k]);
  H.stats.max_front = mf;
  double fl = 0.0;
  long long nz = 0;
  for (int j = 0; j < n; j++) {
    double c = H.S.colcount[j] - 1;
    fl += c * (c + 2.0);
    nz += H.S.colcount[j];
  }
  H.stats.flops_factor = fl;
  H.stats.nnz_l = nz;
  // critical path of the assembly tree in pivot columns, and the sum over
  // levels of the largest panel count, which bounds the panel launches of
  // the level schedule (children are numbered before their parents)
  {
    std::vector<long long> path(H.N, 0);
    long long cp = 0;
    for (int k = 0; k < H.N; k++) {
      long long below = 0;
      for (int q = H.S.children_ptr[k]; q < H.S.children_ptr[k + 1]; q++)
        below = std::max(below, path[H.S.children[q]]);
      path[k] = below + H.ns[k];
      cp = std::max(cp, path[k]);
    }
    H.stats.crit_path_pivots = cp;
    const int nlev = (int)H.lvl_ptr.size() - 1;
    long long lp = 0;
    for (int l = 0; l < nlev; l++) lp += H.large_panels[l];
    H.stats.level_panel_sum = lp;
    long long pl = 0;
    for (int t = 0; t < H.ntick; t++) pl += H.tk_act_ptr[t + 1] > H.tk_act_ptr[t];
    H.stats.panel_launches = pl;
  }
  H.stats.t_analyze_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace kktldl

using kktldl::Error;
using kktldl::Handle;

// --------------------------------------------------------------------------
// Scratch of kkt_ldl_solve_refined_noalloc: the refinement's copy of the
// right-hand sides, its residual / correction and previous iterate (each up
// to kNoAllocEntries = ldx * nrhs entries) and the norm partials of up to
// kNoAllocRhs columns: 25.2 MB, one copy per GPU, held until the process
// ends.  In host code `scratch` is a stand-in for the symbol, not a device
// address: the address is taken per device with hipGetSymbolAddress on the
// entry point's first call there, and only that address is handed to
// kernels and copies.  Every call on a GPU shares its copy, so calls that
// overlap on one GPU race; each call writes every scratch entry it reads
// (copy, residual, partials), so nothing has to start at zero.
// --------------------------------------------------------------------------
constexpr size_t kNoAllocEntries = size_t(1) << 20;
constexpr int kNoAllocRhs = 64;
constexpr size_t kScratchDoubles =
    3 * kNoAllocEntries + (size_t)(kktldl::kNormBlocks + 2) * kNoAllocRhs;
static __device__ __align__(256) double scratch[kScratchDoubles];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[64] = {};  // device address of `scratch`, per device

static double* scratch_address() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error(-2, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(-2, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  return g_scratch_dev[dev];
}

struct kkt_ldl_handle {
  Handle h;
};

template <class Fn>
static int guarded(Fn&& fn) {
  try {
    return fn();
  } catch (const Error& e) {
    kktldl::g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    kktldl::g_last_error = e.what();
    return -100;
  } catch (...) {
    kktldl::g_last_error = "unknown error";
    return -101;
  }
}

extern "C" {

void kkt_ldl_default_options(kkt_ldl_options* o) {
  o->order = KKT_LDL_ORDER_AMD;
  o->solver_type = KKT_LDL_LDL;
  o->ustab = 0.01;
  o->pivot_tol = 1e-14;
  o->delay_growth = 2.0;
  o->precision = KKT_LDL_FP64;
  o->refine_tol = 1e-10;
  o->max_refine = 10;
  o->perm = nullptr;
  o->vals_hint = nullptr;
}

int kkt_ldl_create(int n, const int* rowp, const int* cols,
                   const kkt_ldl_options* opts, kkt_ldl_t* out) {
  return guarded([&] {
    if (!out || !rowp || !cols) throw Error(-1, "null argument");
    std::unique_ptr<kkt_ldl_handle> h(new kkt_ldl_handle());
    kktldl::create_impl(h->h, n, rowp, cols, opts);
    if (const char* e = std::getenv("KKT_LDL_PANEL_SPECULATION")) h->h.panel_spec = std::atoi(e) != 0;
    *out = h.release();
    return 0;
  });
}

int kkt_ldl_factor(kkt_ldl_t h, const double* vals, const double* diag, int mem,
                   void* stream) {
  return guarded([&] {
    if (!h || !vals) throw Error(-1, "null argument");
    return kktldl::factor_impl(h->h, vals, diag, mem, (hipStream_t)stream);
  });
}

int kkt_ldl_refactor(kkt_ldl_t h, const double* diag, int mem, void* stream) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    return kktldl::factor_impl(h->h, nullptr, diag, mem, (hipStream_t)stream);
  });
}

int kkt_ldl_set_pivot_threshold(kkt_ldl_t h, double ustab, double pivot_tol) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    if (std::isnan(ustab) || std::isnan(pivot_tol)) throw Error(-1, "NaN pivot threshold");
    if (ustab >= 0.0) h->h.opt.ustab = ustab;
    if (pivot_tol >= 0.0) h->h.opt.pivot_tol = pivot_tol;
    return 0;
  });
}

int kkt_ldl_solve(kkt_ldl_t h, int nrhs, double* x, int ldx, int mem, void* stream) {
  return guarded([&] {
    if (!h || !x) throw Error(-1, "null argument");
    return kktldl::solve_impl(h->h, nrhs, x, ldx, mem, (hipStream_t)stream, false);
  });
}

int kkt_ldl_solve_refined(kkt_ldl_t h, int nrhs, double* x, int ldx, int mem,
                          void* stream) {
  return guarded([&] {
    if (!h || !x) throw Error(-1, "null argument");
    return kktldl::solve_impl(h->h, nrhs, x, ldx, mem, (hipStream_t)stream, true);
  });
}

int kkt_ldl_solve_refined_noalloc(kkt_ldl_t h, int nrhs, double* x, int ldx, void* stream) {
  return guarded([&] {
    if (!h || !x) throw Error(-1, "null argument");
    Handle& H = h->h;
    if (!H.factored) throw Error(-3, "solve called before factor");
    if (nrhs <= 0) return 0;
    if (ldx < H.n) throw Error(-1, "ldx < n");
    if (nrhs > kNoAllocRhs || (size_t)ldx * (size_t)nrhs > kNoAllocEntries)
      throw Error(-7, "kkt_ldl_solve_refined_noalloc: ldx * nrhs = " +
                          std::to_string((size_t)ldx * (size_t)nrhs) + " with nrhs = " +
                          std::to_string(nrhs) + " exceeds its scratch (" +
                          std::to_string(kNoAllocEntries) + " entries, at most " +
                          std::to_string(kNoAllocRhs) + " right-hand sides)");
    if (H.u_nrhs < nrhs || H.y_nrhs < nrhs)
      throw Error(-7, "kkt_ldl_solve_refined_noalloc: the handle's solve buffers hold " +
                          std::to_string(std::min(H.u_nrhs, H.y_nrhs)) +
                          " right-hand sides; one kkt_ldl_solve or kkt_ldl_solve_refined "
                          "with at least " + std::to_string(nrhs) + " sizes them");
    double* base = scratch_address();
    const kktldl::RefineWork w{base, base + kNoAllocEntries, base + 2 * kNoAllocEntries,
                               base + 3 * kNoAllocEntries};
    // per-kernel timing creates events: off for this call
    const bool prof = H.profiling;
    H.profiling = false;
    int st = 0;
    try {
      st = kktldl::solve_impl(H, nrhs, x, ldx, KKT_LDL_DEVICE, (hipStream_t)stream, true, &w);
    } catch (...) {
      H.profiling = prof;
      throw;
    }
    H.profiling = prof;
    return st;
  });
}

int kkt_ldl_inertia(kkt_ldl_t h, int* npos, int* nneg, int* nzero) {
  return guarded([&] {
    if (!h || !npos || !nneg) throw Error(-1, "null argument");
    Handle& H = h->h;
    if (!H.factored) throw Error(-3, "inertia requested before factor");
    int a = 0, b = 0;
    if (H.opt.solver_type != KKT_LDL_CHOLESKY)
      for (int k = 0; k < H.N; k++) {
        a += H.h_cnt[3 * k];
        b += H.h_cnt[3 * k + 1];
      }
    *npos = a;
    *nneg = b;
    if (nzero) {
      int z = 0;
      for (int k = 0; k < H.N; k++) z += H.h_cnt[3 * (size_t)H.N + 2 * k];
      *nzero = z;
    }
    return 0;
  });
}

int kkt_ldl_get_perm(kkt_ldl_t h, int* perm, int* iperm) {
  return guarded([&] {
    if (!h || !perm || !iperm) throw Error(-1, "null argument");
    const auto& S = h->h.S;
    for (int i = 0; i < S.n; i++) {
      perm[i] = S.permuted ? S.perm[i] : i;
      iperm[i] = S.permuted ? S.iperm[i] : i;
    }
    return 0;
  });
}

int kkt_ldl_get_sizes(kkt_ldl_t h, int* num_nodes, int* contrib_total) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    if (num_nodes) *num_nodes = h->h.S.num_nodes;
    if (contrib_total) *contrib_total = h->h.S.contrib_ptr[h->h.S.num_nodes];
    return 0;
  });
}

int kkt_ldl_get_supernodes(kkt_ldl_t h, int* node_ptr, int* node_to_var,
                           int* children_ptr, int* children, int* contrib_ptr,
                           int* contrib_rows) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    const auto& S = h->h.S;
    auto cp = [](const std::vector<int>& v, int* o) {
      if (o) std::copy(v.begin(), v.end(), o);
    };
    cp(S.node_ptr, node_ptr);
    cp(S.node_to_var, node_to_var);
    cp(S.children_ptr, children_ptr);
    cp(S.children, children);
    cp(S.contrib_ptr, contrib_ptr);
    cp(S.contrib_rows, contrib_rows);
    return 0;
  });
}

int kkt_ldl_get_node_info(kkt_ldl_t h, int* npiv, int* ndel, int* n2x2, int* root) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    Handle& H = h->h;
    if (!H.factored) throw Error(-3, "node info requested before factor");
    for (int k = 0; k < H.N; k++) {
      if (npiv) npiv[k] = H.h_npiv[k];
      if (ndel) ndel[k] = H.h_ndel[k];
      if (n2x2) n2x2[k] = H.h_cnt[3 * k + 2];
      if (root) root[k] = H.S.node_parent[k] < 0;
    }
    return 0;
  });
}

int kkt_ldl_get_stats(kkt_ldl_t h, kkt_ldl_stats* st) {
  return guarded([&] {
    if (!h || !st) throw Error(-1, "null argument");
    h->h.stats.front_bytes = h->h.f_total * (long long)h->h.esize;
    *st = h->h.stats;
    return 0;
  });
}

int kkt_ldl_last_times(kkt_ldl_t h, double* factor_ms, double* solve_ms) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    if (factor_ms) *factor_ms = h->h.last_factor_ms;
    if (solve_ms) *solve_ms = h->h.last_solve_ms;
    return 0;
  });
}

int kkt_ldl_set_profiling(kkt_ldl_t h, int on) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    h->h.profiling = on != 0;
    return 0;
  });
}

int kkt_ldl_get_kernel_times(kkt_ldl_t h, double* ms, int* launches, int nclass) {
  return guarded([&] {
    if (!h) throw Error(-1, "null argument");
    for (int c = 0; c < nclass && c < kktldl::KC_COUNT; c++) {
      if (ms) ms[c] = h->h.prof_ms[c];
      if (launches) launches[c] = h->h.prof_n[c];
    }
    return (int)kktldl::KC_COUNT;
  });
}

const char* kkt_ldl_last_error(void) { return kktldl::g_last_error.c_str(); }

void kkt_ldl_destroy(kkt_ldl_t h) {
  if (!h) return;
  Handle& H = h->h;
  if (H.ev0) (void)hipEventDestroy(H.ev0);
  if (H.ev1) (void)hipEventDestroy(H.ev1);
  if (H.ev2) (void)hipEventDestroy(H.ev2);
  if (H.ev3) (void)hipEventDestroy(H.ev3);
  if (H.ev_fork) (void)hipEventDestroy(H.ev_fork);
  for (auto e : H.lds_ev) (void)hipEventDestroy(e);
  if (H.side) (void)hipStreamDestroy(H.side);
  if (H.h_pin) (void)hipHostFree(H.h_pin);
  if (H.h_norm) (void)hipHostFree(H.h_norm);
  delete h;
}

}  // extern "C"
