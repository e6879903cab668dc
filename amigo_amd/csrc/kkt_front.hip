// Fronts held whole in LDS (dimension <= the LDS front bound, 96 by
// default): assembled, factored and written back by one workgroup each.
//
//  * k_front_rl   -- one workgroup per front of a tick (the bottom of the
//                    tree: many independent small fronts).
//  * k_front_chain -- ONE workgroup walks a list of fronts in elimination
//                    order, each front's children finished before it (a
//                    chain of single-front ticks, as banded KKTs give under
//                    minimum degree: a tree that is one long path).  One
//                    launch replaces two launches and one dependency per
//                    front.
//
// The front is kept in FP64 in LDS whatever the front element type T; the
// L columns are rounded to T as they are produced (and used rounded in the
// trailing updates, as a front stored in T would use them), the rest is
// rounded to T when the front is written back.
//
// Pivoting is the reference's threshold 1x1 / 2x2 test with delayed pivots
// (include/sparse_ldl.h:1316-1553) or Bunch-Kaufman on root fronts (the
// reference calls LAPACK dsytrf there, :1226-1248), the same decisions as
// the blocked panel of the larger fronts (panel_factor), but right-looking:
// after each pivot the whole trailing triangle is updated, which makes a
// pivot step one column reduction, one parallel rank-1 (or rank-2) update
// and ONE workgroup barrier.  Every wave computes the column reductions and
// the pivot decision itself from the same LDS data (they agree bit for
// bit), so the decision needs no barrier; the scaling of a pivot column is
// deferred to the next step, whose reads never touch it.  Rows are swapped
// (delays, 1x1 pivots at imax, 2x2 partners) between two barriers.
#include "kkt_common.hpp"

namespace kktldl {

constexpr int kRLThreads = 256;
constexpr int kRLMaxFront = 96;  // largest front dimension of these kernels
// trailing-triangle entries per thread for fronts of at most MF rows:
// ceil(MF (MF + 1) / 2 / 256)
template <int MF>
constexpr int rl_entries() {
  return (MF * (MF + 1) / 2 + kRLThreads - 1) / kRLThreads;
}
// trailing entries whose loads are issued together: at most 10, in as few
// rounds as that allows, split evenly
template <int MF>
constexpr int rl_batch() {
  constexpr int n = rl_entries<MF>(), rounds = (n + 9) / 10;
  return (n + rounds - 1) / rounds;
}
// launches whose fronts have at most 32, 48, 64 or 80 rows take instances
// with fewer entries per thread (fewer registers: more workgroups per CU;
// the same arithmetic)

template <class T>
__device__ inline double round_t(double x) {
  return (double)(T)x;
}

// Per-thread entries of the trailing triangle.  Entry e <-> (jr, ir), e =
// jr (jr + 1) / 2 + ir, 0 <= ir <= jr, counted from the bottom-right corner
// of the front (row m - 1 - ir, column m - 1 - jr): the entries of the
// trailing triangle from column j0 on are the prefix e < (m - j0)(m - j0 +
// 1) / 2 whatever m, so one table serves every front and step.
template <int MF>
struct TriEntries {
  static constexpr int kN = rl_entries<MF>();
  int pk[kN];
  __device__ void init(int tid) {
#pragma unroll
    for (int r = 0; r < kN; r++) {
      const int e = tid + kRLThreads * r;
      int jr = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while (jr * (jr + 1) / 2 > e) jr--;
      while ((jr + 1) * (jr + 2) / 2 <= e) jr++;
      pk[r] = (jr << 8) | (e - jr * (jr + 1) / 2);
    }
  }
};

struct RLShared {
  int fail, din;
};

// Prepare (delays received from the children, their offsets in this
// front, the front dimension, overflow check), done by thread 0 as
// k_prepare does it for the fronts assembled in HBM.  Returns false when
// the front is not factored (a failed child or an overflow).
__device__ inline bool rl_prepare(const DevPlan& P, int p, RLShared& sh) {
  if (threadIdx.x == 0) {
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
    P.npiv[p] = 0;
    P.ndel[p] = 0;
    if (fail || m > P.cap[p]) {
      P.stat[p] = 1;
      P.need[p] = fail ? -1 : m;
      fail = 1;
    } else {
      P.stat[p] = 0;
    }
    sh.fail = fail;
    sh.din = din;
  }
  __syncthreads();
  return !sh.fail;
}

// One front: assemble into LDS (FP64), factor, write back.  lds: F (m x m)
// | dv | dv2 | vars | pre | ptype.
template <class T, int MF>
__device__ void front_rl(const DevPlan& P, int p, double* lds, const TriEntries<MF>& te,
                         RLShared& sh) {
  // every product rounded before it is added (no fused multiply-adds): the
  // compiler would otherwise fuse them per instance, and the instances for
  // different front bounds would differ in the last bits
#pragma clang fp contract(off)
  constexpr int kRLEntries = rl_entries<MF>();
  constexpr int kRLBatch = rl_batch<MF>();
  const int tid = threadIdx.x, lane = tid % kWave;
  if (!rl_prepare(P, p, sh)) return;
  const int din = sh.din;
  const int ns = P.ns[p], cs = P.cs[p];
  const int m = ns + din + cs, fs = ns + din, mode = P.mode[p];
  const int vp = P.vptr[p], cp = P.cptr[p];
  const int c0 = P.chptr[p], c1 = P.chptr[p + 1];
  double* F = lds;
  double* dv = F + (size_t)m * m;
  double* dv2 = dv + m;
  int* vars = (int*)(dv2 + m);
  int* pre = vars + m;
  int* ptype = pre + m;
  // ---- assembly: zero, children in child order, original entries, diagonal
  for (int e = tid; e < m * m; e += kRLThreads) F[e] = 0.0;
  for (int i = tid; i < ns; i += kRLThreads) vars[i] = P.node_to_var[vp + i];
  for (int i = tid; i < cs; i += kRLThreads) vars[fs + i] = P.contrib_rows[cp + i];
  for (int i = tid; i < m; i += kRLThreads) {
    pre[i] = i;
    ptype[i] = 0;
  }
  __syncthreads();
  for (int q = c0; q < c1; q++) {
    const int c = P.ch[q];
    const int mc = P.m[c], npc = P.npiv[c], ndc = P.ndel[c], doc = P.doff[c];
    const int fsc = P.ns[c] + P.din[c];
    const int cpc = P.cptr[c];
    const T* Fc = front_ptr<T>(P, c);
    const int* cv = P.vars + P.ioff[c] + npc;
    for (int t = tid; t < ndc; t += kRLThreads) vars[ns + doc + t] = cv[t];
    // the child's contribution block: rows npc .. mc-1, lower triangle
    const int w = mc - npc;
    const int ne = w * (w + 1) / 2;
    for (int e = tid; e < ne; e += kRLThreads) {
      int jr = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while (jr * (jr + 1) / 2 > e) jr--;
      while ((jr + 1) * (jr + 2) / 2 <= e) jr++;
      const int ir = e - jr * (jr + 1) / 2;
      const int r1 = mc - 1 - ir, r2 = mc - 1 - jr;  // r1 >= r2
      // parent positions: delayed rows of the child, then its nominal
      // contribution rows through the static forward map
      const int u1 = P.fwdnom[cpc + max(r1 - fsc, 0)];
      const int u2 = P.fwdnom[cpc + max(r2 - fsc, 0)];
      const double v = (double)Fc[r1 + (size_t)r2 * mc];
      const int p1 = r1 < fsc ? ns + doc + (r1 - npc) : (u1 < ns ? u1 : u1 + din);
      const int p2 = r2 < fsc ? ns + doc + (r2 - npc) : (u2 < ns ? u2 : u2 + din);
      const int a = max(p1, p2), b = min(p1, p2);
      F[a + (size_t)b * m] += v;
    }
    __syncthreads();
  }
  for (int e = P.am_ptr[vp] + tid; e < P.am_ptr[vp + ns]; e += kRLThreads) {
    const int u = P.am_row[e], j = P.am_col[e];
    F[(u < ns ? u : u + din) + (size_t)j * m] += P.vals[P.am_src[e]];
  }
  __syncthreads();
  if (P.diag)
    for (int j = tid; j < ns; j += kRLThreads) F[(size_t)j * m + j] += P.diag[P.node_to_var[vp + j]];
  __syncthreads();
  // ---- factorization
  const double alpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8
  const double ustab = P.ustab, ptol = P.ptol;
  int k = 0, nc = mode == 1 ? m : fs;
  int npos = 0, nneg = 0, n2 = 0;
  // the pivot whose L columns are still unscaled in F
  int pk = -1, pstep = 0;
  bool pzero = false;
  double pm0 = 0.0, pm1 = 0.0, pm2 = 0.0;  // 1/d, or the 2x2 inverse i00, i01, i11
  auto finalize = [&]() {
    if (pk < 0) return;
    if (pstep == 1) {
      for (int i = pk + 1 + tid; i < m; i += kRLThreads) {
        const double v = F[i + (size_t)pk * m];
        F[i + (size_t)pk * m] = round_t<T>(pzero ? v : v * pm0);
      }
    } else {
      for (int i = pk + 2 + tid; i < m; i += kRLThreads) {
        const double a = F[i + (size_t)pk * m], b = F[i + (size_t)(pk + 1) * m];
        F[i + (size_t)pk * m] = round_t<T>(a * pm0 + b * pm1);
        F[i + (size_t)(pk + 1) * m] = round_t<T>(a * pm1 + b * pm2);
      }
    }
    pk = -1;
  };
  while (k < nc) {
    finalize();
    // column k: best candidate below the diagonal (first row on ties, NaN
    // never chosen) and the largest entry among the rows that are not
    // candidates.  Lane l holds rows k+1+l and k+65+l: two max reductions,
    // then the first lane holding the maximum, rows k+1.. before k+65..
    double cb[2], rest = 0.0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int i = k + 1 + lane + kWave * h;
      const double a = fabs(F[min(i, m - 1) + (size_t)k * m]);
      cb[h] = (i < nc && a >= 0.0) ? a : -1.0;
      if (i < m && i >= nc) rest = fmax(rest, a);
    }
    const double lbest = wave_max(fmax(cb[0], cb[1]));
    const double lrest = wave_max(rest);
    int lidx = 0x7fffffff;
    if (lbest >= 0.0) {
      const unsigned long long b0 = __ballot(cb[0] == lbest);
      lidx = b0 ? k + 1 + __ffsll((long long)b0) - 1
                : k + 1 + kWave + __ffsll((long long)__ballot(cb[1] == lbest)) - 1;
    }
    const double wkk = F[k + (size_t)k * m];
    const double awkk = fabs(wkk);
    int imax = 0;
    double awik = 0.0;
    if (k + 1 < nc) {
      const bool found = lidx < nc;
      imax = found ? lidx : k + 1;
      awik = found ? lbest : __builtin_nan("");
    }
    const double colmax = fmax(awik, lrest);
    int step = 0, kpiv = -1;
    bool zero_col = false, need_row = false;
    if (mode == 1) {
      if (fmax(awkk, colmax) == 0.0 || isnan(awkk)) {
        step = 1;
        kpiv = k;
        zero_col = true;
      } else if (awkk >= alpha * colmax) {
        step = 1;
        kpiv = k;
      } else {
        need_row = true;
      }
    } else {
      if (awkk >= ptol && awkk >= ustab * colmax) {
        step = 1;
        kpiv = k;
      } else if (k + 1 < nc) {
        need_row = true;
      }
    }
    double wii = 0.0, wik = 0.0;
    if (need_row) {
      // column imax of the trailing matrix, rows k .. m-1 (row imax excluded
      // from the maximum)
      double lrow = 0.0;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int i = min(k + lane + kWave * h, m - 1);
        const double v = i < imax ? F[imax + (size_t)i * m] : F[i + (size_t)imax * m];
        if (k + lane + kWave * h < m && i != imax) lrow = fmax(lrow, fabs(v));
      }
      const double rowmax = wave_max(lrow);
      wii = F[imax + (size_t)imax * m];
      wik = F[imax + (size_t)k * m];
      const double awii = fabs(wii);
      if (mode == 1) {
        if (awkk >= alpha * colmax * (colmax / rowmax)) {
          step = 1;
          kpiv = k;
        } else if (awii >= alpha * rowmax) {
          step = 1;
          kpiv = imax;
        } else {
          step = 2;
          kpiv = imax;
        }
      } else {
        if (awii >= ptol && awii >= ustab * rowmax) {
          step = 1;
          kpiv = imax;
        } else {
          double det = fabs(wkk * wii - wik * wik);
          if (det >= ptol) {
            double delta = det / fmax(awii + awik, awkk + awik);
            if (delta >= ustab * colmax && delta >= ustab * rowmax) {
              step = 2;
              kpiv = imax;
            }
          }
        }
      }
    }
    if (step == 0) {  // delay candidate k (threshold mode only)
      if (k < nc - 1) {
        __syncthreads();
        sym_swap<kRLThreads, double>(F, m, m, vars, pre, k, nc - 1);
        __syncthreads();
      }
      nc -= 1;
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) {
      __syncthreads();
      sym_swap<kRLThreads, double>(F, m, m, vars, pre, ksrc, kpiv);
      __syncthreads();
    }
    // the pivot (after the swap, column k holds it; for 2x2 the partner is
    // column k + 1)
    const int j0 = k + step;
    const int ntri = (m - j0) * (m - j0 + 1) / 2;
    if (step == 1) {
      const double d = kpiv == k ? wkk : wii;
      const double dinv = (zero_col || d == 0.0) ? 0.0 : 1.0 / d;
      for (int r0 = 0; r0 < kRLEntries; r0 += kRLBatch) {
        if (tid + kRLThreads * r0 >= ntri) break;
        // every load of the batch first (clamped indices), then the stores
        int ij[kRLBatch];
        double f[kRLBatch], wi[kRLBatch], wj[kRLBatch];
#pragma unroll
        for (int u = 0; u < kRLBatch; u++) {
          const int r = min(r0 + u, kRLEntries - 1);
          const bool ok = tid + kRLThreads * (r0 + u) < ntri && r0 + u < kRLEntries;
          const int i = ok ? m - 1 - (te.pk[r] & 255) : m - 1;
          const int j = ok ? m - 1 - (te.pk[r] >> 8) : m - 1;
          ij[u] = ok ? i + j * m : -1;
          f[u] = F[i + j * m];
          wi[u] = F[i + k * m];
          wj[u] = F[j + k * m];
        }
#pragma unroll
        for (int u = 0; u < kRLBatch; u++) {
          const double li = round_t<T>(zero_col ? wi[u] : wi[u] * dinv);
          if (ij[u] >= 0) F[ij[u]] = f[u] - li * wj[u];
        }
      }
      if (tid == 0) {
        F[k + (size_t)k * m] = round_t<T>(d);
        ptype[k] = 1;
        dv[k] = d;
      }
      if (d >= 0.0) npos++;
      else nneg++;
      pm0 = dinv;
      pzero = zero_col;
    } else {
      const double d11 = wkk, d21 = wik, d22 = wii;
      const double det = d11 * d22 - d21 * d21;
      const double detinv = 1.0 / det;
      const double i00 = d22 * detinv, i01 = -d21 * detinv, i11 = d11 * detinv;
      for (int r0 = 0; r0 < kRLEntries; r0 += kRLBatch) {
        if (tid + kRLThreads * r0 >= ntri) break;
        int ij[kRLBatch];
        double f[kRLBatch], ai[kRLBatch], bi[kRLBatch], aj[kRLBatch], bj[kRLBatch];
#pragma unroll
        for (int u = 0; u < kRLBatch; u++) {
          const int r = min(r0 + u, kRLEntries - 1);
          const bool ok = tid + kRLThreads * (r0 + u) < ntri && r0 + u < kRLEntries;
          const int i = ok ? m - 1 - (te.pk[r] & 255) : m - 1;
          const int j = ok ? m - 1 - (te.pk[r] >> 8) : m - 1;
          ij[u] = ok ? i + j * m : -1;
          f[u] = F[i + j * m];
          ai[u] = F[i + k * m];
          bi[u] = F[i + (k + 1) * m];
          aj[u] = F[j + k * m];
          bj[u] = F[j + (k + 1) * m];
        }
#pragma unroll
        for (int u = 0; u < kRLBatch; u++) {
          const double l1 = round_t<T>(ai[u] * i00 + bi[u] * i01);
          const double l2 = round_t<T>(ai[u] * i01 + bi[u] * i11);
          double v = f[u];
          v -= l1 * aj[u];
          v -= l2 * bj[u];
          if (ij[u] >= 0) F[ij[u]] = v;
        }
      }
      if (tid == 0) {
        F[k + (size_t)k * m] = round_t<T>(d11);
        F[k + 1 + (size_t)k * m] = round_t<T>(d21);
        F[k + 1 + (size_t)(k + 1) * m] = round_t<T>(d22);
        dv[k] = d11;
        dv2[k] = d21;
        dv[k + 1] = d22;
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
        ptype[k] = 2;
        ptype[k + 1] = 3;
      }
      if (det >= 0.0) {
        if (d11 >= 0.0) npos += 2;
        else nneg += 2;
      } else {
        npos++;
        nneg++;
      }
      n2++;
      pm0 = i00;
      pm1 = i01;
      pm2 = i11;
      pzero = false;
    }
    pk = k;
    pstep = step;
    k += step;
    __syncthreads();
  }
  finalize();
  __syncthreads();
  // ---- write back
  T* Fg = front_ptr<T>(P, p);
  for (int e = tid; e < m * m; e += kRLThreads) Fg[e] = (T)F[e];
  const long long io = P.ioff[p];
  for (int i = tid; i < m; i += kRLThreads) {
    P.vars[io + i] = vars[i];
    P.pre[io + i] = pre[i];
    P.ptype[io + i] = ptype[i];
    P.dv[io + i] = dv[i];
    P.dv2[io + i] = dv2[i];
  }
  if (tid == 0) {
    P.npiv[p] = nc;
    P.ndel[p] = fs - nc;
    P.cnt[3 * p] = npos;
    P.cnt[3 * p + 1] = nneg;
    P.cnt[3 * p + 2] = n2;
  }
}

template <class T, int MF>
__global__ void __launch_bounds__(kRLThreads) k_front_rl(DevPlan P, const int* nodes) {
  extern __shared__ double lds[];
  __shared__ RLShared sh;
  TriEntries<MF> te;
  te.init(threadIdx.x);
  front_rl<T, MF>(P, nodes[blockIdx.x], lds, te, sh);
}

// The fronts of `nodes` one after another (each front's children are
// earlier in the list or finished before the launch).  A front's results
// reach the next front through global memory written by this workgroup:
// the barrier at the end of each front orders them (workgroup scope, one
// CU, whose vector L1 is write-through).
template <class T>
__global__ void __launch_bounds__(kRLThreads) k_front_chain(DevPlan P, const int* nodes,
                                                            int count) {
  extern __shared__ double lds[];
  __shared__ RLShared sh;
  TriEntries<kRLMaxFront> te;
  te.init(threadIdx.x);
  for (int q = 0; q < count; q++) {
    front_rl<T, kRLMaxFront>(P, nodes[q], lds, te, sh);
    __syncthreads();
  }
}

size_t rl_front_bytes(int c) {
  return sizeof(double) * ((size_t)c * c + 2 * (size_t)c) + sizeof(int) * 3 * (size_t)c;
}

int rl_max_front() { return kRLMaxFront; }

template <class T>
static void launch_rl(int cap, int count, bool chain, bool sized, hipStream_t st,
                      const DevPlan& P, const int* nodes) {
  if (cap > kRLMaxFront) throw Error(-2, "LDS front above the right-looking kernel's bound");
  const size_t bytes = rl_front_bytes(cap);
  if (chain)
    hipLaunchKernelGGL(k_front_chain<T>, dim3(1), dim3(kRLThreads), bytes, st, P, nodes, count);
  else if (!sized)
    hipLaunchKernelGGL((k_front_rl<T, kRLMaxFront>), dim3(count), dim3(kRLThreads), bytes, st, P,
                       nodes);
  else if (cap <= 32)
    hipLaunchKernelGGL((k_front_rl<T, 32>), dim3(count), dim3(kRLThreads), bytes, st, P, nodes);
  else if (cap <= 48)
    hipLaunchKernelGGL((k_front_rl<T, 48>), dim3(count), dim3(kRLThreads), bytes, st, P, nodes);
  else if (cap <= 64)
    hipLaunchKernelGGL((k_front_rl<T, 64>), dim3(count), dim3(kRLThreads), bytes, st, P, nodes);
  else if (cap <= 80)
    hipLaunchKernelGGL((k_front_rl<T, 80>), dim3(count), dim3(kRLThreads), bytes, st, P, nodes);
  else
    hipLaunchKernelGGL((k_front_rl<T, kRLMaxFront>), dim3(count), dim3(kRLThreads), bytes, st, P,
                       nodes);
}

void launch_front_rl(const float*, int cap, int count, bool chain, bool sized, hipStream_t st,
                     const DevPlan& P, const int* nodes) {
  launch_rl<float>(cap, count, chain, sized, st, P, nodes);
}
void launch_front_rl(const double*, int cap, int count, bool chain, bool sized, hipStream_t st,
                     const DevPlan& P, const int* nodes) {
  launch_rl<double>(cap, count, chain, sized, st, P, nodes);
}

}  // namespace kktldl
