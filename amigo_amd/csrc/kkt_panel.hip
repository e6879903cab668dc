// Large-front panel kernels of the MI355X LDL^T engine: k_panel_rr (the
// panel's L in LDS) and k_panel (global-memory panel for fronts whose panel
// does not fit LDS).  Built once per front element type: -DKKT_PANEL_T=float
// (FP32 fronts) and -DKKT_PANEL_T=double (FP64 fronts).
#include "kkt_common.hpp"

#ifndef KKT_PANEL_T
#error "build with -DKKT_PANEL_T=float or -DKKT_PANEL_T=double"
#endif

namespace kktldl {

// --------------------------------------------------------------------------
// Large fronts: one panel per launch on one workgroup (state in P.ps), the
// trailing update spread over many workgroups (k_update).
// --------------------------------------------------------------------------
// GV: the front's vars and pre stay in HBM (fronts whose bookkeeping does
// not fit LDS); otherwise they are staged in LDS for the panel.
template <int NT, class T, bool GV>
__global__ void __launch_bounds__(NT) k_panel(DevPlan P, const int* nodes) {
  extern __shared__ int s_iv[];  // vars and pre of the front, staged in LDS
  const int p = nodes[blockIdx.x];
  if (P.stat[p]) return;
  __shared__ FrontShared<NT> sh;
  __shared__ PanelState s_st;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_st = P.ps[p];
    sh.npos = s_st.npos;
    sh.nneg = s_st.nneg;
    sh.n2 = s_st.n2;
  }
  __syncthreads();
  int k = s_st.k, nc = s_st.nc, kw = 0;
  const int kstart = k;
  // a root front in mode 3 turns to Bunch-Kaufman once its threshold phase
  // has left only delayed columns (the phase flag is kept in the state)
  int bkphase = s_st.pad;
  if (P.mode[p] == 3 && !bkphase && k >= nc && nc < P.m[p]) {
    nc = P.m[p];
    bkphase = 1;
  }
  if (k < nc) {
    const int m = P.m[p];
    const int mode = P.mode[p] == 3 ? bkphase : P.mode[p];
    const int nb = max(2, min(kNBLarge, m));
    const long long io = P.ioff[p];
    int* vars = GV ? P.vars + io : s_iv;
    int* pre = GV ? P.pre + io : s_iv + m;
    if (!GV)
      for (int i = k + tid; i < m; i += NT) {
        vars[i] = P.vars[io + i];
        pre[i] = P.pre[io + i];
      }
    __syncthreads();
    panel_factor<NT, T>(front_ptr<T>(P, p), m, m, vars, pre, P.ptype + io, P.dv + io,
                        P.dv2 + io, P.W + P.woff[p], P.cap[p], nb, mode, P.ustab, P.ptol,
                        k, nc, kw, sh);
    if (!GV)
      for (int i = kstart + tid; i < m; i += NT) {
        P.vars[io + i] = vars[i];
        P.pre[io + i] = pre[i];
      }
    if (tid == 0 && k >= nc) {
      const int fs = P.ns[p] + P.din[p];
      P.npiv[p] = nc;
      P.ndel[p] = fs - nc;
      P.cnt[3 * p] = sh.npos;
      P.cnt[3 * p + 1] = sh.nneg;
      P.cnt[3 * p + 2] = sh.n2;
    }
  }
  if (tid == 0) {
    PanelState st = s_st;
    st.k = k;
    st.nc = nc;
    st.kstart = kstart;
    st.kw = kw;  // 0 when the front was already complete: no update
    st.npos = sh.npos;
    st.nneg = sh.nneg;
    st.n2 = sh.n2;
    st.pad = bkphase;
    P.ps[p] = st;
  }
}

// --------------------------------------------------------------------------
// Large-front panel with the panel's L held in LDS (fronts up to R * NT rows).
//
// The panel factorization of a large front runs on one workgroup of NT
// threads; thread t owns rows kstart + t + j * NT (j < R).  The L entries of
// the panel columns live in LDS for the whole panel, so updating front
// column k by the panel so far, W[i, kw] = F[i, k] - sum_c L[i, c] W[k, c],
// reads only the column from HBM (prefetched two columns ahead into
// registers).  The W entries of the positions that can become pivot rows in
// this panel (kstart .. kstart + 17) are cached in LDS by their owners in the
// same barrier interval as the column reduction, and every thread reads the
// pivot row from there.  W is written to HBM as it is produced (the trailing
// update reads it) and L once, at the end of the panel.  The barriers inside
// a pivot step order LDS only, so HBM stores and prefetch loads stay in
// flight across them.  A 1x1 pivot costs one barrier and a 2x2 pivot whose
// partner is the next candidate (imax == k+1, the common case in KKT fronts)
// two; pivots that need a symmetric swap, and delays, take a slower path
// that swaps in HBM and reloads the prefetch window.  The pivot decisions are
// panel_factor's (mode 0 threshold with delays, mode 1 Bunch-Kaufman) on the
// same FP64 values; L entries are formed with explicit FMAs.
// --------------------------------------------------------------------------

// BKC: the launch may hold a root front (Bunch-Kaufman tests, mode 1); the
// instances without it carry no Bunch-Kaufman code, so that the kernel the
// other levels run is smaller.
template <int NT, int R, class T, bool BKC>
__global__ void __launch_bounds__(NT) k_panel_rr(DevPlan P, const int* nodes) {
  constexpr int NB = kNBLarge;
  constexpr int NW = NT / kWave;
  constexpr int RS = R * NT;        // rows of a panel column in LDS
  extern __shared__ double s_rr[];  // Ls[NB][RS] of T, then vars[m], pre[m]
  __shared__ RRShared<NW> sh;
  const int p = nodes[blockIdx.x];
  // everything per node depends on p alone: one round of loads
  const int stat = P.stat[p];
  const PanelState st0 = P.ps[p];
  const int m = P.m[p];
  const int mode0 = P.mode[p];
  const long long io = P.ioff[p];
  T* F = front_ptr<T>(P, p);
  double* Wg = P.W + P.woff[p];
  const int ldw = P.cap[p];
  if (stat) return;
  const int tid = threadIdx.x;
  const int wv = tid / kWave, lane = tid % kWave;
  if (tid == 0) {
    sh.npos = st0.npos;
    sh.nneg = st0.nneg;
    sh.n2 = st0.n2;
  }
  int k = st0.k, nc = st0.nc, kw = 0;
  int kw_spec = 0;  // columns whose L the speculation stored
  // mode 3 (a large root front): threshold tests with delays while
  // candidates remain, then Bunch-Kaufman on the delayed columns; the switch
  // is made at the start of a panel and kept in the state
  int bkphase = st0.pad;
  if (mode0 == 3 && !bkphase && k >= nc && nc < m) {
    nc = m;
    bkphase = 1;
  }
  const int mode = mode0 == 3 ? bkphase : mode0;
  const int kstart = k;
  if (k < nc) {
    const int nb = max(2, min(NB, m));
    const int ld = m;
    const double ustab = P.ustab, ptol = P.ptol;
    const double alpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8
    T* Ls = (T*)s_rr;
    int* vars = (int*)(Ls + (size_t)NB * RS);
    int* pre = vars + m;
    auto ld_index = [&]() {
      for (int i = k + tid; i < m; i += NT) {
        vars[i] = P.vars[io + i];
        pre[i] = P.pre[io + i];
      }
    };
    auto rowof = [&](int j) { return kstart + tid + j * NT; };
    auto lsx = [&](int c, int j) -> T& { return Ls[(size_t)c * RS + j * NT + tid]; };
    auto lsr = [&](int c, int r) -> T& { return Ls[(size_t)c * RS + (r - kstart)]; };
    // front columns k .. k+3 of the owned rows: c0, c1 ready, f2, f3 in
    // flight (loads are unconditional, rows and columns clamped into the
    // front, so that the compiler waits only where a value is first used)
    T c0[R], c1[R], f2[R], f3[R];
    auto ld_col = [&](int col, T (&t)[R]) {
      const size_t co = (size_t)min(col, m - 1) * ld;
#pragma unroll
      for (int j = 0; j < R; j++) t[j] = F[min(rowof(j), m - 1) + co];
    };
    auto reload = [&]() {
      ld_col(k, c0);
      ld_col(k + 1, c1);
      ld_col(k + 2, f2);
      ld_col(k + 3, f3);
    };
    // Speculative panel (threshold pivoting).  The pivots are chosen by wave
    // 0 from its rows alone (positions kstart .. kstart + 63, one per lane):
    // a 1x1 pivot, or a 2x2 pivot with the next position, whenever the tests
    // of the sequential loop below pass against the three rows after the
    // diagonal.  Every row then forms its W and L entries for those choices
    // (the same FMA chains as the sequential loop, so the same bits) and
    // checks them: the choice is the sequential loop's exactly when no row
    // exceeds the bound it implies (|W| <= |d| / ustab for a 1x1 pivot; the
    // pivot row's entry the largest candidate and delta / ustab the bound for
    // a 2x2 pivot; a NaN fails).  Both bounds only grow with more rows, so a
    // test that fails on the leading rows fails on the front.  The first
    // choice that some row contradicts, and everything after it, is left to
    // the sequential loop.  One barrier for the choices and one for the
    // checks replace the per-pivot reductions and barriers.
    const bool spec = P.spec && mode != 2;
    const bool bk = BKC && mode == 1;  // Bunch-Kaufman tests (the root front)
    if (spec) {
      T fin[NB][R];  // front columns kstart .. kstart + NB - 1 of the owned rows
#pragma unroll
      for (int c = 0; c < NB; c++) {
        const size_t co = (size_t)min(kstart + c, m - 1) * ld;
#pragma unroll
        for (int j = 0; j < R; j++) fin[c][j] = F[min(rowof(j), m - 1) + co];
      }
      ld_index();
      if (tid == 0) sh.sp_fail = 0u;
      if (wv == 0) {
        const int r = kstart + lane;
        const bool rin = r < m;
        double xs[NB];  // W columns of the choices made so far, at row r
#pragma unroll
        for (int c = 0; c < NB; c++) xs[c] = 0.0;
        int ns = 0, next = 0;
        bool go = true;
#pragma unroll
        for (int cc = 0; cc < NB - 1; cc++) {
          if (!go || cc != next) continue;
          const int kp = kstart + cc;
          if (cc >= nb - 1 || kp >= nc) {
            go = false;
            continue;
          }
          double x = (double)fin[cc][0];
#pragma unroll
          for (int c = 0; c < cc; c++) x = __builtin_fma(-(double)fin[c][0], lane_d(xs[c], cc), x);
          const double wkk = lane_d(x, cc), awkk = fabs(wkk);
          double lcol = 0.0;
#pragma unroll
          for (int t = 1; t <= 3; t++)
            if (kp + t < m) lcol = fmax(lcol, fabs(lane_d(x, cc + t)));
          int type = 0;
          double thr = 0.0, awik = 0.0, wik = 0.0, wii = 0.0, y = 0.0;
          if (bk ? (awkk > 0.0 && awkk >= alpha * lcol) : (awkk >= ptol && awkk >= ustab * lcol)) {
            type = 1;
            thr = awkk;
          } else if (kp + 1 < nc) {
            wik = lane_d(x, cc + 1);
            awik = fabs(wik);
            bool adj = awik >= 0.0;  // false for NaN
#pragma unroll
            for (int t = 2; t <= 3; t++)
              if (kp + t < nc && fabs(lane_d(x, cc + t)) > awik) adj = false;
            if (adj) {
              // column k+1 updated; its row-k entry is F[k+1, k] with the
              // update chain of row k (as in the sequential loop)
              const double f1 = lane_d((double)fin[cc][0], cc + 1);
              double yv = lane == cc ? f1 : (double)fin[cc + 1][0];
#pragma unroll
              for (int c = 0; c < cc; c++)
                yv = __builtin_fma(-(double)fin[c][0], lane_d(xs[c], cc + 1), yv);
              wii = lane_d(yv, cc + 1);
              double lrow = fabs(lane_d(yv, cc));
#pragma unroll
              for (int t = 2; t <= 3; t++)
                if (kp + t < m) lrow = fmax(lrow, fabs(lane_d(yv, cc + t)));
              const double awii = fabs(wii);
              if (bk) {
                // the second test passes with the leading rows' row maximum,
                // so with the front's (no smaller): 1x1 at k; else a 2x2
                // pivot when the third test fails (a 1x1 pivot at the partner
                // or a zero column is left to the sequential loop)
                if (awik > 0.0 && awkk == awkk) {
                  if (awkk >= alpha * awik * (awik / lrow)) {
                    type = 1;
                    thr = -1.0;  // verified through the partner's maximality alone
                  } else if (!(awii >= alpha * lrow)) {
                    type = 2;
                    y = yv;
                  }
                }
              } else if (!(awii >= ptol && awii >= ustab * lrow)) {
                const double det = fabs(piv_det(wkk, wii, wik));
                if (det >= ptol) {
                  const double delta = piv_delta(det, awkk, awii, awik);
                  if (delta >= ustab * lcol && delta >= ustab * lrow) {
                    type = 2;
                    thr = delta;
                    y = yv;
                  }
                }
              }
            }
          }
          // this wave's rows against the bounds of the choice
          bool bad = false;
          if (rin && type == 1 && r > kp)
            bad = thr < 0.0 ? (r > kp + 1 && !(fabs(x) <= awik))
                            : !((bk ? alpha : ustab) * fabs(x) <= thr);
          if (rin && type == 2 && bk && r >= kp && r != kp + 1) {
            // the second test must fail with this row's entry as the row maximum
            bad = !(awkk < alpha * awik * (awik / fabs(y)));
            if (r > kp + 1) bad = bad || !(fabs(x) <= awik);
          }
          if (rin && type == 2 && !bk && r > kp) {
            bad = !(ustab * fabs(x) <= thr);
            if (r > kp + 1) {
              bad = bad || !(ustab * fabs(y) <= thr);
              if (r < nc) bad = bad || !(fabs(x) <= awik);
            }
          }
          if (type == 0 || __ballot(bad) != 0ull) {
            go = false;
            continue;
          }
          if (type == 1) {
            const double dinv = wkk == 0.0 ? 0.0 : 1.0 / wkk;
            if (rin && r >= kp) {
              const T l = (T)(x * dinv);
              fin[cc][0] = l;
              lsx(cc, 0) = l;
              Wg[r + (size_t)cc * ldw] = x;
              if (lane < kWCRows) sh.wc[lane][cc] = x;
            }
            if (lane == 0) {
              sh.sp_m[ns][0] = dinv;
              sh.sp_d[ns][0] = wkk;
            }
          } else {
            double i00, i01, i11;
            piv_inv2(wkk, wik, wii, i00, i01, i11);
            if (rin && r >= kp) {
              const T l0 = (T)__builtin_fma(x, i00, y * i01);
              const T l1 = (T)__builtin_fma(x, i01, y * i11);
              fin[cc][0] = l0;
              fin[cc + 1][0] = l1;
              lsx(cc, 0) = l0;
              lsx(cc + 1, 0) = l1;
              Wg[r + (size_t)cc * ldw] = x;
              Wg[r + (size_t)(cc + 1) * ldw] = y;
              if (lane < kWCRows) {
                sh.wc[lane][cc] = x;
                sh.wc[lane][cc + 1] = y;
              }
            }
            if (lane == 0) {
              sh.sp_m[ns][0] = i00;
              sh.sp_m[ns][1] = i01;
              sh.sp_m[ns][2] = i11;
              sh.sp_d[ns][0] = wkk;
              sh.sp_d[ns][1] = wik;
              sh.sp_d[ns][2] = wii;
            }
            xs[cc + 1] = y;
          }
          xs[cc] = x;
          if (lane == 0) {
            sh.sp_c[ns] = cc;
            sh.sp_t[ns] = type;
            sh.sp_thr[ns] = thr;
            sh.sp_awik[ns] = awik;
          }
          ns++;
          next = cc + type;
        }
        if (lane == 0) {
          sh.sp_n = ns;
          sh.sp_end = next;
        }
      }
      lds_sync();
      // the other rows (all below position kstart + 63, so below every
      // pivot of the panel): W and L entries of the choices, and the checks
      const int nsp = sh.sp_n;
      unsigned fail = 0u;
      {
        int s = 0, next = 0;
#pragma unroll
        for (int cc = 0; cc < NB - 1; cc++) {
          if (s >= nsp || cc != next) continue;
          const int type = sh.sp_t[s];
          const double thr = sh.sp_thr[s];
          double w0[NB], w1[NB];
#pragma unroll
          for (int c = 0; c < cc; c++) {
            w0[c] = sh.wc[cc][c];
            w1[c] = sh.wc[cc + 1][c];
          }
          if (type == 1) {
            const double dinv = sh.sp_m[s][0];
#pragma unroll
            for (int j = 0; j < R; j++) {
              const int r = rowof(j);
              if ((wv == 0 && j == 0) || r >= m) continue;
              double x = (double)fin[cc][j];
#pragma unroll
              for (int c = 0; c < cc; c++) x = __builtin_fma(-(double)fin[c][j], w0[c], x);
              const T l = (T)(x * dinv);
              fin[cc][j] = l;
              lsx(cc, j) = l;
              Wg[r + (size_t)cc * ldw] = x;
              const bool ok = thr < 0.0 ? fabs(x) <= sh.sp_awik[s] : (bk ? alpha : ustab) * fabs(x) <= thr;
              if (!ok) fail |= 1u << s;
            }
          } else {
            const double awik = sh.sp_awik[s], awkk = fabs(sh.sp_d[s][0]);
            const double i00 = sh.sp_m[s][0], i01 = sh.sp_m[s][1], i11 = sh.sp_m[s][2];
#pragma unroll
            for (int j = 0; j < R; j++) {
              const int r = rowof(j);
              if ((wv == 0 && j == 0) || r >= m) continue;
              double x = (double)fin[cc][j], y = (double)fin[cc + 1][j];
#pragma unroll
              for (int c = 0; c < cc; c++) {
                x = __builtin_fma(-(double)fin[c][j], w0[c], x);
                y = __builtin_fma(-(double)fin[c][j], w1[c], y);
              }
              const T l0 = (T)__builtin_fma(x, i00, y * i01);
              const T l1 = (T)__builtin_fma(x, i01, y * i11);
              fin[cc][j] = l0;
              fin[cc + 1][j] = l1;
              lsx(cc, j) = l0;
              lsx(cc + 1, j) = l1;
              Wg[r + (size_t)cc * ldw] = x;
              Wg[r + (size_t)(cc + 1) * ldw] = y;
              const bool ok =
                  bk ? (fabs(x) <= awik && awkk < alpha * awik * (awik / fabs(y)))
                     : (ustab * fabs(x) <= thr && ustab * fabs(y) <= thr &&
                        !(r < nc && !(fabs(x) <= awik)));
              if (!ok) fail |= 1u << s;
            }
          }
          s++;
          next = cc + type;
        }
      }
      if (fail) atomicOr(&sh.sp_fail, fail);
      lds_sync();
      const unsigned fm = sh.sp_fail;
      const int nok = fm ? min(nsp, (int)__builtin_ctz(fm)) : nsp;
      kw = nok < nsp ? sh.sp_c[nok] : sh.sp_end;
      k = kstart + kw;
      kw_spec = kw;
      // the accepted columns' L entries are final: stored now (ordered
      // before the sequential loop by its full barrier, so that its HBM
      // swaps move them as they move the rest of the rows), not at the end
      for (int q = 0; q < nok; q++) {
        const int c = sh.sp_c[q], t = sh.sp_t[q], pos = kstart + c;
#pragma unroll
        for (int j = 0; j < R; j++) {
          const int r = rowof(j);
          if (r > pos + t - 1 && r < m) {
            F[r + (size_t)pos * ld] = lsx(c, j);
            if (t == 2) F[r + (size_t)(pos + 1) * ld] = lsx(c + 1, j);
          }
        }
      }
      // the accepted steps' D entries, pivot types and inertia, one step per
      // thread (the counts are sums)
      if (tid < nok) {
        const int s = tid;
        const int c = sh.sp_c[s], kp = kstart + c;
        if (sh.sp_t[s] == 1) {
          const double d = sh.sp_d[s][0];
          F[kp + (size_t)kp * ld] = (T)d;
          P.ptype[io + kp] = 1;
          P.dv[io + kp] = d;
          atomicAdd(d >= 0.0 ? &sh.npos : &sh.nneg, 1);
          sh.cT[c] = 1;
        } else {
          const double d11 = sh.sp_d[s][0], d21 = sh.sp_d[s][1], d22 = sh.sp_d[s][2];
          F[kp + (size_t)kp * ld] = (T)d11;
          F[kp + 1 + (size_t)kp * ld] = (T)d21;
          F[kp + 1 + (size_t)(kp + 1) * ld] = (T)d22;
          P.dv[io + kp] = d11;
          P.dv2[io + kp] = d21;
          P.dv[io + kp + 1] = d22;
          vars[kp] = -(vars[kp] + 1);
          vars[kp + 1] = -(vars[kp + 1] + 1);
          P.ptype[io + kp] = 2;
          P.ptype[io + kp + 1] = 3;
          if (piv_det(d11, d22, d21) >= 0.0) atomicAdd(d11 >= 0.0 ? &sh.npos : &sh.nneg, 2);
          else {
            atomicAdd(&sh.npos, 1);
            atomicAdd(&sh.nneg, 1);
          }
          atomicAdd(&sh.n2, 1);
          sh.cT[c] = 2;
          sh.cT[c + 1] = 3;
        }
      }
    } else {
      ld_index();
    }
    // the sequential loop, for what the speculation left
    if (kw < nb - 1 && k < nc) reload();
    __syncthreads();
    int it = 0;
    // Look-ahead: each step also forms the updates of columns k+1 and k+2 by
    // the panel's earlier columns (la1, la2) while its reductions run, so
    // that after a fast 1x1 / adjacent 2x2 pivot the next column is one / two
    // FMAs away (vn); the FMA chain and its order are those of the direct
    // update, so the values are the same bits.
    double vn[R];
    bool la_ok = false;
    while (kw < nb - 1 && k < nc) {
      const int par = it & 1;
      // W[:, kw]: column k updated by the panel so far
      const double* wr = sh.wc[k - kstart];
      double v[R], vp[R];
      if (la_ok) {
#pragma unroll
        for (int j = 0; j < R; j++) {
          v[j] = vn[j];
          vp[j] = 0.0;
        }
      } else {
#pragma unroll
        for (int j = 0; j < R; j++) {
          v[j] = (double)c0[j];
          vp[j] = 0.0;
        }
        for (int c = 0; c < kw; c++) {
          const double w = wr[c];
#pragma unroll
          for (int j = 0; j < R; j++) v[j] = __builtin_fma(-(double)lsx(c, j), w, v[j]);
        }
      }
      double lbest = -1.0, lrest = 0.0;
      int lidx = 0x7fffffff;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int r = rowof(j);
        if (r >= k && r < m) {
          const double x = v[j];
          const double a = fabs(x);
          if (r > k && r < nc) {
            if (a > lbest) {
              lbest = a;
              lidx = r;
            }
          } else if (r >= nc) {
            lrest = fmax(lrest, a);
          }
          if (r == k) sh.wkk[par] = x;
          if (r == k + 1) {
            sh.fk1k[par] = (double)c0[j];
            sh.wik[par] = x;
          }
          if (r - kstart < kWCRows) sh.wc[r - kstart][kw] = x;
        }
      }
      // columns k+1, k+2 updated by the panel columns before kw (W rows of
      // positions k+1, k+2 are cached; k + 2 - kstart <= 16 < kWCRows)
      double la1[R], la2[R];
      {
        const double* w1 = sh.wc[k + 1 - kstart];
        const double* w2 = sh.wc[k + 2 - kstart];
#pragma unroll
        for (int j = 0; j < R; j++) {
          la1[j] = (double)c1[j];
          la2[j] = (double)f2[j];
        }
        for (int c = 0; c < kw; c++) {
          const double a = w1[c], b = w2[c];
#pragma unroll
          for (int j = 0; j < R; j++) {
            const double l = -(double)lsx(c, j);
            la1[j] = __builtin_fma(l, a, la1[j]);
            la2[j] = __builtin_fma(l, b, la2[j]);
          }
        }
      }
      wave_argmax2(lbest, lidx, lrest);
      if (lane == 0) {
        sh.pb[par][wv] = lbest;
        sh.pi[par][wv] = lidx;
        sh.pr[par][wv] = lrest;
      }
      lds_sync();
      {
        double bb = -1.0, rr = 0.0;
        int ii = 0x7fffffff;
        if (lane < NW) {
          bb = sh.pb[par][lane];
          ii = sh.pi[par][lane];
          rr = sh.pr[par][lane];
        }
        row_argmax2(bb, ii, rr);
        lbest = lane_d(bb, 0);
        lidx = __builtin_amdgcn_readlane(ii, 0);
        lrest = lane_d(rr, 0);
      }
      const double wkk = sh.wkk[par];
      const double awkk = fabs(wkk);
      int imax = 0;
      double awik = 0.0;
      if (k + 1 < nc) {
        // lidx keeps its sentinel only when every candidate entry is NaN:
        // take the next candidate, so that every index stays in the front
        const bool found = lidx < nc;
        imax = found ? lidx : k + 1;
        awik = found ? lbest : __builtin_nan("");
      }
      const double colmax = fmax(awik, lrest);
      int step = 0, kpiv = -1;
      bool zero_col = false, need_row = false;
      if (BKC && mode == 1) {
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
        double lrow = 0.0;
        const double* wr2;
        if (imax == k + 1) {
          // the imax column is column k+1, already updated (la1); its entry
          // in row k is F[k+1, k], published with the column reduction, with
          // the update chain of row k
          wr2 = sh.wc[k + 1 - kstart];
#pragma unroll
          for (int j = 0; j < R; j++) {
            vp[j] = la1[j];
            if (rowof(j) == k) {
              double t = sh.fk1k[par];
              for (int c = 0; c < kw; c++) t = __builtin_fma(-(double)lsx(c, j), wr2[c], t);
              vp[j] = t;
            }
          }
        } else {
          // general imax: its W row from the LDS cache or HBM (written by
          // its owner), its column (row imax left of the diagonal) from the
          // front
          const bool wcached = imax - kstart < kWCRows;
#pragma unroll
          for (int j = 0; j < R; j++)
            if (rowof(j) == imax) {
              if (!wcached)
                for (int c = 0; c < kw; c++) sh.xw[c] = Wg[imax + (size_t)c * ldw];
              sh.swik = v[j];
            }
          // all R loads in flight before the barrier; converted after it
          T raw[R];
#pragma unroll
          for (int j = 0; j < R; j++) {
            const int r = min(rowof(j), m - 1);
            raw[j] = r < imax ? F[imax + (size_t)r * ld] : F[r + (size_t)imax * ld];
          }
          __builtin_amdgcn_sched_barrier(0);
          lds_sync();
#pragma unroll
          for (int j = 0; j < R; j++) vp[j] = (double)raw[j];
          wr2 = wcached ? sh.wc[imax - kstart] : sh.xw;
          for (int c = 0; c < kw; c++) {
            const double w = wr2[c];
#pragma unroll
            for (int j = 0; j < R; j++) vp[j] = __builtin_fma(-(double)lsx(c, j), w, vp[j]);
          }
        }
#pragma unroll
        for (int j = 0; j < R; j++) {
          const int r = rowof(j);
          if (r >= k && r < m) {
            const double x = vp[j];
            if (r != imax) lrow = fmax(lrow, fabs(x));
            if (r == imax) sh.wii[par] = x;
            if (imax == k + 1 && r - kstart < kWCRows) sh.wc[r - kstart][kw + 1] = x;
          }
        }
        lrow = wave_max(lrow);
        if (lane == 0) sh.qr[par][wv] = lrow;
        lds_sync();
        wii = sh.wii[par];
        wik = (imax == k + 1) ? sh.wik[par] : sh.swik;
        const double rowmax = lane_d(row_max(lane < NW ? sh.qr[par][lane] : 0.0), 0);
        const double awii = fabs(wii);
        if (BKC && mode == 1) {
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
            const double det = fabs(piv_det(wkk, wii, wik));
            if (det >= ptol) {
              const double delta = piv_delta(det, awkk, awii, awik);
              if (delta >= ustab * colmax && delta >= ustab * rowmax) {
                step = 2;
                kpiv = imax;
              }
            }
          }
        }
      }
      // the decision is the same in every lane; keep it in scalar registers
      step = __builtin_amdgcn_readfirstlane(step);
      kpiv = __builtin_amdgcn_readfirstlane(kpiv);
      const bool fast = step > 0 && kpiv == k + step - 1;
      // a swap of two positions whose W rows are in the LDS cache is done
      // there (and written back), without a round trip through HBM
      bool wswap = false;
      if (!fast) {
        // delay (step 0), 1x1 at imax, or 2x2 whose partner is not k+1
        const int a = (step == 0) ? k : k + step - 1;
        const int b = (step == 0) ? nc - 1 : kpiv;
        const int cnt = kw + step;
        wswap = a != b && b - kstart < kWCRows;
        if (step == 1) {
#pragma unroll
          for (int j = 0; j < R; j++) v[j] = vp[j];  // the pivot column is imax
        }
        if (wswap) {
          for (int c = tid; c < kw; c += NT) {
            const double ta = sh.wc[a - kstart][c], tb = sh.wc[b - kstart][c];
            sh.wc[a - kstart][c] = tb;
            sh.wc[b - kstart][c] = ta;
            Wg[a + (size_t)c * ldw] = tb;
            Wg[b + (size_t)c * ldw] = ta;
          }
          for (int c = tid; c < kw; c += NT) {
            const T tl = lsr(c, a);
            lsr(c, a) = lsr(c, b);
            lsr(c, b) = tl;
          }
          sym_swap<NT, T>(F, ld, m, vars, pre, a, b);
#pragma unroll
          for (int j = 0; j < R; j++) {
            const int r = rowof(j);
            if (r == a || r == b) {
              double* dst = (r == a) ? sh.xa : sh.xb;
              dst[0] = v[j];
              dst[1] = vp[j];
            }
          }
          // (a full barrier: the D entries written below share addresses
          // with the swap's stores)
          __syncthreads();
#pragma unroll
          for (int j = 0; j < R; j++) {
            const int r = rowof(j);
            if (r == a || r == b) {
              const double* src = (r == a) ? sh.xb : sh.xa;
              v[j] = src[0];
              vp[j] = src[1];
            }
          }
        } else {
        if (step > 0) {
#pragma unroll
          for (int j = 0; j < R; j++) {
            const int r = rowof(j);
            if (r >= k && r < m) {
              Wg[r + (size_t)kw * ldw] = v[j];
              if (step == 2) Wg[r + (size_t)(kw + 1) * ldw] = vp[j];
            }
          }
        }
        __syncthreads();
        if (a != b) {
          for (int c = tid; c < cnt; c += NT) {
            const double ta = Wg[a + (size_t)c * ldw];
            Wg[a + (size_t)c * ldw] = Wg[b + (size_t)c * ldw];
            Wg[b + (size_t)c * ldw] = ta;
          }
          for (int c = tid; c < kw; c += NT) {
            const T tl = lsr(c, a);
            lsr(c, a) = lsr(c, b);
            lsr(c, b) = tl;
          }
          sym_swap<NT, T>(F, ld, m, vars, pre, a, b);
#pragma unroll
          for (int j = 0; j < R; j++) {
            const int r = rowof(j);
            if (r == a || r == b) {
              double* dst = (r == a) ? sh.xa : sh.xb;
              dst[0] = v[j];
              dst[1] = vp[j];
            }
          }
          __syncthreads();
#pragma unroll
          for (int j = 0; j < R; j++) {
            const int r = rowof(j);
            if (r == a || r == b) {
              const double* src = (r == a) ? sh.xb : sh.xa;
              v[j] = src[0];
              vp[j] = src[1];
            }
          }
          // cached W rows of the swapped positions, from HBM
          for (int c = tid; c < 2 * cnt; c += NT) {
            const int r = (c < cnt) ? a : b;
            const int cc = (c < cnt) ? c : c - cnt;
            if (r - kstart < kWCRows) sh.wc[r - kstart][cc] = Wg[r + (size_t)cc * ldw];
          }
        }
        }
        if (step == 0) nc -= 1;
      }
      if (step == 1) {
        const double d = (kpiv == k) ? wkk : wii;
        const double dinv = (zero_col || d == 0.0) ? 0.0 : 1.0 / d;
#pragma unroll
        for (int j = 0; j < R; j++) {
          const int r = rowof(j);
          if (r >= k && r < m) {
            lsx(kw, j) = zero_col ? (T)v[j] : (T)(v[j] * dinv);
            if (fast || wswap) Wg[r + (size_t)kw * ldw] = v[j];
            if (!fast && r - kstart < kWCRows) sh.wc[r - kstart][kw] = v[j];
          }
        }
        if (lane == 0) sh.cT[kw] = zero_col ? 4 : 1;  // every wave writes the same value
        if (tid == 0) {
          F[k + (size_t)k * ld] = (T)d;
          P.ptype[io + k] = 1;
          P.dv[io + k] = d;
          if (d >= 0.0) sh.npos++;
          else sh.nneg++;
        }
      } else if (step == 2) {
        const double d11 = wkk, d21 = wik, d22 = wii;
        const double det = piv_det(d11, d22, d21);
        double i00, i01, i11;
        piv_inv2(d11, d21, d22, i00, i01, i11);
#pragma unroll
        for (int j = 0; j < R; j++) {
          const int r = rowof(j);
          if (r >= k && r < m) {
            lsx(kw, j) = (T)__builtin_fma(v[j], i00, vp[j] * i01);
            lsx(kw + 1, j) = (T)__builtin_fma(v[j], i01, vp[j] * i11);
            if (fast || wswap) {
              Wg[r + (size_t)kw * ldw] = v[j];
              Wg[r + (size_t)(kw + 1) * ldw] = vp[j];
            }
            if (!fast && r - kstart < kWCRows) {
              sh.wc[r - kstart][kw] = v[j];
              sh.wc[r - kstart][kw + 1] = vp[j];
            }
          }
        }
        if (lane == 0) {
          sh.cT[kw] = 2;
          sh.cT[kw + 1] = 3;
        }
        if (tid == 0) {
          F[k + (size_t)k * ld] = (T)d11;
          F[k + 1 + (size_t)k * ld] = (T)d21;
          F[k + 1 + (size_t)(k + 1) * ld] = (T)d22;
          P.dv[io + k] = d11;
          P.dv2[io + k] = d21;
          P.dv[io + k + 1] = d22;
          vars[k] = -(vars[k] + 1);
          vars[k + 1] = -(vars[k + 1] + 1);
          P.ptype[io + k] = 2;
          P.ptype[io + k + 1] = 3;
          if (det >= 0.0) {
            if (d11 >= 0.0) sh.npos += 2;
            else sh.nneg += 2;
          } else {
            sh.npos++;
            sh.nneg++;
          }
          sh.n2++;
        }
      }
      if (fast && step == 1) {
        const double a = sh.wc[k + 1 - kstart][kw];
#pragma unroll
        for (int j = 0; j < R; j++) vn[j] = __builtin_fma(-(double)lsx(kw, j), a, la1[j]);
      } else if (fast && step == 2) {
        const double a = sh.wc[k + 2 - kstart][kw], b = sh.wc[k + 2 - kstart][kw + 1];
#pragma unroll
        for (int j = 0; j < R; j++)
          vn[j] = __builtin_fma(-(double)lsx(kw + 1, j), b,
                                __builtin_fma(-(double)lsx(kw, j), a, la2[j]));
      }
      la_ok = fast;
      k += step;
      kw += step;
      if (!fast) {
        // the swaps and the cached rows are complete for every thread
        __syncthreads();
        reload();
      } else if (step == 1) {
#pragma unroll
        for (int j = 0; j < R; j++) {
          c0[j] = c1[j];
          c1[j] = f2[j];
          f2[j] = f3[j];
        }
        ld_col(k + 3, f3);
      } else {
#pragma unroll
        for (int j = 0; j < R; j++) {
          c0[j] = f2[j];
          c1[j] = f3[j];
        }
        ld_col(k + 2, f2);
        ld_col(k + 3, f3);
      }
      it++;
    }
    lds_sync();
    // L of the panel below each pivot block (the D entries were written at
    // the pivot), for the columns of the sequential loop
    for (int c = kw_spec; c < kw; c++) {
      const int pos = kstart + c;
      const int last = (sh.cT[c] == 2) ? pos + 1 : pos;
#pragma unroll
      for (int j = 0; j < R; j++) {
        const int r = rowof(j);
        if (r > last && r < m) F[r + (size_t)pos * ld] = lsx(c, j);
      }
    }
    for (int i = kstart + tid; i < m; i += NT) {
      P.vars[io + i] = vars[i];
      P.pre[io + i] = pre[i];
    }
    if (tid == 0 && k >= nc) {
      const int fs = P.ns[p] + P.din[p];
      P.npiv[p] = nc;
      P.ndel[p] = fs - nc;
      P.cnt[3 * p] = sh.npos;
      P.cnt[3 * p + 1] = sh.nneg;
      P.cnt[3 * p + 2] = sh.n2;
    }
  }
  if (tid == 0) {
    PanelState st = st0;
    st.k = k;
    st.nc = nc;
    st.kstart = kstart;
    st.kw = kw;
    st.npos = sh.npos;
    st.nneg = sh.nneg;
    st.n2 = sh.n2;
    st.pad = bkphase;
    P.ps[p] = st;
  }
}


// LDS-resident panel for a level whose largest front slot has maxcap rows:
// 512 threads with R = 1, 2 or 4 rows each.  False (use the global-memory
// panel) when the panel's L does not fit LDS: above 2048 rows for FP32
// fronts, above 1024 rows for FP64 fronts.
template <class T, bool BKC>
static bool launch_panel_rr(int maxcap, dim3 grid, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                            const DevPlan& P, const int* nodes) {
  constexpr int NT = 512;
  auto lds = [&](int r) {
    return (uint32_t)(sizeof(T) * kNBLarge * (size_t)r * NT + sizeof(int) * 2 * (size_t)maxcap);
  };
  if (maxcap <= NT / 2) {
    // fronts of at most 256 rows: 4 waves, every one holding rows
    constexpr int NH = NT / 2;
    const uint32_t b = (uint32_t)(sizeof(T) * kNBLarge * (size_t)NH + sizeof(int) * 2 * (size_t)maxcap);
    hipExtLaunchKernelGGL((k_panel_rr<NH, 1, T, BKC>), grid, dim3(NH), b, st, e0, e1, 0, P, nodes);
  } else if (maxcap <= NT) {
    hipExtLaunchKernelGGL((k_panel_rr<NT, 1, T, BKC>), grid, dim3(NT), lds(1), st, e0, e1, 0, P, nodes);
  } else if (maxcap <= 2 * NT) {
    hipExtLaunchKernelGGL((k_panel_rr<NT, 2, T, BKC>), grid, dim3(NT), lds(2), st, e0, e1, 0, P, nodes);
  } else if (maxcap <= 4 * NT && sizeof(T) == 4) {
    if constexpr (sizeof(T) == 4)
      hipExtLaunchKernelGGL((k_panel_rr<NT, 4, T, BKC>), grid, dim3(NT), lds(4), st, e0, e1, 0, P,
                            nodes);
  } else {
    return false;
  }
  return true;
}


bool launch_large_panel(const KKT_PANEL_T*, bool bkc, int maxcap, dim3 grid, hipStream_t st,
                        hipEvent_t e0, hipEvent_t e1, const DevPlan& P, const int* nodes) {
  using T = KKT_PANEL_T;
  const bool rr = bkc ? launch_panel_rr<T, true>(maxcap, grid, st, e0, e1, P, nodes)
                      : launch_panel_rr<T, false>(maxcap, grid, st, e0, e1, P, nodes);
  if (!rr) {
    const size_t iv = sizeof(int) * 2 * (size_t)maxcap;
    if (iv <= 96 * 1024)
      hipExtLaunchKernelGGL((k_panel<kPanelThreads, T, false>), grid, dim3(kPanelThreads),
                            (uint32_t)iv, st, e0, e1, 0, P, nodes);
    else
      hipExtLaunchKernelGGL((k_panel<kPanelThreads, T, true>), grid, dim3(kPanelThreads), 0, st,
                            e0, e1, 0, P, nodes);
  }
  return rr;
}

}  // namespace kktldl
