// Declarations shared by the translation units of the MI355X LDL^T engine:
// the device plan, the per-front panel state, and the device helpers used by
// the factor kernels (kkt_ldl.hip) and by the large-front panel kernels
// (kkt_panel.hip, compiled once per front element type so that the two
// largest groups of kernel instances build in parallel).
#ifndef KKT_COMMON_HPP
#define KKT_COMMON_HPP

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/kkt_ldl.h"

namespace kktldl {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& s) : std::runtime_error(s), code(c) {}
};

#define HIPCHECK(expr)                                                        \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      throw Error(-2, std::string("HIP error ") + hipGetErrorString(_e) +     \
                          " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

constexpr int kNB = 64;        // panel width bound (reference block_size)
constexpr int kTPB = 256;      // threads per workgroup
constexpr int kWave = 64;
constexpr int kAsmCols = 4;    // columns per assembly workgroup
constexpr int kPanelThreads = 1024;  // workgroup of the large-front panel
constexpr int kNBLarge = 16;    // panel width of large fronts (narrow: the panel re-reads it per column)
constexpr int kLdsFront = 96;   // default bound of fronts assembled and factored in LDS

// --------------------------------------------------------------------------
// Device-side plan (pointers into device memory).
// --------------------------------------------------------------------------
// Factorization state of a large front between panel launches.
struct PanelState {
  int k, nc, kstart, kw;
  int npos, nneg, n2, pad;
};

struct DevPlan {
  // static per node
  const int* ns;       // own (fully summed nominal) columns
  const int* cs;       // nominal contribution rows
  const int* cap;      // front slot dimension
  const int* vptr;     // node_ptr
  const int* cptr;     // contrib_ptr
  const int* chptr;    // children list pointer [N+1]
  const int* ch;       // children
  const long long* invoff;  // per node (as child): offset of its inverse map
  const long long* foff;    // front offset (elements)
  const long long* ioff;    // vars/pre offset
  const long long* woff;    // panel scratch offset (elements)
  const long long* uoff;    // solve update-vector offset (elements, per rhs)
  const int* mode;     // 0 threshold, 1 Bunch-Kaufman root, 2 Cholesky
  const int* node_to_var;
  const int* contrib_rows;
  const int* invnom;   // for each child: parent's nominal position -> child contrib index or -1
  const int* fwdnom;   // per contribution row of a child (contrib_rows order): the
                       // parent's nominal position of that variable
  const int* am_ptr;   // per own column (global index g in node_to_var order)
  const int* am_src;   // CSR index into vals
  const int* am_row;   // nominal row in the front
  const int* am_col;   // own column of the entry inside its front
  int lds_t;           // small-front solve kernels: doubles of LDS before the staged block
  int lds_b;           // small-front solve kernels: leading dimension of the staged block
  // dynamic per node
  int* din;            // delayed pivots received
  int* m;              // actual front dimension
  int* npiv;
  int* ndel;
  int* doff;           // offset of this node's delays inside the parent's section
  int* stat;           // 0 ok, 1 overflow / failed
  int* need;           // front dimension required when overflowing
  int* cnt;            // [3N]: npos, nneg, n2x2
  PanelState* ps;      // [N]
  // arenas
  void* F;             // fronts, element type T (double or float)
  int* vars;
  int* pre;
  int* ptype;          // per front position: 1 = 1x1, 2/3 = first/second of 2x2
  double* dv;          // D in FP64 per front position: d11 / d22 (and 1x1 d)
  double* dv2;         // FP64 off-diagonal d21 of a 2x2 block, at its first position
  double* W;
  double* U;
  // large nodes: inverses of the kPB x kPB diagonal blocks of L11, two
  // copies per block (column major, then transposed), and per pair of
  // blocks the coupling term of the composed inverse, at pinvoff[node]
  double* pinv;
  const long long* pinvoff;
  // numeric input
  const double* vals;
  const double* diag;
  double ustab, ptol;
  int spec;            // large-front panels choose their pivots speculatively
};

__device__ inline int unpiv(int v) { return v >= 0 ? v : -v - 1; }

// v - sum_{c < n} a[c * sa] * b[c], accumulated in index order, with the
// loads of a (global or LDS) issued in batches of 8 so that their latency
// is paid once per batch rather than once per term.
template <class TA>
__device__ inline double sub_dot(double v, const TA* a, size_t sa,
                                 const double* b, int n) {
  int c = 0;
  for (; c + 8 <= n; c += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = (double)a[(size_t)(c + u) * sa];
#pragma unroll
    for (int u = 0; u < 8; u++) v -= x[u] * b[c + u];
  }
  if (c < n) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = (c + u < n) ? (double)a[(size_t)(c + u) * sa] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (c + u < n) v -= x[u] * b[c + u];
  }
  return v;
}

// sum_{i} a[i] * b[i] over i = i0, i0 + st, ... < i1, loads batched by 8.
template <class TA>
__device__ inline double strided_dot(const TA* a, const double* b, int i0, int i1,
                                     int st) {
  double s = 0.0;
  int i = i0;
  for (; i + 7 * st < i1; i += 8 * st) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = (double)a[i + u * st];
#pragma unroll
    for (int u = 0; u < 8; u++) s += x[u] * b[i + u * st];
  }
  for (; i < i1; i += st) s += (double)a[i] * b[i];
  return s;
}

// Row of child c's front (post-pivot) that holds the parent's pre-pivot
// position i, or -1.  Children contribute their delayed pivots (rows
// npiv_c .. fs_c-1) and their nominal contribution rows (fs_c + q).
template <class T>
struct ChildView {
  const T* F;
  int m, fs, npiv, ndel, doff;
  long long inv;
};

template <class T>
__device__ inline T* front_ptr(const DevPlan& P, int p) {
  return (T*)P.F + P.foff[p];
}

template <class T>
__device__ inline ChildView<T> load_child(const DevPlan& P, int c) {
  ChildView<T> v;
  v.F = front_ptr<T>(P, c);
  v.m = P.m[c];
  v.fs = P.ns[c] + P.din[c];
  v.npiv = P.npiv[c];
  v.ndel = P.ndel[c];
  v.doff = P.doff[c];
  v.inv = P.invoff[c];
  return v;
}

template <class CV>
__device__ inline int child_row(const DevPlan& P, const CV& cv, int i, int ns, int din) {
  if (i < ns) {
    int q = P.invnom[cv.inv + i];
    return q >= 0 ? cv.fs + q : -1;
  }
  if (i < ns + din) {
    int d = i - ns - cv.doff;
    return (d >= 0 && d < cv.ndel) ? cv.npiv + d : -1;
  }
  int q = P.invnom[cv.inv + (i - din)];
  return q >= 0 ? cv.fs + q : -1;
}

// child_row without branches: the map entry is loaded for every position
// (index clamped into the map), so that several lookups can be in flight
template <class CV>
__device__ inline int child_row_nb(const DevPlan& P, const CV& cv, int i, int ns, int din) {
  const long long inv = cv.inv;
  const int fs = cv.fs, npiv = cv.npiv, ndel = cv.ndel, doff = cv.doff;
  const int q = P.invnom[inv + (i < ns ? i : max(i - din, 0))];
  const int d = i - ns - doff;
  const int rd = (unsigned)d < (unsigned)ndel ? npiv + d : -1;
  const int rq = q >= 0 ? fs + q : -1;
  return (i >= ns && i < ns + din) ? rd : rq;
}


// --------------------------------------------------------------------------
// Workgroup reductions.
// --------------------------------------------------------------------------
template <int NT>
struct RedBuf {
  double v[NT / kWave];
  int i[NT / kWave];
  double v2[NT / kWave];
};

// Cross-lane moves by DPP (VALU speed): quad_perm [1,0,3,2] and [2,3,0,1],
// row_half_mirror and row_mirror pair every lane of a 16-lane row with every
// other after four steps.
template <int CTRL>
__device__ inline int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ inline double dpp_d(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = dpp_i<CTRL>((int)b);
  const int hi = dpp_i<CTRL>((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double lane_d(double x, int l) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}

// (max |v| with the first row on ties, max of a second value); the order of
// combination does not change the result.
__device__ inline void am_combine(double& b, int& i, double& r, double ob, int oi, double orr) {
  if (ob > b || (ob == b && oi < i)) {
    b = ob;
    i = oi;
  }
  r = fmax(r, orr);
}
template <int CTRL>
__device__ inline void am_step(double& b, int& i, double& r) {
  am_combine(b, i, r, dpp_d<CTRL>(b), dpp_i<CTRL>(i), dpp_d<CTRL>(r));
}
__device__ inline void row_argmax2(double& b, int& i, double& r) {
  am_step<0xB1>(b, i, r);
  am_step<0x4E>(b, i, r);
  am_step<0x141>(b, i, r);
  am_step<0x140>(b, i, r);
}
// wave-wide result, uniform in every lane
__device__ inline void wave_argmax2(double& b, int& i, double& r) {
  row_argmax2(b, i, r);
  double rb = lane_d(b, 0), rr = lane_d(r, 0);
  int ri = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
  for (int q = 1; q < kWave / 16; q++)
    am_combine(rb, ri, rr, lane_d(b, 16 * q), __builtin_amdgcn_readlane(i, 16 * q),
               lane_d(r, 16 * q));
  b = rb;
  i = ri;
  r = rr;
}
__device__ inline double row_max(double x) {
  x = fmax(x, dpp_d<0xB1>(x));
  x = fmax(x, dpp_d<0x4E>(x));
  x = fmax(x, dpp_d<0x141>(x));
  x = fmax(x, dpp_d<0x140>(x));
  return x;
}
__device__ inline double wave_max(double x) {
  x = row_max(x);
  double r = lane_d(x, 0);
#pragma unroll
  for (int q = 1; q < kWave / 16; q++) r = fmax(r, lane_d(x, 16 * q));
  return r;
}

// Block-wide (max |x|, first index) and second plain max; results broadcast.
// Each wave reduces by DPP, then the waves' results are combined in wave
// order; the combination is order independent, so the result is the
// lexicographic optimum whatever the wave count.
template <int NT>
__device__ inline void block_argmax2(double& val, int& idx, double& val2,
                                     RedBuf<NT>& rb) {
  wave_argmax2(val, idx, val2);
  if constexpr (NT == kWave) return;
  const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
  __syncthreads();
  if (l == 0) {
    rb.v[w] = val;
    rb.i[w] = idx;
    rb.v2[w] = val2;
  }
  __syncthreads();
  val = rb.v[0];
  idx = rb.i[0];
  val2 = rb.v2[0];
  for (int t = 1; t < NT / kWave; t++) am_combine(val, idx, val2, rb.v[t], rb.i[t], rb.v2[t]);
}

// --------------------------------------------------------------------------
// Front factorization (one workgroup).  F: front (ld = m), W: panel
// workspace (ld = ldw, kNB columns).
// --------------------------------------------------------------------------
template <int NT>
struct FrontShared {
  double wr[kNB];
  double wr2[kNB];
  double wkk, wik, wii;  // W[k,kw], W[imax,kw], W1[imax] stashed by their owners
  RedBuf<NT> rb;
  int npos, nneg, n2;
};

template <int NT, class T>
__device__ inline void sym_swap(T* F, int ld, int m, int* vars, int* pre,
                                int a, int b) {
  const int tid = threadIdx.x;
  if (a < 0 || b >= m || a >= b) return;  // positions outside the front: no swap
  if (tid == 0) {
    int t = vars[a];
    vars[a] = vars[b];
    vars[b] = t;
    t = pre[a];
    pre[a] = pre[b];
    pre[b] = t;
    T d = F[a + (size_t)a * ld];
    F[a + (size_t)a * ld] = F[b + (size_t)b * ld];
    F[b + (size_t)b * ld] = d;
  }
  for (int j = tid; j < a; j += NT) {
    T t = F[a + (size_t)j * ld];
    F[a + (size_t)j * ld] = F[b + (size_t)j * ld];
    F[b + (size_t)j * ld] = t;
  }
  for (int j = a + 1 + tid; j < b; j += NT) {
    T t = F[j + (size_t)a * ld];
    F[j + (size_t)a * ld] = F[b + (size_t)j * ld];
    F[b + (size_t)j * ld] = t;
  }
  for (int i = b + 1 + tid; i < m; i += NT) {
    T t = F[i + (size_t)a * ld];
    F[i + (size_t)a * ld] = F[i + (size_t)b * ld];
    F[i + (size_t)b * ld] = t;
  }
}

// One panel of at most nb-1 pivots starting at column k.  Returns with k,
// nc advanced; W[k0:, 0:kw] and F[:, kstart:k] hold the panel for the
// trailing update.  mode 0: threshold pivoting with delays; mode 1:
// Bunch-Kaufman (every column is a candidate, nothing is delayed).
template <int NT, class T>
__device__ void panel_factor(T* F, int ld, int m, int* vars, int* pre,
                             int* ptype, double* dv, double* dv2, double* W, int ldw,
                             int nb, int mode,
                             double ustab,
                             double ptol, int& k, int& nc, int& kw_out,
                             FrontShared<NT>& sh) {
  const int tid = threadIdx.x;
  const double alpha = 0.6403882032022076;  // (1 + sqrt(17)) / 8
  const int kstart = k;
  int kw = 0;
  while (kw < nb - 1 && k < nc) {
    // W[k:m, kw] = F[k:m, k] - F[k:m, kstart:k] W[k, 0:kw]^T
    for (int c = tid; c < kw; c += NT) sh.wr[c] = W[k + (size_t)c * ldw];
    __syncthreads();
    double lbest = -1.0, lrest = 0.0;
    int lidx = 0x7fffffff;
    for (int i = k + tid; i < m; i += NT) {
      double v = (double)F[i + (size_t)k * ld];
      v = sub_dot(v, F + i + (size_t)kstart * ld, ld, sh.wr, kw);
      W[i + (size_t)kw * ldw] = v;
      if (i == k) sh.wkk = v;
      double a = fabs(v);
      if (i > k && i < nc) {
        if (a > lbest) {
          lbest = a;
          lidx = i;
        }
      } else if (i >= nc) {
        lrest = fmax(lrest, a);
      }
    }
    block_argmax2<NT>(lbest, lidx, lrest, sh.rb);
    const double wkk = sh.wkk;
    const double awkk = fabs(wkk);
    int imax = 0;
    double awik = 0.0;
    if (k + 1 < nc) {
      // lidx keeps its sentinel only when every candidate entry is NaN (no
      // comparison succeeds): take the next candidate, so that the row
      // search and the swaps stay inside the front
      const bool found = lidx < nc;
      imax = found ? lidx : k + 1;
      awik = found ? lbest : __builtin_nan("");  // |W[imax, kw]|
    }
    const double colmax = fmax(awik, lrest);
    int step = 0, kpiv = -1;
    bool zero_col = false;
    bool need_row = false;
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
    if (need_row) {
      // W[k:m, kw+1] = column imax updated by the panel so far
      for (int c = tid; c < kw; c += NT) sh.wr2[c] = W[imax + (size_t)c * ldw];
      __syncthreads();
      double* W1 = W + (size_t)(kw + 1) * ldw;
      double lrow = 0.0;
      for (int i = k + tid; i < m; i += NT) {
        double v = (double)(i < imax ? F[imax + (size_t)i * ld] : F[i + (size_t)imax * ld]);
        v = sub_dot(v, F + i + (size_t)kstart * ld, ld, sh.wr2, kw);
        W1[i] = v;
        if (i == imax) sh.wii = v;
        if (i == imax) sh.wik = W[imax + (size_t)kw * ldw];
        if (i != imax) lrow = fmax(lrow, fabs(v));
      }
      double dummy = -1.0;
      int di = 0;
      block_argmax2<NT>(dummy, di, lrow, sh.rb);
      const double rowmax = lrow;
      const double awii = fabs(sh.wii);
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
          const double wik = sh.wik;
          double det = fabs(wkk * sh.wii - wik * wik);
          if (det >= ptol) {
            double delta = det / fmax(awii + awik, awkk + awik);
            if (delta >= ustab * colmax && delta >= ustab * rowmax) {
              step = 2;
              kpiv = imax;
            }
          }
        }
      }
      if (step == 1 && kpiv == imax) {
        for (int i = k + tid; i < m; i += NT) W[i + (size_t)kw * ldw] = W1[i];
      }
      __syncthreads();
    }
    if (step == 0) {  // delay candidate k (threshold mode only)
      if (k < nc - 1) {
        sym_swap<NT, T>(F, ld, m, vars, pre, k, nc - 1);
        for (int c = tid; c < kw; c += NT) {
          double t = W[k + (size_t)c * ldw];
          W[k + (size_t)c * ldw] = W[nc - 1 + (size_t)c * ldw];
          W[nc - 1 + (size_t)c * ldw] = t;
        }
      }
      nc -= 1;
      __syncthreads();
      continue;
    }
    const int ksrc = k + step - 1;
    if (kpiv != ksrc) {
      sym_swap<NT, T>(F, ld, m, vars, pre, ksrc, kpiv);
      for (int c = tid; c < kw + step; c += NT) {
        double t = W[ksrc + (size_t)c * ldw];
        W[ksrc + (size_t)c * ldw] = W[kpiv + (size_t)c * ldw];
        W[kpiv + (size_t)c * ldw] = t;
      }
      __syncthreads();
    }
    if (step == 1) {
      const double d = (kpiv == k) ? wkk : sh.wii;
      const double dinv = (zero_col || d == 0.0) ? 0.0 : 1.0 / d;
      for (int i = k + tid; i < m; i += NT) {
        double v = W[i + (size_t)kw * ldw];
        F[i + (size_t)k * ld] = (T)((i == k || zero_col) ? v : v * dinv);
      }
      if (tid == 0) {
        ptype[k] = 1;
        dv[k] = d;
        if (d >= 0.0) sh.npos++;
        else sh.nneg++;
      }
    } else {
      const double d11 = wkk;       // W[k, kw]
      const double d21 = sh.wik;    // W[imax, kw], now in row k+1
      const double d22 = sh.wii;    // W1[imax], now W[k+1, kw+1]
      const double det = d11 * d22 - d21 * d21;
      const double detinv = 1.0 / det;
      const double i00 = d22 * detinv, i01 = -d21 * detinv, i11 = d11 * detinv;
      for (int i = k + 2 + tid; i < m; i += NT) {
        double a = W[i + (size_t)kw * ldw], b = W[i + (size_t)(kw + 1) * ldw];
        F[i + (size_t)k * ld] = (T)(a * i00 + b * i01);
        F[i + (size_t)(k + 1) * ld] = (T)(a * i01 + b * i11);
      }
      if (tid == 0) {
        F[k + (size_t)k * ld] = (T)d11;
        F[k + 1 + (size_t)k * ld] = (T)d21;
        F[k + 1 + (size_t)(k + 1) * ld] = (T)d22;
        dv[k] = d11;
        dv2[k] = d21;
        dv[k + 1] = d22;
        vars[k] = -(vars[k] + 1);
        vars[k + 1] = -(vars[k + 1] + 1);
        ptype[k] = 2;
        ptype[k + 1] = 3;
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
    k += step;
    kw += step;
    __syncthreads();
  }
  kw_out = kw;
}

// F[k:m, k:m] (lower) -= F[k:m, kstart:kstart+kw] * W[k:m, 0:kw]^T, one
// workgroup, 4x4 register tiles.
template <int NT, class T>
__device__ __attribute__((always_inline)) void trailing_update_wg(T* F, int ld, int m, const double* W,
                                   int ldw, int kstart, int k, int kw) {
  const int nt = m - k;
  if (nt <= 0 || kw == 0) return;
  const int nb4 = (nt + 3) / 4;
  const int ntiles = nb4 * (nb4 + 1) / 2;
  for (int t = threadIdx.x; t < ntiles; t += NT) {
    // map t -> (bi >= bj)
    int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (bi * (bi + 1) / 2 > t) bi--;
    while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
    const int bj = t - bi * (bi + 1) / 2;
    const int i0 = k + 4 * bi, j0 = k + 4 * bj;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
    for (int c0 = 0; c0 < kw; c0 += 4) {
      double l[4][4], w[4][4];  // [column][row]: 32 loads issued together
#pragma unroll
      for (int cc = 0; cc < 4; cc++) {
        const bool okc = c0 + cc < kw;
        const T* Lc = F + (size_t)(kstart + c0 + cc) * ld;
        const double* Wc = W + (size_t)(c0 + cc) * ldw;
#pragma unroll
        for (int a = 0; a < 4; a++) {
          l[cc][a] = (okc && i0 + a < m) ? (double)Lc[i0 + a] : 0.0;
          w[cc][a] = (okc && j0 + a < m) ? Wc[j0 + a] : 0.0;
        }
      }
#pragma unroll
      for (int cc = 0; cc < 4; cc++)
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++) acc[a][b] += l[cc][a] * w[cc][b];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        int i = i0 + a, j = j0 + b;
        if (i < m && j < m && i >= j)
          F[i + (size_t)j * ld] = (T)((double)F[i + (size_t)j * ld] - acc[a][b]);
      }
  }
  __syncthreads();
}

// Cholesky mode (include/sparse_ldl.h:2010-2044): right-looking, one
// workgroup; used by the reference's tests only.
template <int NT, class T>
__device__ void cholesky_front(T* F, int ld, int m, int fs) {
  const int tid = threadIdx.x;
  for (int k = 0; k < fs; k++) {
    const double d = sqrt((double)F[k + (size_t)k * ld]);
    __syncthreads();
    for (int i = k + 1 + tid; i < m; i += NT) F[i + (size_t)k * ld] = (T)((double)F[i + (size_t)k * ld] / d);
    if (tid == 0) F[k + (size_t)k * ld] = (T)d;
    __syncthreads();
    const int nt = m - k - 1;
    for (int e = tid; e < nt * nt; e += NT) {
      int jj = e / nt, ii = e % nt;
      if (ii < jj) continue;
      int i = k + 1 + ii, j = k + 1 + jj;
      F[i + (size_t)j * ld] = (T)((double)F[i + (size_t)j * ld] -
                                  (double)F[i + (size_t)k * ld] * (double)F[j + (size_t)k * ld]);
    }
    __syncthreads();
  }
}


constexpr int kWCRows = 2 * kNBLarge;  // W rows cached in LDS (panel positions and swap partners)

// Pivot arithmetic shared by the sequential and the speculative panel paths,
// evaluated without contraction so that both make the same decisions.
__device__ inline double piv_det(double d11, double d22, double d21) {
#pragma clang fp contract(off)
  return d11 * d22 - d21 * d21;
}
__device__ inline double piv_delta(double det, double awkk, double awii, double awik) {
#pragma clang fp contract(off)
  return det / fmax(awii + awik, awkk + awik);
}
__device__ inline void piv_inv2(double d11, double d21, double d22, double& i00, double& i01,
                                double& i11) {
#pragma clang fp contract(off)
  const double detinv = 1.0 / piv_det(d11, d22, d21);
  i00 = d22 * detinv;
  i01 = -d21 * detinv;
  i11 = d11 * detinv;
}

template <int NW>
struct RRShared {
  double pb[2][NW];  // column reduction: best candidate |v|
  int pi[2][NW];     //   its row
  double pr[2][NW];  //   max |v| over the rows below the candidates
  double qr[2][NW];  // row reduction of the imax column
  double wkk[2], fk1k[2], wik[2], wii[2];
  double wc[kWCRows][kNBLarge + 1];  // W rows of positions kstart .. kstart + kWCRows - 1 (padded)
  double xw[kNBLarge];           // W row of a general imax
  double xa[kNBLarge + 2], xb[kNBLarge + 2];
  double swii, swik;
  int cT[kNBLarge];  // per panel column: 1 1x1, 2/3 first/second of 2x2, 4 unscaled
  int npos, nneg, n2;
  // speculative panel: the steps chosen from the leading rows
  int sp_n, sp_end;
  unsigned sp_fail;
  int sp_c[kNBLarge], sp_t[kNBLarge];
  double sp_thr[kNBLarge], sp_awik[kNBLarge];
  double sp_m[kNBLarge][3];  // 1/d, or the 2x2 inverse i00, i01, i11
  double sp_d[kNBLarge][3];  // d, or d11, d21, d22
};

// Workgroup barrier that orders LDS only: it waits for this wave's LDS
// operations, not for its HBM loads and stores (a full __syncthreads()
// waits for every outstanding global access, which would stall each pivot
// step on the W stores and on the prefetch loads in flight).
__device__ inline void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}


// One panel launch over the large fronts of a level (kkt_panel.hip): the
// LDS-resident panel k_panel_rr where the panel's L fits LDS, else the
// global-memory panel k_panel.  bkc: the level holds a root front
// (Bunch-Kaufman tests).  e0/e1 time the kernel itself.  Returns false when
// the global-memory panel was launched.
bool launch_large_panel(const float*, bool bkc, int maxcap, dim3 grid, hipStream_t st,
                        hipEvent_t e0, hipEvent_t e1, const DevPlan& P, const int* nodes);
bool launch_large_panel(const double*, bool bkc, int maxcap, dim3 grid, hipStream_t st,
                        hipEvent_t e0, hipEvent_t e1, const DevPlan& P, const int* nodes);

// Fronts held whole in LDS (kkt_front.hip), threshold / Bunch-Kaufman modes:
// `count` fronts of at most `cap` rows, one workgroup each (chain = false),
// or one workgroup factoring them in list order (chain = true).  sized:
// the kernel instance for the launch's front bound (else the one for the
// largest fronts; the same results).
void launch_front_rl(const float*, int cap, int count, bool chain, bool sized, hipStream_t st,
                     const DevPlan& P, const int* nodes);
void launch_front_rl(const double*, int cap, int count, bool chain, bool sized, hipStream_t st,
                     const DevPlan& P, const int* nodes);
int rl_max_front();

}  // namespace kktldl

#endif  // KKT_COMMON_HPP
