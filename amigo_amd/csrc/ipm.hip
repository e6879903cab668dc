// Interior-point condensation on the GPU: the barrier diagonal Sigma and the
// condensed right-hand side (the factorization's and the solve's inputs), the
// bound-dual back-substitution, the fraction-to-the-boundary step and the
// complementarity / KKT-error reductions.
//
// The reference runs these as loops over the primal list
// (include/interior_point_backend.h:146-369), writing Sigma and the residual
// by scatter into a zeroed vector (include/interior_point_optimizer.h:
// 233-285).  Here the host inverts the index lists once per problem: for
// every entry of the full KKT vector, the primal position and the constraint
// position that write it.  Sigma and the residual are then one coalesced pass
// over the full vector (no zero fill, no scatter); the per-primal quantities
// run over the primal list.  Reductions are deterministic: a fixed grid of
// kRedBlocks blocks, each thread accumulating its elements in index order,
// fixed-shape tree reductions in LDS, and one final block.
//
// This file is compiled with -ffp-contract=off: every elementwise result is
// the reference's expression evaluated in the same order with IEEE double
// operations, so the vectors are bitwise equal to the serial loops'.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kkt_ipm.h"

namespace kktipm {

static thread_local std::string g_err;

#define ICHECK(expr)                                                               \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e)); \
  } while (0)

constexpr int kTPB = 256;
constexpr int kRedBlocks = 256;  // partial results per reduction
static_assert(kRedBlocks == kTPB, "the final stage reduces one partial per thread");

struct Info {
  int n, np, nc;
  const int* pidx;  // [np] primal positions
  const double* lb;
  const double* ub;
  const int* cidx;  // [nc] constraint positions
  const double* lbh;
  const int* pof;   // [n] primal position writing entry idx, or -1
  const int* cof;   // [n] constraint position writing entry idx, or -1
};

// Fixed-shape tree reduction of one value per thread into s[0].
template <class T, class Op>
__device__ inline void block_reduce(T* s, T v, Op op) {
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int w = kTPB / 2; w > 0; w >>= 1) {
    if (t < w) s[t] = op(s[t], s[t + w]);
    __syncthreads();
  }
}

struct Sum {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct Max {
  __device__ double operator()(double a, double b) const { return a < b ? b : a; }
};
struct Min {
  __device__ double operator()(double a, double b) const { return b < a ? b : a; }
};

// Sigma at the primal positions (zl/gap_l + zu/gap_u), 0 elsewhere.
__global__ void __launch_bounds__(kTPB) k_sigma(Info I, const double* __restrict__ x,
                                                const double* __restrict__ zl,
                                                const double* __restrict__ zu,
                                                double* __restrict__ diag) {
  const int idx = blockIdx.x * kTPB + threadIdx.x;
  if (idx >= I.n) return;
  const int i = I.pof[idx];
  double sigma = 0.0;
  if (i >= 0) {
    const double xv = x[idx], lb = I.lb[i], ub = I.ub[i];
    if (!isinf(lb)) {
      const double gap = xv - lb;
      sigma += zl[i] / gap;
    }
    if (!isinf(ub)) {
      const double gap = ub - xv;
      sigma += zu[i] / gap;
    }
  }
  diag[idx] = sigma;
}

// Condensed right-hand side, and per-block partial sums of its squares.
__global__ void __launch_bounds__(kTPB) k_residual(Info I, double mu,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ zl,
                                                   const double* __restrict__ zu,
                                                   const double* __restrict__ grad,
                                                   double* __restrict__ res,
                                                   double* __restrict__ part) {
  __shared__ double s[kTPB];
  double acc = 0.0;
  for (int idx = blockIdx.x * kTPB + threadIdx.x; idx < I.n; idx += kRedBlocks * kTPB) {
    double out = 0.0;
    const int j = I.cof[idx];
    const int i = I.pof[idx];
    if (j >= 0) {
      out = -(grad[idx] - I.lbh[j]);
    } else if (i >= 0) {
      const double xv = x[idx], lb = I.lb[i], ub = I.ub[i];
      double r = grad[idx] - zl[i] + zu[i];
      if (!isinf(lb)) {
        const double gap = xv - lb;
        r += (gap * zl[i] - mu) / gap;
      }
      if (!isinf(ub)) {
        const double gap = ub - xv;
        r -= (gap * zu[i] - mu) / gap;
      }
      out = -r;
    }
    res[idx] = out;
    acc += out * out;
  }
  block_reduce(s, acc, Sum());
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kTPB) k_bound_dual(Info I, double mu,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ zl,
                                                     const double* __restrict__ zu,
                                                     const double* __restrict__ px,
                                                     double* __restrict__ dzl,
                                                     double* __restrict__ dzu) {
  const int i = blockIdx.x * kTPB + threadIdx.x;
  if (i >= I.np) return;
  const int idx = I.pidx[i];
  const double xv = x[idx], dx = px[idx], lb = I.lb[i], ub = I.ub[i];
  double a = 0.0, b = 0.0;
  if (!isinf(lb)) {
    const double gap = xv - lb;
    const double rhs = gap * zl[i] - mu;
    a = -(rhs + zl[i] * dx) / gap;
  }
  if (!isinf(ub)) {
    const double gap = ub - xv;
    const double rhs = gap * zu[i] - mu;
    b = -(rhs - zu[i] * dx) / gap;
  }
  dzl[i] = a;
  dzu[i] = b;
}

// Fraction to the boundary: per block, the smallest primal and dual step
// bound and the first primal position attaining it (positions compared on
// ties, so the result is the serial loop's first minimum).
__global__ void __launch_bounds__(kTPB) k_max_step(Info I, double tau,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ zl,
                                                   const double* __restrict__ zu,
                                                   const double* __restrict__ dxv,
                                                   const double* __restrict__ dzl,
                                                   const double* __restrict__ dzu,
                                                   double* __restrict__ part,
                                                   int* __restrict__ ipart) {
  __shared__ double sv[kTPB];
  __shared__ int si[kTPB];
  double bx = INFINITY, bz = INFINITY;
  int ix = INT_MAX, iz = INT_MAX;
  auto take = [](double a, int i, double& b, int& bi) {
    if (a < b) {
      b = a;
      bi = i;
    }
  };
  for (int i = blockIdx.x * kTPB + threadIdx.x; i < I.np; i += kRedBlocks * kTPB) {
    const int idx = I.pidx[i];
    const double xv = x[idx], dx = dxv[idx], lb = I.lb[i], ub = I.ub[i];
    if (!isinf(lb)) {
      if (dx < 0.0) take(-tau * (xv - lb) / dx, i, bx, ix);
      if (dzl[i] < 0.0) take(-tau * zl[i] / dzl[i], i, bz, iz);
    }
    if (!isinf(ub)) {
      if (dx > 0.0) take(tau * (ub - xv) / dx, i, bx, ix);
      if (dzu[i] < 0.0) take(-tau * zu[i] / dzu[i], i, bz, iz);
    }
  }
  // (value, position) minimum, the smaller position on equal values
  for (int q = 0; q < 2; q++) {
    double v = q == 0 ? bx : bz;
    int vi = q == 0 ? ix : iz;
    sv[threadIdx.x] = v;
    si[threadIdx.x] = vi;
    __syncthreads();
    for (int w = kTPB / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        const double ov = sv[threadIdx.x + w];
        const int oi = si[threadIdx.x + w];
        if (ov < sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
          sv[threadIdx.x] = ov;
          si[threadIdx.x] = oi;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[2 * blockIdx.x + q] = sv[0];
      ipart[2 * blockIdx.x + q] = si[0];
    }
    __syncthreads();
  }
}

// Complementarity products: per block, their sum, count and minimum.
__global__ void __launch_bounds__(kTPB) k_compl(Info I, const double* __restrict__ x,
                                                const double* __restrict__ zl,
                                                const double* __restrict__ zu,
                                                double* __restrict__ part) {
  __shared__ double s[kTPB];
  double sum = 0.0, cnt = 0.0, mn = DBL_MAX;
  for (int i = blockIdx.x * kTPB + threadIdx.x; i < I.np; i += kRedBlocks * kTPB) {
    const int idx = I.pidx[i];
    const double xv = x[idx], lb = I.lb[i], ub = I.ub[i];
    if (!isinf(lb)) {
      const double c = (xv - lb) * zl[i];
      sum += c;
      cnt += 1.0;
      mn = c < mn ? c : mn;
    }
    if (!isinf(ub)) {
      const double c = (ub - xv) * zu[i];
      sum += c;
      cnt += 1.0;
      mn = c < mn ? c : mn;
    }
  }
  block_reduce(s, sum, Sum());
  if (threadIdx.x == 0) part[3 * blockIdx.x] = s[0];
  __syncthreads();
  block_reduce(s, cnt, Sum());
  if (threadIdx.x == 0) part[3 * blockIdx.x + 1] = s[0];
  __syncthreads();
  block_reduce(s, mn, Min());
  if (threadIdx.x == 0) part[3 * blockIdx.x + 2] = s[0];
}

// Infinity-norm optimality errors: per block, the dual, primal and
// complementarity maxima.
__global__ void __launch_bounds__(kTPB) k_kkt_error(Info I, double mu,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ zl,
                                                    const double* __restrict__ zu,
                                                    const double* __restrict__ grad,
                                                    double* __restrict__ part) {
  __shared__ double s[kTPB];
  double dual = 0.0, primal = 0.0, comp = 0.0;
  for (int i = blockIdx.x * kTPB + threadIdx.x; i < I.np; i += kRedBlocks * kTPB) {
    const int idx = I.pidx[i];
    const double xv = x[idx], lb = I.lb[i], ub = I.ub[i];
    dual = fmax(dual, fabs(grad[idx] - zl[i] + zu[i]));
    if (!isinf(lb)) comp = fmax(comp, fabs((xv - lb) * zl[i] - mu));
    if (!isinf(ub)) comp = fmax(comp, fabs((ub - xv) * zu[i] - mu));
  }
  for (int j = blockIdx.x * kTPB + threadIdx.x; j < I.nc; j += kRedBlocks * kTPB)
    primal = fmax(primal, fabs(grad[I.cidx[j]] - I.lbh[j]));
  block_reduce(s, dual, Max());
  if (threadIdx.x == 0) part[3 * blockIdx.x] = s[0];
  __syncthreads();
  block_reduce(s, primal, Max());
  if (threadIdx.x == 0) part[3 * blockIdx.x + 1] = s[0];
  __syncthreads();
  block_reduce(s, comp, Max());
  if (threadIdx.x == 0) part[3 * blockIdx.x + 2] = s[0];
}

// Final stage: out[c] = reduction over blocks b of part[stride * b + c], in a
// fixed tree order (kind 0 sum, 1 min, 2 max per column).
__global__ void __launch_bounds__(kTPB) k_final(const double* __restrict__ part, int stride,
                                                const int* __restrict__ kind,
                                                double* __restrict__ out) {
  __shared__ double s[kTPB];
  for (int c = 0; c < stride; c++) {
    const double v = part[stride * threadIdx.x + c];
    const int k = kind[c];
    if (k == 0) block_reduce(s, v, Sum());
    else if (k == 1) block_reduce(s, v, Min());
    else block_reduce(s, v, Max());
    if (threadIdx.x == 0) out[c] = s[0];
    __syncthreads();
  }
}

// Final stage of k_max_step: (value, position) minima of the two columns.
__global__ void __launch_bounds__(kTPB) k_final_argmin(const double* __restrict__ part,
                                                       const int* __restrict__ ipart,
                                                       double* __restrict__ out,
                                                       int* __restrict__ iout) {
  __shared__ double sv[kTPB];
  __shared__ int si[kTPB];
  for (int q = 0; q < 2; q++) {
    sv[threadIdx.x] = part[2 * threadIdx.x + q];
    si[threadIdx.x] = ipart[2 * threadIdx.x + q];
    __syncthreads();
    for (int w = kTPB / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
        const double ov = sv[threadIdx.x + w];
        const int oi = si[threadIdx.x + w];
        if (ov < sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
          sv[threadIdx.x] = ov;
          si[threadIdx.x] = oi;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      out[q] = sv[0];
      iout[q] = si[0];
    }
    __syncthreads();
  }
}

template <class T>
struct DArr {
  T* p = nullptr;
  ~DArr() {
    if (p) (void)hipFree(p);
  }
  void put(const std::vector<T>& v) {
    ICHECK(hipMalloc(&p, sizeof(T) * std::max<size_t>(1, v.size())));
    if (!v.empty()) ICHECK(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  }
};

struct Handle {
  int n = 0, np = 0, nc = 0;
  std::vector<int> h_pidx;
  DArr<int> pidx, cidx, pof, cof, kinds, ipart, iout;
  DArr<double> lb, ub, lbh, part, out;
  Info info() const {
    return Info{n, np, nc, pidx.p, lb.p, ub.p, cidx.p, lbh.p, pof.p, cof.p};
  }
};

}  // namespace kktipm

struct kkt_ipm_handle {
  kktipm::Handle h;
};

namespace {

using kktipm::kRedBlocks;
using kktipm::kTPB;

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    kktipm::g_err = e.what();
    return -1;
  }
}

unsigned blocks_for(int count) { return (unsigned)std::max(1, (count + kTPB - 1) / kTPB); }

// run the final stage and bring `cols` doubles to the host
void finish(kktipm::Handle& h, int stride, const int* kinds_host, double* host,
            hipStream_t st) {
  std::vector<int> kinds(kinds_host, kinds_host + stride);
  ICHECK(hipMemcpyAsync(h.kinds.p, kinds.data(), sizeof(int) * stride, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(kktipm::k_final, dim3(1), dim3(kTPB), 0, st, h.part.p, stride, h.kinds.p,
                     h.out.p);
  ICHECK(hipGetLastError());
  ICHECK(hipMemcpyAsync(host, h.out.p, sizeof(double) * stride, hipMemcpyDeviceToHost, st));
  ICHECK(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

int kkt_ipm_create(int n, int num_primals, const int* primal_indices, const double* lbx,
                   const double* ubx, int num_constraints, const int* constraint_indices,
                   const double* lbh, kkt_ipm_t* out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("null argument");
    if (n < 0 || num_primals < 0 || num_constraints < 0) throw std::runtime_error("negative size");
    if ((num_primals > 0 && (!primal_indices || !lbx || !ubx)) ||
        (num_constraints > 0 && (!constraint_indices || !lbh)))
      throw std::runtime_error("null argument");
    auto H = std::make_unique<kkt_ipm_handle>();
    kktipm::Handle& h = H->h;
    h.n = n;
    h.np = num_primals;
    h.nc = num_constraints;
    std::vector<int> pof(std::max(n, 1), -1), cof(std::max(n, 1), -1);
    for (int i = 0; i < num_primals; i++) {
      const int idx = primal_indices[i];
      if (idx < 0 || idx >= n) throw std::runtime_error("primal index out of range");
      pof[idx] = i;  // the last occurrence writes last
    }
    for (int j = 0; j < num_constraints; j++) {
      const int idx = constraint_indices[j];
      if (idx < 0 || idx >= n) throw std::runtime_error("constraint index out of range");
      cof[idx] = j;
    }
    h.h_pidx.assign(primal_indices, primal_indices + num_primals);
    h.pidx.put(h.h_pidx);
    h.lb.put(std::vector<double>(lbx, lbx + num_primals));
    h.ub.put(std::vector<double>(ubx, ubx + num_primals));
    h.cidx.put(std::vector<int>(constraint_indices, constraint_indices + num_constraints));
    h.lbh.put(std::vector<double>(lbh, lbh + num_constraints));
    h.pof.put(pof);
    h.cof.put(cof);
    h.part.put(std::vector<double>(3 * (size_t)kRedBlocks, 0.0));
    h.ipart.put(std::vector<int>(2 * (size_t)kRedBlocks, 0));
    h.out.put(std::vector<double>(4, 0.0));
    h.iout.put(std::vector<int>(2, 0));
    h.kinds.put(std::vector<int>(4, 0));
    *out = H.release();
    return 0;
  });
}

int kkt_ipm_diagonal(kkt_ipm_t H, const double* x, const double* zl, const double* zu,
                     double* diag, void* stream) {
  return guarded([&] {
    if (!H || !diag || (H->h.np > 0 && (!x || !zl || !zu))) throw std::runtime_error("null argument");
    kktipm::Handle& h = H->h;
    if (h.n > 0)
      hipLaunchKernelGGL(kktipm::k_sigma, dim3(blocks_for(h.n)), dim3(kTPB), 0,
                         (hipStream_t)stream, h.info(), x, zl, zu, diag);
    ICHECK(hipGetLastError());
    return 0;
  });
}

int kkt_ipm_residual(kkt_ipm_t H, double mu, const double* x, const double* zl,
                     const double* zu, const double* grad, double* res, double* norm,
                     void* stream) {
  return guarded([&] {
    if (!H || !res || !grad || (H->h.np > 0 && (!x || !zl || !zu)))
      throw std::runtime_error("null argument");
    kktipm::Handle& h = H->h;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(kktipm::k_residual, dim3(kRedBlocks), dim3(kTPB), 0, st, h.info(), mu, x,
                       zl, zu, grad, res, h.part.p);
    ICHECK(hipGetLastError());
    const int kind[1] = {0};
    double s = 0.0;
    finish(h, 1, kind, &s, st);
    if (norm) *norm = std::sqrt(s);
    return 0;
  });
}

int kkt_ipm_bound_dual_step(kkt_ipm_t H, double mu, const double* x, const double* zl,
                            const double* zu, const double* px, double* dzl, double* dzu,
                            void* stream) {
  return guarded([&] {
    if (!H || (H->h.np > 0 && (!x || !zl || !zu || !px || !dzl || !dzu)))
      throw std::runtime_error("null argument");
    kktipm::Handle& h = H->h;
    if (h.np > 0)
      hipLaunchKernelGGL(kktipm::k_bound_dual, dim3(blocks_for(h.np)), dim3(kTPB), 0,
                         (hipStream_t)stream, h.info(), mu, x, zl, zu, px, dzl, dzu);
    ICHECK(hipGetLastError());
    return 0;
  });
}

int kkt_ipm_max_step(kkt_ipm_t H, double tau, const double* x, const double* zl,
                     const double* zu, const double* dx, const double* dzl,
                     const double* dzu, double* ax, int* xi, double* az, int* zi,
                     void* stream) {
  return guarded([&] {
    if (!H || (H->h.np > 0 && (!x || !zl || !zu || !dx || !dzl || !dzu)))
      throw std::runtime_error("null argument");
    kktipm::Handle& h = H->h;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(kktipm::k_max_step, dim3(kRedBlocks), dim3(kTPB), 0, st, h.info(), tau, x,
                       zl, zu, dx, dzl, dzu, h.part.p, h.ipart.p);
    hipLaunchKernelGGL(kktipm::k_final_argmin, dim3(1), dim3(kTPB), 0, st, h.part.p, h.ipart.p,
                       h.out.p, h.iout.p);
    ICHECK(hipGetLastError());
    double v[2];
    int iv[2];
    ICHECK(hipMemcpyAsync(v, h.out.p, sizeof(v), hipMemcpyDeviceToHost, st));
    ICHECK(hipMemcpyAsync(iv, h.iout.p, sizeof(iv), hipMemcpyDeviceToHost, st));
    ICHECK(hipStreamSynchronize(st));
    // the serial loop starts from 1 and keeps the first strictly smaller term
    const bool bx = v[0] < 1.0, bz = v[1] < 1.0;
    if (ax) *ax = bx ? v[0] : 1.0;
    if (xi) *xi = bx ? h.h_pidx[iv[0]] : -1;
    if (az) *az = bz ? v[1] : 1.0;
    if (zi) *zi = bz ? h.h_pidx[iv[1]] : -1;
    return 0;
  });
}

int kkt_ipm_complementarity(kkt_ipm_t H, const double* x, const double* zl,
                            const double* zu, double* avg, double* xi, void* stream) {
  return guarded([&] {
    if (!H || (H->h.np > 0 && (!x || !zl || !zu))) throw std::runtime_error("null argument");
    kktipm::Handle& h = H->h;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(kktipm::k_compl, dim3(kRedBlocks), dim3(kTPB), 0, st, h.info(), x, zl, zu,
                       h.part.p);
    ICHECK(hipGetLastError());
    const int kind[3] = {0, 0, 1};
    double r[3];
    finish(h, 3, kind, r, st);
    const double a = r[1] > 0.0 ? r[0] / r[1] : 0.0;
    double u = 1.0;
    if (a > 0.0) u = std::max(0.0, std::min(1.0, r[2] / a));
    if (avg) *avg = a;
    if (xi) *xi = u;
    return 0;
  });
}

int kkt_ipm_kkt_error(kkt_ipm_t H, double mu, const double* x, const double* zl,
                      const double* zu, const double* grad, double* dual, double* primal,
                      double* comp, void* stream) {
  return guarded([&] {
    if (!H || !grad || (H->h.np > 0 && (!x || !zl || !zu))) throw std::runtime_error("null argument");
    kktipm::Handle& h = H->h;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(kktipm::k_kkt_error, dim3(kRedBlocks), dim3(kTPB), 0, st, h.info(), mu, x,
                       zl, zu, grad, h.part.p);
    ICHECK(hipGetLastError());
    const int kind[3] = {2, 2, 2};
    double r[3];
    finish(h, 3, kind, r, st);
    if (dual) *dual = r[0];
    if (primal) *primal = r[1];
    if (comp) *comp = r[2];
    return 0;
  });
}

const char* kkt_ipm_last_error(void) { return kktipm::g_err.c_str(); }

void kkt_ipm_destroy(kkt_ipm_t h) { delete h; }

}  // extern "C"
