// KKT value assembly: element Hessian contributions -> CSR values, on the GPU.
//
// The reference adds one Hessian-vector product per (element, column) into
// the CSR: serially with a binary search per entry (CSRMat::add_row,
// include/csr_matrix.h:555-567), by element colour with OpenMP
// (include/component_group.h:491-528), or with atomicAdd at precomputed
// positions on CUDA (include/cuda/component_group_backend.cuh:159-189); it
// then zeroes the rows and columns of fixed variables and puts 1 on their
// diagonal (include/fixed_variables.h:66-69).
//
// Here the host sorts every contribution by its CSR slot once per pattern
// (a segmented gather map).  The kernel then produces each CSR value with one
// thread that sums its slot's contributions in contribution order: writes are
// coalesced, no atomics are used, and the result is bitwise repeatable.
// Fixed rows and columns are folded into a per-slot mode byte.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "kkt_asm.h"

namespace kktasm {

static thread_local std::string g_err;

#define ACHECK(expr)                                                               \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e)); \
  } while (0)

enum : unsigned char { kSum = 0, kZero = 1, kOne = 2 };

// One thread per CSR slot s: vals[s] = sum of cv[order[q]], q in
// [sptr[s], sptr[s+1]), unless the slot belongs to a fixed row/column.
__global__ void k_gather_sum(int nnz, const int* __restrict__ sptr, const int* __restrict__ order,
                             const unsigned char* __restrict__ mode,
                             const double* __restrict__ cv, double* __restrict__ vals) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int a = sptr[s], b = sptr[s + 1];
  double v = 0.0;
  for (int q = a; q < b; q++) v += cv[order[q]];
  const unsigned char md = mode[s];
  vals[s] = md == kSum ? v : (md == kOne ? 1.0 : 0.0);
}

struct Handle {
  int n = 0, nnz = 0;
  long long ncontrib = 0, ndropped = 0;
  int* d_sptr = nullptr;
  int* d_order = nullptr;
  unsigned char* d_mode = nullptr;
  double* d_cv = nullptr;    // staging for host contributions
  double* d_vals = nullptr;  // staging for host output
  ~Handle() {
    if (d_sptr) (void)hipFree(d_sptr);
    if (d_order) (void)hipFree(d_order);
    if (d_mode) (void)hipFree(d_mode);
    if (d_cv) (void)hipFree(d_cv);
    if (d_vals) (void)hipFree(d_vals);
  }
};

}  // namespace kktasm

struct kkt_asm_handle {
  kktasm::Handle h;
};

extern "C" {

int kkt_asm_create(int n, const int* rowp, const int* cols, long long ncontrib,
                   const int* crow, const int* ccol, int nfixed, const int* fixed,
                   kkt_asm_t* out) {
  try {
    if (!rowp || !cols || !out || (ncontrib > 0 && (!crow || !ccol)) ||
        (nfixed > 0 && !fixed))
      throw std::runtime_error("null argument");
    if (n < 0 || ncontrib < 0 || nfixed < 0) throw std::runtime_error("negative size");
    if (ncontrib > INT_MAX) throw std::runtime_error("more than 2^31-1 contributions");
    auto H = std::make_unique<kkt_asm_handle>();
    kktasm::Handle& h = H->h;
    h.n = n;
    h.nnz = rowp[n];
    h.ncontrib = ncontrib;
    for (int r = 0; r < n; r++)
      if (rowp[r + 1] < rowp[r]) throw std::runtime_error("rowp not monotone");
    // CSR slot of each contribution; -1 when outside the pattern (dropped)
    std::vector<int> slot(ncontrib);
    for (long long q = 0; q < ncontrib; q++) {
      const int r = crow[q], c = ccol[q];
      if (r < 0 || r >= n || c < 0 || c >= n)
        throw std::runtime_error("contribution index out of range");
      const int* b = cols + rowp[r];
      const int* e = cols + rowp[r + 1];
      const int* it = std::lower_bound(b, e, c);
      slot[q] = (it != e && *it == c) ? (int)(it - cols) : -1;
      if (slot[q] < 0) h.ndropped++;
    }
    // counting sort by slot keeps contribution order inside each slot
    std::vector<int> sptr(h.nnz + 1, 0);
    for (long long q = 0; q < ncontrib; q++)
      if (slot[q] >= 0) sptr[slot[q] + 1]++;
    for (int s = 0; s < h.nnz; s++) sptr[s + 1] += sptr[s];
    std::vector<int> fill(sptr.begin(), sptr.end() - 1);
    std::vector<int> order(std::max<long long>(1, ncontrib - h.ndropped));
    for (long long q = 0; q < ncontrib; q++)
      if (slot[q] >= 0) order[fill[slot[q]]++] = (int)q;
    std::vector<unsigned char> isfix(n, 0), mode(std::max(1, h.nnz), kktasm::kSum);
    for (int i = 0; i < nfixed; i++)  // out-of-range indices are ignored, as the reference does
      if (fixed[i] >= 0 && fixed[i] < n) isfix[fixed[i]] = 1;
    for (int r = 0; r < n; r++)
      for (int p = rowp[r]; p < rowp[r + 1]; p++)
        if (isfix[r] || isfix[cols[p]])
          mode[p] = (r == cols[p] && isfix[r]) ? kktasm::kOne : kktasm::kZero;
    ACHECK(hipMalloc(&h.d_sptr, sizeof(int) * (h.nnz + 1)));
    ACHECK(hipMalloc(&h.d_order, sizeof(int) * order.size()));
    ACHECK(hipMalloc(&h.d_mode, mode.size()));
    ACHECK(hipMemcpy(h.d_sptr, sptr.data(), sizeof(int) * (h.nnz + 1), hipMemcpyHostToDevice));
    ACHECK(hipMemcpy(h.d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    ACHECK(hipMemcpy(h.d_mode, mode.data(), mode.size(), hipMemcpyHostToDevice));
    *out = H.release();
    return 0;
  } catch (const std::exception& e) {
    kktasm::g_err = e.what();
    return -1;
  }
}

int kkt_asm_values(kkt_asm_t H, const double* contrib_vals, double* vals, int mem,
                   void* stream) {
  try {
    if (!H || !vals || (H->h.ncontrib > 0 && !contrib_vals))
      throw std::runtime_error("null argument");
    if (mem != 0 && mem != 1) throw std::runtime_error("mem must be 0 (host) or 1 (device)");
    kktasm::Handle& h = H->h;
    hipStream_t st = (hipStream_t)stream;
    const double* cv = contrib_vals;
    double* dst = vals;
    if (mem == 0) {
      if (!h.d_cv) ACHECK(hipMalloc(&h.d_cv, sizeof(double) * std::max(1LL, h.ncontrib)));
      if (!h.d_vals) ACHECK(hipMalloc(&h.d_vals, sizeof(double) * std::max(1, h.nnz)));
      if (h.ncontrib)
        ACHECK(hipMemcpyAsync(h.d_cv, contrib_vals, sizeof(double) * h.ncontrib,
                              hipMemcpyHostToDevice, st));
      cv = h.d_cv;
      dst = h.d_vals;
    }
    constexpr int tpb = 256;
    if (h.nnz > 0) {
      const unsigned nb = (unsigned)((h.nnz + tpb - 1) / tpb);
      hipLaunchKernelGGL(kktasm::k_gather_sum, dim3(nb), dim3(tpb), 0, st, h.nnz, h.d_sptr,
                         h.d_order, h.d_mode, cv, dst);
      ACHECK(hipGetLastError());
    }
    if (mem == 0) {
      ACHECK(hipMemcpyAsync(vals, h.d_vals, sizeof(double) * h.nnz, hipMemcpyDeviceToHost, st));
      ACHECK(hipStreamSynchronize(st));
    }
    return 0;
  } catch (const std::exception& e) {
    kktasm::g_err = e.what();
    return -2;
  }
}

int kkt_asm_get_info(kkt_asm_t H, long long* nnz, long long* ncontrib, long long* ndropped) {
  if (!H) {
    kktasm::g_err = "null handle";
    return -1;
  }
  if (nnz) *nnz = H->h.nnz;
  if (ncontrib) *ncontrib = H->h.ncontrib;
  if (ndropped) *ndropped = H->h.ndropped;
  return 0;
}

const char* kkt_asm_last_error(void) { return kktasm::g_err.c_str(); }

void kkt_asm_destroy(kkt_asm_t h) { delete h; }

}  // extern "C"
