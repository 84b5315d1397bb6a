// Shared definitions for the gfx950 TDVP kernels and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace qca {

// Complex numbers are stored interleaved (re, im), the layout of
// std::complex<T> and of torch.complex64 / complex128.
template <typename T>
struct cplx {
  T re, im;
};
using c64 = cplx<float>;
using c128 = cplx<double>;

// Return code of a numerical method that stopped at its iteration limit
// (QCA_ERR_NOT_CONVERGED in qca.h).
constexpr int QCA_ERR_NOT_CONVERGED_CODE = -4;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// On failure the runtime's sticky last-error is cleared before throwing, so
// that the error is reported once, through this library's return code.
#define QCA_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      (void)hipGetLastError();                                                     \
      throw ::qca::Error(-3, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    }                                                                              \
  } while (0)

#define QCA_CHECK(cond, msg)                                                       \
  do {                                                                             \
    if (!(cond)) throw ::qca::Error(-2, std::string("invalid argument: ") + (msg)); \
  } while (0)

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Raises the dynamic-LDS limit of kernel `fn` on the current device to at
// least `bytes` (once per kernel, device and size; safe from several threads).
inline void ensure_dynamic_lds(const void* fn, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> granted;
  int dev = 0;
  QCA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  size_t& g = granted[{fn, dev}];
  if (bytes > g) {
    QCA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    g = bytes;
  }
}

// ---------------------------------------------------------------- GEMM descriptor
constexpr int SEG_STARTS_MAX = 4;  // two-site blocks: 4 physical index pairs
// C[b](m,n) = alpha * sum_k opA(A)[b](m,k) * opB(B)[b](k,n) + beta * C[b](m,n)
// with every index split into at most two levels so that contractions over
// composite indices (e.g. (MPO bond, bond) or (physical, bond)) map onto one
// GEMM without a transpose copy.  k = k1*K2 + k2.  batch = b1*NB2 + b2.
//
// The K range is walked as segments of K2 (k1 = segment).  In gather form
// (seg_rows > 0) the segments of batch b2 (NB1 == 1) are instead the nonzeros
// e of rows [b2*seg_rows, (b2+1)*seg_rows) of a sparse matrix (CSR seg_rowptr,
// seg_col, seg_outv = row % seg_rows, seg_val): segment e contributes
//   sum_k2 seg_val[e] * A[m*a_m + seg_col[e]*a_k1 + k2*a_k2]
//                     * B[seg_outv[e]*b_k1 + k2*b_k2 + n*b_n],
// i.e. opA is a sparse mix of A's slices formed on the fly.  seg_max bounds
// the segments of any batch (K = seg_max * K2 for planning).
struct GemmDesc {
  int M, N, K, K2;  // K2 divides K
  int NB1, NB2;     // batch = NB1 * NB2
  long a_m, a_k1, a_k2, a_b1, a_b2;
  long b_k1, b_k2, b_n, b_b1, b_b2;
  long c_m, c_n, c_b1, c_b2;
  int conj_a, conj_b;
  double alpha_re, alpha_im;
  int accumulate;  // beta = 1 when set, else 0
  const int* skip; // device flag: launch is a no-op when *skip != 0 (may be null)
  int seg_rows, seg_max;
  const int* seg_rowptr;
  const int* seg_col;
  const int* seg_outv;
  const void* seg_val;  // cplx<T> of the GEMM's precision
  // seg_start[b2] = seg_rowptr[b2 * seg_rows] for b2 <= NB2 when seg_starts
  // is set (NB2 <= SEG_STARTS_MAX): the kernel then reads its segment range
  // from its arguments instead of from memory
  int seg_starts;
  int seg_start[SEG_STARTS_MAX + 1];
};

void gemm_c128(const GemmDesc& d, const c128* A, const c128* B, c128* C, hipStream_t s,
               c128* workspace, long workspace_elems);
void gemm_c64(const GemmDesc& d, const c64* A, const c64* B, c64* C, hipStream_t s,
              c64* workspace, long workspace_elems);

}  // namespace qca
