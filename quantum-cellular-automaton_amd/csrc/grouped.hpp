// Grouped launches: one kernel launch runs the same operation for many
// independent TDVP instances, each with its own operands and shapes.
//
// Every grouped kernel takes an array of records in device memory.  A record
// holds the arguments the single-instance launch would have taken plus that
// launch's grid (its "virtual" grid) and the first flat block of the record in
// the grouped launch.  The grouped kernel finds its record by binary search
// over the flat block index and runs the same device body as the
// single-instance kernel with the virtual block index, so every instance
// computes exactly what it computes alone.
#pragma once

#include <vector>

#include "kernels.hpp"

namespace qca {

struct VGrid {
  unsigned x = 1, y = 1, z = 1;
  unsigned blocks() const { return x * y * z; }
};

template <typename A>
struct GRec {
  A a;
  VGrid grid;
  unsigned begin;  // first flat block of this record
  unsigned pad;
};

// Buffers named in records are device (global) memory, but a pointer read
// from memory is generic to the compiler: it emits flat accesses, which count
// against both the vector-memory and the LDS wait counters and are ordered
// against LDS traffic.  Rebuilding the pointer in the global address space
// lets the grouped kernels use the same global accesses as the
// single-instance ones.  p must be uniform over the workgroup (a record's
// pointers are); the readfirstlane keeps the compiler from folding the
// round trip away.
template <typename T>
__device__ __forceinline__ T* gbl(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned long long u =
      ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(v >> 32)) << 32) |
      (unsigned)__builtin_amdgcn_readfirstlane((unsigned)v);
  return (T*)(__attribute__((address_space(1))) T*)u;
}

// Finds the record holding flat block `flat` (records sorted by begin) and
// the block's virtual index within it.
template <typename A>
__device__ __forceinline__ const GRec<A>& grec_find(const GRec<A>* __restrict__ recs, int n,
                                                   unsigned flat, unsigned& bx, unsigned& by,
                                                   unsigned& bz) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (recs[mid].begin <= flat) lo = mid; else hi = mid - 1;
  }
  const GRec<A>& r = recs[lo];
  unsigned local = flat - r.begin;
  bx = local % r.grid.x;
  local /= r.grid.x;
  by = local % r.grid.y;
  bz = local / r.grid.y;
  return r;
}

// Host staging of records: pinned host memory mirrored by a device buffer.
// push() appends a record list to the host side and returns its device
// address; upload() copies everything pushed since the last upload in one
// stream-ordered transfer.  reset() may only be called after the stream has
// drained the previous uploads.
class RecordArena {
 public:
  RecordArena() = default;
  ~RecordArena() { release(); }
  RecordArena(const RecordArena&) = delete;
  RecordArena& operator=(const RecordArena&) = delete;
  void reserve(size_t bytes);
  void release();
  template <typename R>
  const R* push(const std::vector<R>& recs) {
    return reinterpret_cast<const R*>(push_bytes(recs.data(), sizeof(R) * recs.size()));
  }
  void upload(hipStream_t s);
  void reset() { used_ = uploaded_ = 0; }
  size_t capacity() const { return cap_; }

 private:
  const void* push_bytes(const void* src, size_t bytes);
  unsigned char* host_ = nullptr;
  unsigned char* dev_ = nullptr;
  size_t cap_ = 0, used_ = 0, uploaded_ = 0;
};

// Assigns begin offsets to a record list; returns the total flat blocks.
template <typename A>
unsigned grec_layout(std::vector<GRec<A>>& recs) {
  unsigned b = 0;
  for (auto& r : recs) {
    r.begin = b;
    b += r.grid.blocks();
  }
  return b;
}

// ------------------------------------------------------------------ GEMM
template <typename T>
struct GemmArgs {
  GemmDesc d;
  const cplx<T>* A;
  const cplx<T>* B;
  cplx<T>* C;
  cplx<T>* ws;
  long a_jstride, b_jstride;   // A, B advance by j * a_jstride, j * b_jstride elements in launch(s, j)
  int splits, kps;  // kps: K-tiles per split
  int fused;        // one workgroup per output tile takes all splits (no partials, no reduce)
};

// One GEMM of one instance, with the split-K workspace it would use alone.
// a_jstride, b_jstride: the A (B) operand of iteration j of a repeated launch
// is A + j * a_jstride (B + j * b_jstride): a Krylov basis vector; 0 for a
// fixed operand.
template <typename T>
struct GemmJob {
  GemmDesc d;
  const cplx<T>* A;
  const cplx<T>* B;
  cplx<T>* C;
  cplx<T>* ws;
  long ws_elems;
  long b_jstride = 0;
  long a_jstride = 0;
};

template <typename R>
struct GLaunch {
  const GRec<R>* recs = nullptr;
  int n = 0;
  unsigned blocks = 0;
};

// A set of GEMMs (one per instance, any shapes) as at most sixteen grouped
// launches (one per kernel variant) plus one grouped split-K reduction.
// prepare() pushes the records; launch() may run any number of times once the
// arena has been uploaded.
template <typename T>
struct GemmGroup {
  GLaunch<GemmArgs<T>> launches[16];   // by (A m-contiguous, B n-contiguous, two K-groups, fused)
  GLaunch<GemmArgs<T>> reduce;
  void prepare(const std::vector<GemmJob<T>>& jobs, RecordArena& arena);
  void launch(hipStream_t s, int j = 0) const;
};

// ------------------------------------------------------------------ blocked QR steps
struct QrPanelArgs {
  c128* M;
  int rows, j0, pb;
  c128* tau;
  c128* V;
  c128* T;
  c128* VT;   // V T of the panel, laid out like V
};
struct QrIdentArgs {
  c128* Q;
  int rows, cols;
};
constexpr int QR_MAX_PANELS = 32;                       // 512 / 16
constexpr int QR_SLOT_FACTOR = 0;                       // 4 per panel
constexpr int QR_SLOT_IDENTITY = 4 * QR_MAX_PANELS;
constexpr int QR_SLOT_Q = QR_SLOT_IDENTITY + 1;         // 3 per panel, last panel first
struct QrStep {
  int slot;
  int kind;        // 0 panel factorisation, 1 GEMM, 2 identity start of Q
  QrPanelArgs panel;
  GemmJob<double> gemm;
  QrIdentArgs ident;
};
std::vector<QrStep> blocked_qr_steps(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols,
                                     c128* ws, size_t ws_elems);

// ------------------------------------------------------------------ SVD
// One launch_jacobi_svd call of one instance.
struct SvdJob {
  c128* X;
  int rows, cols;
  double* sigma;
  int* perm;
  SvdResult* res;
  int chi_max;
  double eps;
  int rank_cap;
  JacobiWork* w;
};
// Device and pinned host words for batch-wide convergence checks.
struct GroupFlags {
  int* dev = nullptr;
  int* host = nullptr;
  int cap = 0;
};
// launch_jacobi_svd for every job, grouped: the same kernels and rotations per
// job, the sweep rounds of all jobs in the same launches, one host check per
// chunk of sweeps for the whole group.  Leaves the results (SvdResult, sigma,
// perm, X) on the device.
void grouped_jacobi_svd(const std::vector<SvdJob>& jobs, RecordArena& arena, GroupFlags& flags,
                        hipStream_t s);

// ------------------------------------------------------------------ other kernels
// A list of records of one kernel, pushed to the arena.
template <typename A>
struct KGroup {
  GLaunch<A> l;
  void prepare(std::vector<GRec<A>>& recs, RecordArena& arena) {
    l.n = (int)recs.size();
    l.blocks = grec_layout(recs);
    l.recs = recs.empty() ? nullptr : arena.push(recs);
  }
  bool empty() const { return l.n == 0; }
};

template <typename A>
GRec<A> grec(const A& a, unsigned gx, unsigned gy = 1, unsigned gz = 1) {
  GRec<A> r{};
  r.a = a;
  r.grid.x = gx;
  r.grid.y = gy;
  r.grid.z = gz;
  return r;
}

struct ConvArgs {          // complex128 -> complex64 copy
  const c128* in;
  c64* out;
  long n;
  const int* skip;
};
struct ThetaSvdArgs {      // launch_theta_to_svd
  const c128* th;
  c128* X;
  int chil, chir, mode;
};
struct GatherIsoArgs {     // launch_gather_isometry
  const c128* X;
  const int* perm;
  const double* sigma;
  c128* A;
  int rows, chi_bond, k, mode;
};
template <typename T>
struct WmixArgs {          // launch_wmix
  const cplx<T>* Tin;
  long in_stride, nxc;
  int nin;
  cplx<T>* U;
  int nq, Dout;
  const int* rowptr;
  const int* col;
  const cplx<T>* val;
  const int* skip;
};

GRec<ConvArgs> conv_rec(const c128* in, c64* out, long n, const int* skip);
GRec<ThetaSvdArgs> theta_svd_rec(const c128* th, c128* X, int chil, int chir, int mode);
GRec<GatherIsoArgs> gather_iso_rec(const c128* X, const int* perm, const double* sigma, c128* A,
                                   int rows, int chi_bond, int k, int mode);
template <typename T>
GRec<WmixArgs<T>> wmix_rec(const cplx<T>* Tin, long in_stride, long nxc, int nin, cplx<T>* U, int nq,
                           int Dout, const WSparseDev<T>& w, const int* skip);
void launch_conv_grouped(const KGroup<ConvArgs>& g, hipStream_t s);
void launch_theta_svd_grouped(const KGroup<ThetaSvdArgs>& g, hipStream_t s);
void launch_gather_iso_grouped(const KGroup<GatherIsoArgs>& g, hipStream_t s);
// lds: the largest dynamic LDS of the records (wmix_lds_bytes of each)
template <typename T>
void launch_wmix_grouped(const KGroup<WmixArgs<T>>& g, size_t lds, hipStream_t s);
template <typename T>
size_t wmix_lds_bytes(int nin, int nq, int Dout);

// One Lanczos exponential per instance (the buffers of its LanczosWork).
struct LzArgs {
  LanczosCtrl* ctrl;
  const c128* x;          // input of lanczos_begin
  c128* y;                // output of lanczos_finish
  c128* V;
  c128* w64;
  c128* part;
  c128* part2;
  double* part_norm;
  c128* coef1;
  c128* coef2;
  const void* w;          // matvec output, cplx<TW>
  c64* v32;               // complex64 copy of each new basis vector (mixed path) or null
  long n;
  int nblk;               // lanczos_nblk(n)
  int max_m;
  double tau, tol;
};

// The Krylov exponentials of many instances in lockstep: the same kernels
// and reductions as lanczos_begin / lanczos_iteration_tail / lanczos_finish.
struct LzGroup {
  KGroup<LzArgs> nb, gf, one;   // grids lanczos_nblk(n), grid_for(n), 1
  void prepare(const std::vector<LzArgs>& jobs, RecordArena& arena);
  void begin(hipStream_t s) const;
  template <typename TW>
  void tail(int j, hipStream_t s) const;
  void finish(hipStream_t s) const;
  bool empty() const { return one.empty(); }
};

// out[i * words + w] = ((const unsigned long long*)src[i])[w], w < words:
// small device words of many instances gathered for one host transfer.
void launch_gather_words(const void* const* src, int words, int n, unsigned long long* out, hipStream_t s);

// One-launch Krylov exponentials, one workgroup each.  fused_krylov_place
// fills in the LDS placement of one exponential and returns its LDS bytes.
template <typename TE>
size_t fused_krylov_place(const FusedKrylovArgs<TE>& a, FusedKrylovArgs<TE>& placed);
template <typename TE>
void launch_fused_krylov_grouped(const KGroup<FusedKrylovArgs<TE>>& g, size_t lds, hipStream_t s);

}  // namespace qca
