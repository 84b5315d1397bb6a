// Device-side data structures and kernel launchers of the TDVP path.
#pragma once

#include "qca_common.hpp"

#include <vector>

namespace qca {

// Maximum Krylov dimension (one wave holds the tridiagonal matrix).
constexpr int LZ_MAXK = 64;

// Lanczos control block, device resident.  Written by the reduction kernels,
// read back by the host only once per exponential.
struct LanczosCtrl {
  double beta0;
  double err;
  int done;
  int m;
  int zero;
  int hit_limit;
  double alpha[LZ_MAXK];
  double beta[LZ_MAXK];
  c128 z[LZ_MAXK];
  double wnorm2;      // |w|^2 of the iteration (two-level reduction of long vectors)
};

// Workspace of one Lanczos exponential (vectors of length <= max_n).
struct LanczosWork {
  LanczosCtrl* ctrl = nullptr;
  c128* V = nullptr;      // [LZ_MAXK][max_n] Krylov basis
  c128* w64 = nullptr;    // [max_n] orthogonalised residual
  c128* part = nullptr;   // [max_blocks][LZ_MAXK] partial projections
  c128* part2 = nullptr;  // same size: the second pass's partials (read and written apart)
  double* part_norm = nullptr;
  c128* coef1 = nullptr;  // [LZ_MAXK]
  c128* coef2 = nullptr;
  long max_n = 0;
  int max_blocks = 0;
  int max_m = LZ_MAXK;    // Krylov dimension limit of this workspace (<= LZ_MAXK)
};

// MPO tensor in compressed sparse rows over output channels: a channel is
// (physical, MPO bond) and the value of each stored entry is a W element.
template <typename T>
struct WSparseDev {
  int* rowptr = nullptr;
  int* col = nullptr;
  int* outv = nullptr;   // MPO bond index of each nonzero's row (row % group_rows)
  cplx<T>* val = nullptr;
  int nout = 0, nin = 0, nnz = 0;
  int group_rows = 0;    // rows per output physical index (the output MPO dimension)
  int group_max = 0;     // most nonzeros in any group of group_rows rows
  int ngroups = 0;       // nout / group_rows when at most SEG_STARTS_MAX, else 0
  int group_start[SEG_STARTS_MAX + 1] = {};  // rowptr[g * group_rows], g <= ngroups
};
// Fills ngroups and group_start of w from the host row pointer.
template <typename T>
inline void set_group_starts(WSparseDev<T>& w, const std::vector<int>& rowptr, int group_rows) {
  const int nout = (int)rowptr.size() - 1;
  w.ngroups = 0;
  if (group_rows <= 0 || nout % group_rows != 0 || nout / group_rows > SEG_STARTS_MAX) return;
  w.ngroups = nout / group_rows;
  for (int g = 0; g <= w.ngroups; ++g) w.group_start[g] = rowptr[(size_t)g * group_rows];
}

template <typename T>
void launch_wmix(const cplx<T>* Tin, long in_stride, long nxc, int nin, cplx<T>* U, int nq,
                 int Dout, const WSparseDev<T>& w, hipStream_t s, const int* skip = nullptr);

void launch_convert_c128_c64(const c128* in, c64* out, long n, hipStream_t s,
                             const int* skip = nullptr);
void launch_convert_c64_c128(const c64* in, c128* out, long n, hipStream_t s,
                             const int* skip = nullptr);
void launch_to_colmajor(const c128* in, c128* out, int rows, int cols, hipStream_t s);
void launch_site_to_rightqr(const c128* A, c128* X, int chil, int chir, hipStream_t s);
void launch_qr_gather_left(const c128* Q, const c128* X, c128* Anew, c128* C, int rows, int cols,
                           int qcols, int target, hipStream_t s);
void launch_qr_gather_right(const c128* Q, const c128* X, c128* Anew, c128* C, int chir, int cols,
                            int qcols, int target, hipStream_t s);
void launch_theta_to_svd(const c128* th, c128* X, int chil, int chir, int mode, hipStream_t s);
void launch_gather_isometry(const c128* X, const int* perm, const double* sigma, c128* A, int rows,
                            int chi_bond, int k, int mode, hipStream_t s);

int lanczos_nblk(long n);
const int* lanczos_done_flag(const LanczosWork& lw);
// v32 (optional): also write each new basis vector as complex64 there, the
// operand of the mixed path's next matvec.
void lanczos_begin(const c128* x, long n, LanczosWork& lw, hipStream_t s, c64* v32 = nullptr);
template <typename TW>
void lanczos_iteration_tail(const cplx<TW>* w, long n, int j, double tau, double tol,
                            LanczosWork& lw, hipStream_t s, c64* v32 = nullptr);
void lanczos_finish(long n, c128* y, LanczosWork& lw, hipStream_t s);

// ---------------------------------------------------------------- one-launch Krylov
// Longest vector of the one-workgroup exponential (4 residual elements per
// thread of 1024).
constexpr long FUSED_KRYLOV_MAX_N = 4096;

// Device-side running totals of the one-launch exponentials of one engine.
struct KrylovTally {
  long matvecs[3];          // by kind: 2-site, 1-site, bond
  long limit_hits;
  long calls;
  unsigned long long ticks; // summed kernel time in wall-clock ticks
  double flops;
};

// x <- exp(-i tau H_eff) x, H_eff of `kind` (0: 2-site x[q][a][c], q = (p1,p2);
// 1: 1-site x[p][a][c]; 2: bond x[a][c]) with environments L[a][w][a*] and
// R[c][v][c*] of element type TE and the sparse MPO (forward form of the
// merged pair for kind 0, of the site for kind 1; unused for kind 2).  The
// contractions run in TE, the recurrence in FP64 whatever TE is.
template <typename TE>
struct FusedKrylovArgs {
  int kind = 0;
  int chil = 1, chir = 1;
  int Din = 1, Dout = 1;
  long n = 0;
  const cplx<TE>* L = nullptr;
  const cplx<TE>* R = nullptr;
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const cplx<TE>* val = nullptr;
  int nnz = 0;                 // entries of the sparse MPO
  c128* x = nullptr;           // in/out, n elements
  c128* V = nullptr;           // basis, (max_m - 1) * n elements
  cplx<TE>* T1 = nullptr;      // dq * Din * chil * chir elements
  cplx<TE>* U = nullptr;       // dq * Dout * chil * chir elements
  LanczosCtrl* ctrl = nullptr; // optional: m, error, limit flag of this call
  KrylovTally* tally = nullptr;// optional running totals
  double tau = 0, tol = 0;
  double flops_per_matvec = 0;
  int max_m = LZ_MAXK;
  // Placement in the workgroup's LDS, filled by launch_fused_krylov: byte
  // offsets into the dynamic LDS (-1: the operand stays in global memory) and
  // the number of leading basis vectors held there.
  long offL = -1, offR = -1, offRow = -1, offCol = -1, offVal = -1, offU = -1, offT1 = -1;
  long offV = -1;
  long offW = 0;               // residual staging (n elements), always in LDS
  int vcap = 0;
};

template <typename TE>
void launch_fused_krylov(const FusedKrylovArgs<TE>& a, hipStream_t s);

// population, rounded population and single-site entropy of n sites from their
// 2x2 density matrices; out holds 3n doubles.
void launch_observables(const c128* rho, int n, double* out, hipStream_t s);
void launch_rho(const c128* A, const c128* T, long nxb, c128* rho, hipStream_t s);

int exact_rho_blocks(int ncells);
void launch_exact_rho(const c128* psi, int ncells, double* part, c128* rho, hipStream_t s);

template <typename T>
void launch_exact_apply(const cplx<T>* psi, cplx<T>* out, int ncells, int distance, int lo, int hi,
                        hipStream_t s);

// ---------------------------------------------------------------- SVD / QR
struct SvdResult {
  int k;              // truncation rank
  int nsweeps;        // Jacobi sweeps used
  int converged;
  int pad;
  double norm_kept;   // ||s[:k]||
  double tail;        // ||s[k:]||
};

// Workspace of the Jacobi SVD: device memory of at least
// jacobi_workspace_bytes(rows, cols) (control block, partial Grams, block
// rotations), a pinned host word pair, and the sweep count of the previous call
// (the number of sweeps launched before the first convergence check).
struct JacobiWork {
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* host = nullptr;
  int pred_sweeps = 8;
  // QR preconditioning (matrices wider than 32 columns): at least
  // jacobi_precond_elems(rows, cols) complex128, or null to skip it
  c128* pre = nullptr;
  size_t pre_elems = 0;
  int pred_sweeps_pre = 6;
  // rotation threshold relative to the column norms; 0: eps * sqrt(rows)
  double tol = 0;
};

struct JacobiLayout {
  int bs = 1, nb = 2, kspl = 1;
  long gpart_elems = 0, j_elems = 0;
  size_t bytes = 0;
};
JacobiLayout jacobi_layout(int rows, int cols);
size_t jacobi_workspace_bytes(int max_rows, int max_cols);
size_t jacobi_precond_elems(int max_rows, int max_cols);

// One-sided (block) Jacobi SVD of the column-major matrix X (rows x cols,
// FP64), in place.  On exit the columns are orthogonal; sigma (sorted
// descending), perm (column index of each sorted value) and the truncation
// result are written to device memory.
void launch_jacobi_svd(c128* X, int rows, int cols, double* sigma, int* perm, SvdResult* res,
                       int chi_max, double eps, int rank_cap, hipStream_t s, JacobiWork& w);

// Householder QR with LAPACK's zgeqr2/zung2r conventions (real beta, beta =
// -sign(Re alpha)*norm) of the column-major matrix M (rows x cols), in place:
// R on and above the diagonal, reflectors below it, tau[min(rows, cols)].
// Q's first qcols columns are written column-major to Q (skipped when null).
// Blocked Householder QR with LAPACK zgeqrf/zungqr semantics (same reflectors
// and signs as launch_householder_qr, panels of 16 columns, trailing updates
// and Q formation as GEMMs): M (rows <= 512 x cols, column-major) -> R on/above
// the diagonal, reflectors below, tau; Q (rows x qcols) when Q != null.
// ws holds at least qr_workspace_elems(rows, cols, qcols) complex128.
size_t qr_workspace_elems(int max_rows, int max_cols, int max_qcols);
void launch_blocked_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols, c128* ws,
                       size_t ws_elems, hipStream_t s);
// QR dispatch: the blocked version for min(rows, cols) > 32 when ws is large
// enough, else the single-workgroup one.
void launch_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols, c128* ws,
               size_t ws_elems, hipStream_t s);
void launch_householder_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols,
                           hipStream_t s);

}  // namespace qca
