// FP64 factorisations of the TDVP split (gfx950).
//
// jacobi_svd: block one-sided (Hestenes) Jacobi on the columns of a
// column-major matrix X, with the rotations of each block pair computed on its
// Gram matrix.  The columns are cut into blocks of bs <= 16; the blocks meet in
// the round-robin tournament, one outer round per kernel triple:
//   jac_gram   G_p = X_p^H X_p for every block pair p (2bs columns), each
//              workgroup one chunk of 64 rows on the FP64 matrix cores,
//              partial Grams to a workspace;
//   jac_inner  one workgroup per pair sums its partial Grams and runs one
//              cyclic sweep of two-sided Jacobi rotations on G_p in LDS
//              (G <- R^H G R), accumulating J_p = product of the rotations;
//              two rounds of a sweep sweep all pairs of the 2bs columns,
//              the others only those across the two blocks;
//   jac_apply  X_p <- X_p J_p, each workgroup one chunk of rows, on the
//              FP64 matrix cores.
// Stream order separates the rounds.  The Gram is rebuilt from the columns in
// every round, so the test for a converged pair, |x_i^H x_j| <=
// tol*|x_i||x_j|, and the singular values (final column norms) carry the
// accuracy of the one-sided method: working precision relative to the largest
// singular value, which the 1e-10 truncation tail test needs.  A sweep without
// rotations above the convergence margin (jac_conv_tol) sets a device flag that
// turns the remaining launches into no-ops.
// A pair whose smaller column is below 16 eps of the matrix norm is left alone
// (JAC_NOISE2): such columns are rounding noise of the rotations, their
// contribution to any tail is < 4e-15, and rotating them against the signal
// columns of a matrix with more columns than rows would go on indefinitely.  jac_finalize sorts the singular values and evaluates
// the truncation rank (first i with ||s[i:]|| < eps, capped; reference
// algorithms/tdvp.py:289-292).
//
// householder_qr: zgeqr2 + zung2r semantics (reference tensor_networks/mps.py:86
// calls numpy.linalg.qr -> LAPACK), including LAPACK's zlarfg sign convention
// beta = -sign(Re alpha)*|x|, so R has a real diagonal and Q matches LAPACK's
// complete-QR complement columns.
#include "grouped.hpp"
#include "wave.hpp"

#include <algorithm>

namespace qca {
namespace {

constexpr int JAC_NT = 1024;         // threads of every Jacobi kernel
constexpr int JAC_MAXCOLS = 512;
constexpr int JAC_BS = 16;           // block size: pairs of 32 columns
constexpr int JAC_LMAX = 2 * JAC_BS;
constexpr int JAC_RC = 64;           // rows per Gram / apply chunk
constexpr int JAC_GRAM_NT = 256;     // jac_gram: one wave per 16x16 block of the 32x32 Gram
constexpr int JAC_APPLY_NT = 512;    // jac_apply: one wave per 16x16 block of the 32x64 product
constexpr int JAC_MAX_SWEEPS = 60;
constexpr int JAC_MAXKSPL = 8;       // row chunks of a Gram: rows <= 512
// Noise floor of a column, squared and relative to ||X||_F^2: (16 eps)^2.
// Columns below 16 eps ||X|| (3.6e-15 ||X||) are not orthogonalised further;
// their singular values are at the accuracy floor of the method, far below any
// truncation cutoff the engine is used with (1e-10).
constexpr double JAC_NOISE2 = 256.0 * 1.1102230246251565e-16 * 1.1102230246251565e-16;
// A pair is rotated when |g_ij| > tol sqrt(g_ii g_jj); a sweep counts as
// converged when no rotation had |g_ij| > conv sqrt(g_ii g_jj), conv =
// max(tol, 4 eps sqrt(rows)).  At the working-precision tolerance eps sqrt(rows)
// the inner products of freshly rotated columns carry rounding noise of about
// that size, so without the margin a pair can sit just above the threshold
// sweep after sweep; its last rotations are still applied.
__device__ __forceinline__ double jac_conv_tol(double tol, int rows) {
  return fmax(tol, 4.0 * 1.1102230246251565e-16 * sqrt((double)rows));
}

__device__ __forceinline__ double wsum(double v) { return wave_sum64(v); }

// Sum over rows r = lane, lane + 64, ... < rows (rows <= 512) of |col[r]|^2 in
// row order, with all of the lane's loads issued before the first add.
__device__ __forceinline__ double lane_norm2(const c128* __restrict__ col, int rows, int lane) {
  c128 v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (lane + 64 * i < rows) v[i] = col[lane + 64 * i];
  double a = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (lane + 64 * i < rows) a += v[i].re * v[i].re + v[i].im * v[i].im;
  return a;
}

// Sum of |x[e]|^2 over e = t, t + stride, ... < n in that order, eight loads
// in flight at a time.
__device__ __forceinline__ double strided_norm2(const c128* __restrict__ x, long n, long t, long stride) {
  double part = 0;
  for (long e0 = t; e0 < n; e0 += 8 * stride) {
    c128 v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (e0 + i * stride < n) v[i] = x[e0 + i * stride];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (e0 + i * stride < n) part += v[i].re * v[i].re + v[i].im * v[i].im;
  }
  return part;
}

// Position t of round r in the round-robin tournament over n (even) players.
__device__ __forceinline__ int rr_player(int t, int r, int n) {
  return (t == 0) ? 0 : ((t - 1 + r) % (n - 1)) + 1;
}

struct JacobiCtrl {
  double fro;                  // ||X||_F^2
  int done;                    // set by the first sweep without a rotation above jac_conv_tol
  int sweeps;                  // sweeps executed
  int rot[JAC_MAX_SWEEPS];     // rotations above jac_conv_tol per sweep
};

struct JacobiGeom {
  int rows, cols, bs, nb, kspl;
};

// Stages rows [r0, r0+nr) of the 2bs columns of pair p into a row-major LDS
// tile with a padded row stride (JAC_LMAX + 1) that spreads the banks; rows
// beyond nr and columns beyond 2bs (up to JAC_LMAX) are zero.  Lane bits give
// the row, the wave index walks the columns: no divisions, and each wave reads
// one contiguous 64-row run of a column.  NT is the workgroup size: all of a
// thread's loads are issued before its first LDS store, so the tile costs one
// memory latency, not one per column it holds.
template <int NT>
__device__ __forceinline__ void stage_tile(const c128* __restrict__ X, const JacobiGeom& g, int p,
                                           int round, int r0, int nr, c128* tile) {
  static_assert(JAC_RC == 64, "one row per lane");
  static_assert(NT % 64 == 0 && JAC_LMAX % (NT / 64) == 0, "whole columns per wave");
  constexpr int NW = NT / 64, IT = JAC_LMAX / NW;
  const int L = 2 * g.bs;
  const int bi = rr_player(p, round, g.nb), bj = rr_player(g.nb - 1 - p, round, g.nb);
  const int r = threadIdx.x & 63, w = threadIdx.x >> 6;
  c128 v[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int c = w + i * NW;
    const int gc = c < g.bs ? bi * g.bs + c : bj * g.bs + (c - g.bs);
    v[i] = (c < L && gc < g.cols && r < nr) ? X[(long)gc * g.rows + r0 + r] : c128{0.0, 0.0};
  }
#pragma unroll
  for (int i = 0; i < IT; ++i) tile[r * (JAC_LMAX + 1) + w + i * NW] = v[i];
}

// The pair products of a round on the FP64 matrix cores: v_mfma_f64_16x16x4f64
// takes A[m][k] from lane m + 16k and B[k][n] from lane n + 16k, and register
// r of lane l holds D[(l / 16) + 4r][l % 16].  A complex product is four real
// MFMAs.  One wave per 16x16 output tile, the whole K range in that wave, in
// order: the sums are deterministic.
typedef double jac_f64x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ jac_f64x4 jac_mfma(double a, double b, jac_f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ||X||_F^2 and reset of the control block (one workgroup).
__device__ __forceinline__ void jac_init_body(const c128* __restrict__ X, long n,
                                              JacobiCtrl* __restrict__ ctrl) {
  __shared__ double red[JAC_NT / 64];
  double part = 0;
  part = strided_norm2(X, n, threadIdx.x, JAC_NT);
  part = wsum(part);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < JAC_NT / 64; ++w) s += red[w];
    ctrl->fro = s;
    ctrl->done = 0;
    ctrl->sweeps = 0;
  }
  if (threadIdx.x < JAC_MAX_SWEEPS) ctrl->rot[threadIdx.x] = 0;
}
__global__ __launch_bounds__(JAC_NT) void jac_init(const c128* __restrict__ X, long n,
                                                   JacobiCtrl* __restrict__ ctrl) {
  jac_init_body(X, n, ctrl);
}

// Partial Gram of pair blockIdx.x over row chunk blockIdx.y:
// Gpart[p][kc](i, j) = sum_r conj(X(r, c_i)) X(r, c_j), stored column-major.
// Four waves, wave w computing the 16x16 block (j in 16(w/2).., i in 16(w%2)..)
// as D[j][i] = sum_r X(r, c_j) conj(X(r, c_i)) over the chunk's 64 rows.
__device__ __forceinline__ void jac_gram_body(const c128* __restrict__ X, const JacobiGeom& g, int round,
                                              c128* __restrict__ Gpart,
                                              const JacobiCtrl* __restrict__ ctrl, int p, int kc) {
  // the control word is read alongside the tile, not before it: the round
  // kernels are latency bound, and a no-op round only stages a tile
  const int done = ctrl->done;
  __shared__ c128 tile[JAC_RC * (JAC_LMAX + 1)];
  const int L = 2 * g.bs;
  const int r0 = kc * JAC_RC, nr = min(JAC_RC, g.rows - r0);
  stage_tile<JAC_GRAM_NT>(X, g, p, round, r0, nr, tile);
  __syncthreads();
  if (done) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tj = (wave >> 1) * 16, ti = (wave & 1) * 16;
  if (tj >= L || ti >= L) return;
  const int q = lane >> 4, c = lane & 15;
  jac_f64x4 gre = {0, 0, 0, 0}, gim = {0, 0, 0, 0};
  const int kmax = (nr + 3) & ~3;
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    const c128 a = tile[(k0 + q) * (JAC_LMAX + 1) + tj + c];   // X(r, c_j)
    const c128 b = tile[(k0 + q) * (JAC_LMAX + 1) + ti + c];   // X(r, c_i), conjugated below
    gre = jac_mfma(a.re, b.re, gre);
    gre = jac_mfma(a.im, b.im, gre);
    gim = jac_mfma(a.im, b.re, gim);
    gim = jac_mfma(-a.re, b.im, gim);
  }
  c128* out = Gpart + ((long)p * g.kspl + kc) * L * L;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int j = tj + q + 4 * rr, i = ti + c;
    if (i < L && j < L) out[i + j * L] = c128{gre[rr], gim[rr]};
  }
}
__global__ __launch_bounds__(JAC_GRAM_NT) void jac_gram(const c128* __restrict__ X, JacobiGeom g,
                                                   int round, c128* __restrict__ Gpart,
                                                   const JacobiCtrl* __restrict__ ctrl) {
  jac_gram_body(X, g, round, Gpart, ctrl, blockIdx.x, blockIdx.y);
}

// One cyclic sweep of two-sided Jacobi on the Gram of pair blockIdx.x.
// Rotation of columns (i, j), with g = G(i, j) = x_i^H x_j, d = |x_j|^2 - |x_i|^2,
// h = sqrt(d^2 + 4|g|^2), w = h + |d|:
//   x_i' = c x_i - conj(s e) x_j,   x_j' = (s e) x_i + c x_j,
//   c = sqrt(w / (2h)),  s e = sign(d) g sqrt(2 / (h w)),
// the Rutishauser rotation t = sign(d)/(|zeta| + sqrt(1 + zeta^2)), zeta = d/(2|g|),
// c = 1/sqrt(1 + t^2), s = t c, rewritten so that c^2 + |s e|^2 = 1 exactly in
// real arithmetic on non-negative operands; c and |s| are evaluated as w y and
// 2 y with y = 1/sqrt(2hw), so the critical path is sqrt -> rsqrt.
// G <- R^H G R and J <- J R.
// Per round: wave 0 computes the rotations of round r while the other waves
// apply round r-1 to J; after a barrier all threads apply round r to G.
// The pair indices are recomputed in registers, the rotation of a pair is one
// 32-byte LDS record.
struct alignas(16) JRot {
  double c, sr, si;   // c, s e
  int act, pad;
};
static_assert(JAC_NT >= 64 + JAC_LMAX * (JAC_LMAX / 2), "jac_inner needs one thread per J item");
static_assert(JAC_NT >= (JAC_LMAX / 2) * (JAC_LMAX / 2), "jac_inner needs one thread per G block");

// Round-robin position rr_player(t, r, n) without integer division (t < n, r < n - 1).
__device__ __forceinline__ int rr_fast(int t, int r, int n) {
  if (t == 0) return 0;
  int x = t - 1 + r;
  x -= (x >= n - 1) ? (n - 1) : 0;
  return x + 1;
}
// Pair t of inner round r.  Full ordering: the round-robin tournament over all
// L = 2H columns (L - 1 rounds, every pair once).  Cross ordering: column t of
// the first block meets column H + (t + r) mod H of the second (H rounds, only
// the pairs across the two blocks).
template <bool CROSS>
__device__ __forceinline__ int pair_first(int t, int r, int L) {
  return CROSS ? t : rr_fast(t, r, L);
}
template <bool CROSS>
__device__ __forceinline__ int pair_second(int t, int r, int L) {
  if (!CROSS) return rr_fast(L - 1 - t, r, L);
  const int H = L >> 1;
  int x = t + r;
  x -= (x >= H) ? H : 0;
  return H + x;
}
// One cyclic sweep of the rotations on G (L x L, column-major, in LDS),
// accumulating J <- J R; rot is a 2 x (L/2) LDS scratch.  CROSS restricts the
// sweep to the pairs across the two halves (see pair_first).  Needs blockDim.x
// == JAC_NT and G, J visible to the whole workgroup on entry (it ends with a
// barrier).  Returns the rotations this thread computed.
template <bool CROSS>
__device__ int inner_sweep(c128* G, c128* J, JRot (*rot)[JAC_LMAX / 2], int L, int H, double tol2,
                           double conv2, double negl) {
  const int tid = threadIdx.x;
  const int NR = CROSS ? H : L - 1;
  int mine = 0;
  // fixed roles: lanes < H of wave 0 compute the rotations; threads 64.. own
  // one J item (row r, pair Q) each; threads < H*H own one 2x2 block of G
  const int jit = tid - 64;
  const bool has_j = jit >= 0 && jit < L * H;
  const int jr = has_j ? jit % L : 0, jq = has_j ? jit / L : 0;
  const bool has_g = tid < H * H;
  const int gP = tid % H, gQ = tid / H;
  for (int round = 0; round <= NR; ++round) {
    const int cur = round & 1, prev = cur ^ 1;
    if (tid < 64) {
      if (round < NR && tid < H) {
        const int i = pair_first<CROSS>(tid, round, L), j = pair_second<CROSS>(tid, round, L);
        const double a = G[i * L + i].re, b = G[j * L + j].re;
        const c128 gij = G[j * L + i];
        const double ag2 = gij.re * gij.re + gij.im * gij.im;
        // a pair with a column at the noise floor is done: rotating noise
        // against a signal column only re-creates noise, and for matrices with
        // more columns than rows such pairs can rotate indefinitely
        const int act = !(ag2 <= tol2 * a * b || ag2 < 1e-300 || fmin(a, b) < negl);
        JRot q{1.0, 0.0, 0.0, act, 0};
        if (act) {
          const double d = b - a;
          const double h = sqrt(d * d + 4.0 * ag2);
          const double w = h + fabs(d);
          // c = sqrt(w / 2h) = w y and |s| = sqrt(2 / hw) = 2 y with y =
          // 1/sqrt(2hw): one reciprocal square root in place of two divides
          // and two square roots on the serial path of the round.  Hardware
          // estimate, two Newton steps, and a last one whose residual
          // 1 - t y^2 includes the exact error of t y, so that y carries no
          // bias (c and s share y: a biased y drifts the column norms)
          const double t = 2.0 * h * w;
          double y = __builtin_amdgcn_rsq(t);
          y = y * fma(-0.5 * t * y, y, 1.5);
          y = y * fma(-0.5 * t * y, y, 1.5);
          const double ty = t * y;
          const double e = fma(-ty, y, 1.0) - fma(t, y, -ty) * y;
          y = fma(0.5 * y, e, y);
          q.c = w * y;
          const double s = (d >= 0 ? 2.0 : -2.0) * y;
          q.sr = s * gij.re;
          q.si = s * gij.im;
          // low half: rotations applied; high half: rotations above the
          // convergence threshold, which keep the sweeps going
          mine += ag2 > conv2 * a * b ? 0x10001 : 1;
        }
        rot[cur][tid] = q;
      }
    } else if (round > 0 && has_j) {
      const int qi = pair_first<CROSS>(jq, round - 1, L), qj = pair_second<CROSS>(jq, round - 1, L);
      const c128 x = J[qi * L + jr], y = J[qj * L + jr];
      const JRot Q = rot[prev][jq];
      if (Q.act) {
        J[qi * L + jr] = c128{Q.c * x.re - (Q.sr * y.re + Q.si * y.im), Q.c * x.im - (Q.sr * y.im - Q.si * y.re)};
        J[qj * L + jr] = c128{(Q.sr * x.re - Q.si * x.im) + Q.c * y.re, (Q.sr * x.im + Q.si * x.re) + Q.c * y.im};
      }
    }
    __syncthreads();
    if (round == NR) break;
    if (has_g) {
      const int pi = pair_first<CROSS>(gP, round, L), pj = pair_second<CROSS>(gP, round, L);
      const int qi = pair_first<CROSS>(gQ, round, L), qj = pair_second<CROSS>(gQ, round, L);
      const c128 t00 = G[qi * L + pi], t10 = G[qi * L + pj], t01 = G[qj * L + pi], t11 = G[qj * L + pj];
      const JRot P = rot[cur][gP], Q = rot[cur][gQ];
      if (P.act || Q.act) {
        const c128 u00{Q.c * t00.re - (Q.sr * t01.re + Q.si * t01.im), Q.c * t00.im - (Q.sr * t01.im - Q.si * t01.re)};
        const c128 u10{Q.c * t10.re - (Q.sr * t11.re + Q.si * t11.im), Q.c * t10.im - (Q.sr * t11.im - Q.si * t11.re)};
        const c128 u01{(Q.sr * t00.re - Q.si * t00.im) + Q.c * t01.re, (Q.sr * t00.im + Q.si * t00.re) + Q.c * t01.im};
        const c128 u11{(Q.sr * t10.re - Q.si * t10.im) + Q.c * t11.re, (Q.sr * t10.im + Q.si * t10.re) + Q.c * t11.im};
        G[qi * L + pi] = c128{P.c * u00.re - (P.sr * u10.re - P.si * u10.im), P.c * u00.im - (P.sr * u10.im + P.si * u10.re)};
        G[qi * L + pj] = c128{(P.sr * u00.re + P.si * u00.im) + P.c * u10.re, (P.sr * u00.im - P.si * u00.re) + P.c * u10.im};
        G[qj * L + pi] = c128{P.c * u01.re - (P.sr * u11.re - P.si * u11.im), P.c * u01.im - (P.sr * u11.im + P.si * u11.re)};
        G[qj * L + pj] = c128{(P.sr * u01.re + P.si * u01.im) + P.c * u11.re, (P.sr * u01.im - P.si * u01.re) + P.c * u11.im};
      }
    }
    __syncthreads();
  }
  return mine;
}

// Outer rounds 0 and (nb - 1) / 2 of a sweep run the full inner ordering, which
// also visits the pairs inside each block (every block takes part in every
// outer round); the other rounds visit only the pairs across their two blocks.
// A sweep thus meets every column pair at least once, in 2 * 31 + 13 * 16 inner
// rounds for 16 blocks of 16 instead of 15 * 31.  Inside pairs once per sweep
// (round 0 only) cost about one sweep more; twice keeps the sweep count of the
// full ordering within a fraction of a sweep.
__device__ __forceinline__ void jac_inner_body(const JacobiGeom& g, int round, int sweep,
                                               const c128* __restrict__ Gpart,
                                               c128* __restrict__ Jout, int* __restrict__ rotated,
                                               double tol, JacobiCtrl* __restrict__ ctrl, int p) {
  const int done = ctrl->done;  // read alongside the partial Grams
  constexpr int HM = JAC_LMAX / 2;
  __shared__ c128 G[JAC_LMAX * JAC_LMAX];
  __shared__ c128 J[JAC_LMAX * JAC_LMAX];
  __shared__ JRot rot[2][HM];
  __shared__ int total, total_sig;
  const int L = 2 * g.bs, H = g.bs, tid = threadIdx.x;
  const double negl = JAC_NOISE2 * ctrl->fro;
  const double conv = jac_conv_tol(tol, g.rows);
  static_assert(JAC_NT >= JAC_LMAX * JAC_LMAX, "one Gram entry per thread");
  if (tid < L * L) {
    // all chunk partials are loaded before the first add (one memory latency),
    // then summed in chunk order
    c128 v[JAC_MAXKSPL];
#pragma unroll
    for (int kc = 0; kc < JAC_MAXKSPL; ++kc)
      if (kc < g.kspl) v[kc] = Gpart[((long)p * g.kspl + kc) * L * L + tid];
    double re = 0, im = 0;
#pragma unroll
    for (int kc = 0; kc < JAC_MAXKSPL; ++kc)
      if (kc < g.kspl) {
        re += v[kc].re;
        im += v[kc].im;
      }
    G[tid] = c128{re, im};
    J[tid] = c128{(tid % L) == (tid / L) ? 1.0 : 0.0, 0.0};
  }
  if (tid == 0) {
    total = 0;
    total_sig = 0;
  }
  __syncthreads();
  if (done) return;
  const bool cross = round != 0 && round != (g.nb - 1) / 2;
  const int mine = cross ? inner_sweep<true>(G, J, rot, L, H, tol * tol, conv * conv, negl)
                         : inner_sweep<false>(G, J, rot, L, H, tol * tol, conv * conv, negl);
  if (mine & 0xffff) atomicAdd(&total, mine & 0xffff);
  if (mine >> 16) atomicAdd(&total_sig, mine >> 16);
  __syncthreads();
  c128* jo = Jout + (long)p * L * L;
  for (int e = tid; e < L * L; e += JAC_NT) jo[e] = J[e];
  if (tid == 0) {
    rotated[p] = total;
    if (total_sig) atomicAdd(&ctrl->rot[sweep], total_sig);
  }
}
// Two workgroups per CU (64 VGPRs): grouped launches hold a pair per workgroup.
__global__ __launch_bounds__(JAC_NT) __attribute__((amdgpu_waves_per_eu(8))) void jac_inner(JacobiGeom g, int round, int sweep,
                                                    const c128* __restrict__ Gpart,
                                                    c128* __restrict__ Jout,
                                                    int* __restrict__ rotated, double tol,
                                                    JacobiCtrl* __restrict__ ctrl) {
  jac_inner_body(g, round, sweep, Gpart, Jout, rotated, tol, ctrl, blockIdx.x);
}

// X_p <- X_p J_p on row chunk blockIdx.y of pair blockIdx.x; chunks from
// g.kspl on belong to the second matrix Z (geometry gz), which receives the
// same rotations (gz.kspl == 0 when there is none).
__device__ __forceinline__ void jac_apply_body(c128* __restrict__ Xa, const JacobiGeom& ga,
                                               c128* __restrict__ Za, const JacobiGeom& gz, int round,
                                               const c128* __restrict__ Jin,
                                               const int* __restrict__ rotated,
                                               const JacobiCtrl* __restrict__ ctrl, int p, int by) {
  // the control words are read alongside the tile and J, not before them
  const int skip = ctrl->done | (rotated[p] == 0);
  const bool second = by >= ga.kspl;
  c128* __restrict__ X = second ? Za : Xa;
  const JacobiGeom g = second ? gz : ga;
  const int kc = second ? by - ga.kspl : by;
  __shared__ c128 tile[JAC_RC * (JAC_LMAX + 1)];
  __shared__ c128 J[JAC_LMAX * (JAC_LMAX + 1)];   // J(k, c) at c * (JAC_LMAX + 1) + k, zero-padded
  const int L = 2 * g.bs;
  const int r0 = kc * JAC_RC, nr = min(JAC_RC, g.rows - r0);
  stage_tile<JAC_APPLY_NT>(X, g, p, round, r0, nr, tile);
  const c128* jp = Jin + (long)p * L * L;
  constexpr int JIT = JAC_LMAX * JAC_LMAX / JAC_APPLY_NT;
  c128 jv[JIT];
#pragma unroll
  for (int i = 0; i < JIT; ++i) {
    const int e = threadIdx.x + i * JAC_APPLY_NT, k = e % JAC_LMAX, c = e / JAC_LMAX;
    jv[i] = (k < L && c < L) ? jp[c * L + k] : c128{0.0, 0.0};
  }
#pragma unroll
  for (int i = 0; i < JIT; ++i) {
    const int e = threadIdx.x + i * JAC_APPLY_NT, k = e % JAC_LMAX, c = e / JAC_LMAX;
    J[c * (JAC_LMAX + 1) + k] = jv[i];
  }
  __syncthreads();
  if (skip) return;
  // Eight waves: wave w computes the 16x16 block (columns c in 16(w%2)..,
  // rows r in 16(w/2)..) of the transposed product D[c][r] = sum_k J(k, c) X(r, k),
  // so that consecutive lanes store consecutive rows of a column.
  const int bi = rr_player(p, round, g.nb), bj = rr_player(g.nb - 1 - p, round, g.nb);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tc = (wave & 1) * 16, tr = (wave >> 1) * 16;
  if (tc >= L || tr >= nr) return;
  const int q = lane >> 4, cl = lane & 15;
  jac_f64x4 dre = {0, 0, 0, 0}, dim = {0, 0, 0, 0};
  const int kmax = (L + 3) & ~3;
  for (int k0 = 0; k0 < kmax; k0 += 4) {
    const c128 a = J[(tc + cl) * (JAC_LMAX + 1) + k0 + q];      // J(k, c)
    const c128 b = tile[(tr + cl) * (JAC_LMAX + 1) + k0 + q];   // X(r, k)
    dre = jac_mfma(a.re, b.re, dre);
    dre = jac_mfma(-a.im, b.im, dre);
    dim = jac_mfma(a.re, b.im, dim);
    dim = jac_mfma(a.im, b.re, dim);
  }
  const int r = tr + cl;
  if (r < nr)
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int c = tc + q + 4 * rr;
      const int gc = c < g.bs ? bi * g.bs + c : bj * g.bs + (c - g.bs);
      if (c < L && gc < g.cols) X[(long)gc * g.rows + r0 + r] = c128{dre[rr], dim[rr]};
    }
}
__global__ __launch_bounds__(JAC_APPLY_NT) void jac_apply(c128* __restrict__ Xa, JacobiGeom ga,
                                                    c128* __restrict__ Za, JacobiGeom gz, int round,
                                                    const c128* __restrict__ Jin,
                                                    const int* __restrict__ rotated,
                                                    const JacobiCtrl* __restrict__ ctrl) {
  jac_apply_body(Xa, ga, Za, gz, round, Jin, rotated, ctrl, blockIdx.x, blockIdx.y);
}

// Closes a sweep: converged when it made no rotation.
__global__ void jac_check(JacobiCtrl* __restrict__ ctrl, int sweep) {
  if (threadIdx.x != 0 || ctrl->done) return;
  ctrl->sweeps = sweep + 1;
  if (ctrl->rot[sweep] == 0) ctrl->done = 1;
}

// Column norms, descending sort and truncation rank (one workgroup of JAC_NT).
__device__ void finalize_body(const c128* __restrict__ X, int rows, int cols,
                              double* __restrict__ sigma, int* __restrict__ perm,
                              SvdResult* __restrict__ res, int chi_max, double eps, int rank_cap,
                              int sweeps, int converged) {
  __shared__ double skey[JAC_MAXCOLS];
  __shared__ int sidx[JAC_MAXCOLS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = JAC_NT / 64;
  int np2 = 1;
  while (np2 < cols) np2 <<= 1;
  for (int c = wave; c < np2; c += nw) {
    double a = (c < cols) ? lane_norm2(X + (long)c * rows, rows, lane) : 0.0;
    a = wsum(a);
    if (lane == 0) {
      skey[c] = (c < cols) ? sqrt(a) : -1.0;
      sidx[c] = c;
    }
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid; e < np2; e += JAC_NT) {
        const int q = e ^ stride;
        if (q > e) {
          const bool desc = ((e & size) == 0);
          const double ke = skey[e], kq = skey[q];
          const bool swap = desc ? (ke < kq || (ke == kq && sidx[e] > sidx[q]))
                                 : (ke > kq || (ke == kq && sidx[e] < sidx[q]));
          if (swap) {
            skey[e] = kq; skey[q] = ke;
            const int ti = sidx[e]; sidx[e] = sidx[q]; sidx[q] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < cols; e += JAC_NT) {
    sigma[e] = skey[e];
    perm[e] = sidx[e];
  }
  __syncthreads();
  if (tid == 0) {
    // first i (< cap) with ||s[i:]|| < eps over the min(rows, cols) values a
    // reduced SVD returns; the tail norms form a non-increasing sequence
    const int ns = rank_cap;
    const int cap = min(chi_max, ns);
    int k = cap;
    double suffix = 0;
    for (int e = ns - 1; e >= 0; --e) {
      suffix += skey[e] * skey[e];
      if (e < cap && sqrt(suffix) < eps) k = e;
    }
    double kept = 0, tail = 0;
    for (int e = 0; e < k; ++e) kept += skey[e] * skey[e];
    for (int e = k; e < ns; ++e) tail += skey[e] * skey[e];
    res->k = k;
    res->nsweeps = sweeps;
    res->converged = converged;
    res->norm_kept = sqrt(kept);
    res->tail = sqrt(tail);
  }
}

__global__ __launch_bounds__(JAC_NT) void jac_finalize(const c128* __restrict__ X, int rows, int cols,
                                                       double* __restrict__ sigma,
                                                       int* __restrict__ perm,
                                                       SvdResult* __restrict__ res, int chi_max,
                                                       double eps, int rank_cap,
                                                       const JacobiCtrl* __restrict__ ctrl) {
  finalize_body(X, rows, cols, sigma, perm, res, chi_max, eps, rank_cap, ctrl->sweeps, ctrl->done);
}

// The whole SVD of a matrix with at most JAC_LMAX columns in one workgroup:
// sweeps of (Gram from the columns -> one rotation sweep on it -> X <- X J)
// until a sweep makes no rotation, then finalize_body.  Same rotations, test
// and accuracy as the kernel triple, without launches between rounds.
__device__ __forceinline__ void jac_small_body(c128* __restrict__ X, int rows, int cols,
                                               double* __restrict__ sigma, int* __restrict__ perm,
                                               SvdResult* __restrict__ res, int chi_max, double eps,
                                               int rank_cap, double tol) {
  constexpr int HM = JAC_LMAX / 2;
  __shared__ c128 tile[JAC_RC * (JAC_LMAX + 1)];
  __shared__ c128 G[JAC_LMAX * JAC_LMAX];
  __shared__ c128 J[JAC_LMAX * JAC_LMAX];
  __shared__ JRot rot[2][HM];
  __shared__ double red[JAC_NT / 64];
  __shared__ int total, total_sig;
  const int tid = threadIdx.x;
  const int H = (cols + 1) / 2, L = 2 * H;
  const double conv = jac_conv_tol(tol, rows);
  double part = 0;
  part = strided_norm2(X, (long)rows * cols, tid, JAC_NT);
  part = wsum(part);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  double fro = 0;
  for (int w = 0; w < JAC_NT / 64; ++w) fro += red[w];
  const double negl = JAC_NOISE2 * fro;
  const JacobiGeom g{rows, cols, H, 2, 1};  // one block pair: columns 0..H-1 and H..L-1
  int sweep = 0, converged = 0;
  for (; sweep < JAC_MAX_SWEEPS; ++sweep) {
    // Gram from the current columns, 64-row chunks through the tile
    double gre = 0, gim = 0;
    const int gi = tid % L, gj = tid / L;
    for (int r0 = 0; r0 < rows; r0 += JAC_RC) {
      const int nr = min(JAC_RC, rows - r0);
      __syncthreads();
      stage_tile<JAC_NT>(X, g, 0, 0, r0, nr, tile);
      __syncthreads();
      if (tid < L * L)
        for (int r = 0; r < nr; ++r) {
          const c128 u = tile[r * (JAC_LMAX + 1) + gi], v = tile[r * (JAC_LMAX + 1) + gj];
          gre += u.re * v.re + u.im * v.im;
          gim += u.re * v.im - u.im * v.re;
        }
    }
    if (tid < L * L) {
      G[tid] = c128{gre, gim};
      J[tid] = c128{gi == gj ? 1.0 : 0.0, 0.0};
    }
    if (tid == 0) {
      total = 0;
      total_sig = 0;
    }
    __syncthreads();
    const int mine = inner_sweep<false>(G, J, rot, L, H, tol * tol, conv * conv, negl);
    if (mine & 0xffff) atomicAdd(&total, mine & 0xffff);
    if (mine >> 16) atomicAdd(&total_sig, mine >> 16);
    __syncthreads();
    if (total == 0) {
      converged = 1;
      break;
    }
    // X <- X J, chunk by chunk (each chunk is read into the tile before it is written)
    for (int r0 = 0; r0 < rows; r0 += JAC_RC) {
      const int nr = min(JAC_RC, rows - r0);
      stage_tile<JAC_NT>(X, g, 0, 0, r0, nr, tile);
      __syncthreads();
      const int r = tid & 63;
      if (r < nr)
        for (int c = tid >> 6; c < cols; c += JAC_NT / 64) {
          double re = 0, im = 0;
          for (int k = 0; k < L; ++k) {
            const c128 x = tile[r * (JAC_LMAX + 1) + k], jj = J[c * L + k];
            re += x.re * jj.re - x.im * jj.im;
            im += x.re * jj.im + x.im * jj.re;
          }
          X[(long)c * rows + r0 + r] = c128{re, im};
        }
      __syncthreads();
    }
    if (total_sig == 0) {
      converged = 1;
      break;
    }
  }
  __syncthreads();
  finalize_body(X, rows, cols, sigma, perm, res, chi_max, eps, rank_cap,
                converged ? sweep + 1 : sweep, converged);
}
__global__ __launch_bounds__(JAC_NT) void jac_small(c128* __restrict__ X, int rows, int cols,
                                                    double* __restrict__ sigma,
                                                    int* __restrict__ perm,
                                                    SvdResult* __restrict__ res, int chi_max,
                                                    double eps, int rank_cap, double tol) {
  jac_small_body(X, rows, cols, sigma, perm, res, chi_max, eps, rank_cap, tol);
}

// ------------------------------------------------------------------ QR
constexpr int QR_THREADS = 1024;
constexpr int QR_WAVES = QR_THREADS / 64;

__device__ double qr_block_sum(double v, double* sh) {
  v = wsum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = (l < QR_WAVES) ? sh[l] : 0.0;
  r = wsum(r);
  __syncthreads();
  return r;
}

// M (rows x cols, column-major) -> reflectors below the diagonal, R on and
// above it (zgeqr2), then Q's first qcols columns (zung2r) into Q.
__global__ __launch_bounds__(QR_THREADS) void householder_qr_kernel(c128* __restrict__ M, int rows,
                                                                    int cols, c128* __restrict__ tau,
                                                                    c128* __restrict__ Q, int qcols) {
  __shared__ double sh[QR_WAVES];
  __shared__ c128 s_tau;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int kmax = min(rows, cols);
  for (int j = 0; j < kmax; ++j) {
    c128* x = M + (long)j * rows;
    double part = 0;
    for (int r = j + 1 + tid; r < rows; r += QR_THREADS) part += x[r].re * x[r].re + x[r].im * x[r].im;
    const double xnorm2 = qr_block_sum(part, sh);
    const c128 alpha = x[j];
    c128 tj{0.0, 0.0};
    bool identity = (xnorm2 == 0.0 && alpha.im == 0.0);
    double beta = 0.0;
    c128 scale{0.0, 0.0};
    if (!identity) {
      const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
      beta = (alpha.re >= 0.0) ? -nrm : nrm;   // -SIGN(dlapy3, alphr)
      tj = c128{(beta - alpha.re) / beta, -alpha.im / beta};
      // scale = 1 / (alpha - beta)
      const double dr = alpha.re - beta, di = alpha.im;
      const double den = dr * dr + di * di;
      scale = c128{dr / den, -di / den};
    }
    __syncthreads();
    if (!identity) {
      for (int r = j + 1 + tid; r < rows; r += QR_THREADS) {
        const c128 v = x[r];
        x[r] = c128{v.re * scale.re - v.im * scale.im, v.re * scale.im + v.im * scale.re};
      }
    }
    __syncthreads();
    if (tid == 0) {
      if (!identity) x[j] = c128{beta, 0.0};
      tau[j] = tj;
      s_tau = tj;
    }
    __syncthreads();
    if (!identity) {
      // C <- (I - conj(tau) v v^H) C on columns j+1..cols-1, v_j = 1
      const c128 ct{s_tau.re, -s_tau.im};
      for (int c = j + 1 + wave; c < cols; c += QR_WAVES) {
        c128* y = M + (long)c * rows;
        double wr = 0, wi = 0;  // w = v^H y
        for (int r = j + lane; r < rows; r += 64) {
          const c128 v = (r == j) ? c128{1.0, 0.0} : x[r];
          wr += v.re * y[r].re + v.im * y[r].im;
          wi += v.re * y[r].im - v.im * y[r].re;
        }
        wr = wsum(wr);
        wi = wsum(wi);
        // f = conj(tau) * w
        const double fr = ct.re * wr - ct.im * wi, fi = ct.re * wi + ct.im * wr;
        for (int r = j + lane; r < rows; r += 64) {
          const c128 v = (r == j) ? c128{1.0, 0.0} : x[r];
          y[r].re -= v.re * fr - v.im * fi;
          y[r].im -= v.re * fi + v.im * fr;
        }
      }
    }
    __syncthreads();
  }
  if (Q == nullptr || qcols == 0) return;
  // Q = H_0 ... H_{kk-1} [e_0 .. e_{qcols-1}]; reflectors with index >= qcols
  // leave those columns unchanged.
  const int kk = min(kmax, qcols);
  for (long e = tid; e < (long)rows * qcols; e += QR_THREADS) {
    const int c = (int)(e / rows), r = (int)(e - (long)c * rows);
    Q[e] = c128{(r == c) ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();
  for (int i = kk - 1; i >= 0; --i) {
    const c128* v = M + (long)i * rows;
    const c128 ti = tau[i];
    for (int c = i + wave; c < qcols; c += QR_WAVES) {
      c128* y = Q + (long)c * rows;
      double wr = 0, wi = 0;
      for (int r = i + lane; r < rows; r += 64) {
        const c128 vv = (r == i) ? c128{1.0, 0.0} : v[r];
        wr += vv.re * y[r].re + vv.im * y[r].im;
        wi += vv.re * y[r].im - vv.im * y[r].re;
      }
      wr = wsum(wr);
      wi = wsum(wi);
      const double fr = ti.re * wr - ti.im * wi, fi = ti.re * wi + ti.im * wr;
      for (int r = i + lane; r < rows; r += 64) {
        const c128 vv = (r == i) ? c128{1.0, 0.0} : v[r];
        y[r].re -= vv.re * fr - vv.im * fi;
        y[r].im -= vv.re * fi + vv.im * fr;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------- blocked QR
// One panel of the blocked Householder QR (zgeqrf = zgeqr2 on panels of
// QRB_PB columns + zlarft + zlarfb): columns [j0, j0+pb) of M (rows x cols,
// column-major), rows [j0, rows), staged in LDS.  Per column: zlarfg with
// LAPACK's beta = -sign(Re alpha)*|x|, then H^H = I - conj(tau) v v^H applied to
// the panel's remaining columns.  Writes R (on/above the diagonal) and v
// (below) back to M, tau, the unit-diagonal V of the panel into V (same
// layout as M, zero above the diagonal), and T (pb x pb, column-major with
// leading dimension QRB_PB, upper triangular) with H_j0 ... H_{j0+pb-1} =
// I - V T V^H (zlarft, forward, columnwise).
constexpr int QRB_PB = 16;
constexpr int QRB_MAXROWS = 512;

__device__ __forceinline__ void qr_panel_body(c128* __restrict__ M, int rows, int j0, int pb,
                                              c128* __restrict__ tau, c128* __restrict__ V,
                                              c128* __restrict__ T, c128* __restrict__ VT) {
  __shared__ c128 P[QRB_PB][QRB_MAXROWS];   // panel, local row r = global row j0 + r
  __shared__ c128 s_tau[QRB_PB];
  __shared__ c128 s_vv[QRB_PB][QRB_PB];     // V^H V of the panel
  __shared__ c128 s_T[QRB_PB][QRB_PB];      // s_T[i][k] = T(k, i)
  __shared__ double s_xnorm2;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nr = rows - j0;
  {
    // wave w stages column w, lane l rows l, l + 64, ...: all of a thread's
    // loads are issued before its first LDS store.  Slots past the panel
    // (columns >= pb, rows >= nr) load a clamped valid element and are never
    // read, so neither loads nor stores branch.
    static_assert(QR_WAVES == QRB_PB && QRB_MAXROWS % 64 == 0, "one wave per panel column");
    constexpr int IT = QRB_MAXROWS / 64;
    const int c = min(wave, pb - 1);
    const c128* src = M + (long)(j0 + c) * rows + j0;
    c128 v[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) v[i] = src[min(lane + 64 * i, nr - 1)];
#pragma unroll
    for (int i = 0; i < IT; ++i) P[wave][lane + 64 * i] = v[i];
  }
  __syncthreads();
  for (int c = 0; c < pb; ++c) {
    c128* x = P[c];
    // |x[c+1:]|^2 by wave 0 (nr <= 512: at most 8 elements per lane)
    if (wave == 0) {
      double part = 0;
      for (int r = c + 1 + lane; r < nr; r += 64) part += x[r].re * x[r].re + x[r].im * x[r].im;
      part = wsum(part);
      if (lane == 0) s_xnorm2 = part;
    }
    __syncthreads();
    // zlarfg, evaluated redundantly by every thread from the shared norm
    const double xnorm2 = s_xnorm2;
    const c128 alpha = x[c];
    const bool identity = (xnorm2 == 0.0 && alpha.im == 0.0);
    c128 tj{0.0, 0.0};
    if (!identity) {
      const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
      const double beta = (alpha.re >= 0.0) ? -nrm : nrm;
      tj = c128{(beta - alpha.re) / beta, -alpha.im / beta};
      const double dr = alpha.re - beta, di = alpha.im;
      const double den = dr * dr + di * di;
      const c128 scale{dr / den, -di / den};
      for (int r = c + 1 + tid; r < nr; r += QR_THREADS) {
        const c128 v = x[r];
        x[r] = c128{v.re * scale.re - v.im * scale.im, v.re * scale.im + v.im * scale.re};
      }
      __syncthreads();  // uniform: identity is the same in every thread
      if (tid == 0) x[c] = c128{beta, 0.0};
    }
    if (tid == 0) {
      s_tau[c] = tj;
      tau[j0 + c] = tj;
    }
    if (!identity) {
      // C <- (I - conj(tau) v v^H) C on the panel's columns c+1.., v_c = 1
      const c128 ct{tj.re, -tj.im};
      for (int cc = c + 1 + wave; cc < pb; cc += QR_WAVES) {
        c128* y = P[cc];
        double wr = 0, wi = 0;
        for (int r = c + lane; r < nr; r += 64) {
          const c128 v = (r == c) ? c128{1.0, 0.0} : x[r];
          wr += v.re * y[r].re + v.im * y[r].im;
          wi += v.re * y[r].im - v.im * y[r].re;
        }
        wr = wsum(wr);
        wi = wsum(wi);
        const double fr = ct.re * wr - ct.im * wi, fi = ct.re * wi + ct.im * wr;
        for (int r = c + lane; r < nr; r += 64) {
          const c128 v = (r == c) ? c128{1.0, 0.0} : x[r];
          y[r].re -= v.re * fr - v.im * fi;
          y[r].im -= v.re * fi + v.im * fr;
        }
      }
    }
    __syncthreads();
  }
  // write back R / v, and the explicit unit-diagonal V of the panel
  for (int e = tid; e < pb * rows; e += QR_THREADS) {
    const int c = e / rows, g = e - c * rows;   // g: global row
    const int r = g - j0;
    const long idx = (long)(j0 + c) * rows + g;
    if (r >= 0) M[idx] = P[c][r];
    c128 v{0.0, 0.0};
    if (r == c) v = c128{1.0, 0.0};
    else if (r > c) v = P[c][r];
    V[idx] = v;
  }
  // V^H V of the panel (k < i), one wave per entry
  for (int e = wave; e < pb * pb; e += QR_WAVES) {
    const int k = e % pb, i = e / pb;
    if (k >= i) continue;
    double wr = 0, wi = 0;
    for (int r = i + lane; r < nr; r += 64) {   // V_k and V_i overlap on rows >= i
      const c128 a = (r == k) ? c128{1.0, 0.0} : P[k][r];
      const c128 b = (r == i) ? c128{1.0, 0.0} : P[i][r];
      wr += a.re * b.re + a.im * b.im;
      wi += a.re * b.im - a.im * b.re;
    }
    wr = wsum(wr);
    wi = wsum(wi);
    if (lane == 0) s_vv[k][i] = c128{wr, wi};
  }
  __syncthreads();
  // T (zlarft forward columnwise): T(i,i) = tau_i,
  // T(0:i, i) = -tau_i T(0:i, 0:i) (V(:, 0:i)^H v_i); one lane per row k of T,
  // the columns in order (T(k, l) for l < i written by the same lane earlier)
  if (tid < pb) {
    const int k = tid;
    for (int i = 0; i < pb; ++i) {
      c128 out{0.0, 0.0};
      const c128 ti = s_tau[i];
      if (k < i) {
        double ar = 0, ai = 0;
        for (int l = k; l < i; ++l) {
          const c128 t = s_T[l][k], w = s_vv[l][i];
          ar += t.re * w.re - t.im * w.im;
          ai += t.re * w.im + t.im * w.re;
        }
        out = c128{-(ti.re * ar - ti.im * ai), -(ti.re * ai + ti.im * ar)};
      } else if (k == i) {
        out = ti;
      }
      s_T[i][k] = out;
    }
  }
  __syncthreads();
  if (tid < pb * QRB_PB) {
    const int i = tid / QRB_PB, k = tid - i * QRB_PB;
    T[(long)i * QRB_PB + k] = k < pb ? s_T[i][k] : c128{0.0, 0.0};
  }
  // VT = V T on the panel's rows (both triangular: V(r, k) = 0 for r < k,
  // T(k, c) = 0 for k > c), stored like V.  The trailing update uses
  // A -= V ((V T)^H A) and the accumulation of Q uses Q -= (V T)(V^H Q), so
  // neither needs a separate product with T.
  for (int e = tid; e < nr * pb; e += QR_THREADS) {
    const int c = e / nr, r = e - c * nr;
    double ar = 0, ai = 0;
    const int kmax = min(c, r);
    for (int k = 0; k <= kmax; ++k) {
      const c128 v = (r == k) ? c128{1.0, 0.0} : P[k][r];
      const c128 t = s_T[c][k];
      ar += v.re * t.re - v.im * t.im;
      ai += v.re * t.im + v.im * t.re;
    }
    VT[(long)(j0 + c) * rows + j0 + r] = c128{ar, ai};
  }
}
// One workgroup per CU (the panel fills LDS): a register budget for 4 waves per
// SIMD lets the panel load keep all its loads in flight.
__global__ __launch_bounds__(QR_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void qr_panel_kernel(c128* __restrict__ M, int rows, int j0,
                                                              int pb, c128* __restrict__ tau,
                                                              c128* __restrict__ V,
                                                              c128* __restrict__ T,
                                                              c128* __restrict__ VT) {
  qr_panel_body(M, rows, j0, pb, tau, V, T, VT);
}

__device__ __forceinline__ void set_identity_body(c128* __restrict__ Q, int rows, int cols, unsigned blk,
                                                  unsigned nblk) {
  for (long e = blk * (long)blockDim.x + threadIdx.x; e < (long)rows * cols;
       e += (long)nblk * blockDim.x) {
    const int c = (int)(e / rows), r = (int)(e - (long)c * rows);
    Q[e] = c128{r == c ? 1.0 : 0.0, 0.0};
  }
}
__global__ void set_identity(c128* __restrict__ Q, int rows, int cols) {
  set_identity_body(Q, rows, cols, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------- QR preconditioning
// order[c]: columns of X (rows x cols) by descending norm (ties by index).
__device__ __forceinline__ void col_norm_order_body(const c128* __restrict__ X, int rows, int cols,
                                                    int* __restrict__ order) {
  __shared__ double skey[JAC_MAXCOLS];
  __shared__ int sidx[JAC_MAXCOLS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = JAC_NT / 64;
  int np2 = 1;
  while (np2 < cols) np2 <<= 1;
  for (int c = wave; c < np2; c += nw) {
    double a = (c < cols) ? lane_norm2(X + (long)c * rows, rows, lane) : 0.0;
    a = wsum(a);
    if (lane == 0) {
      skey[c] = (c < cols) ? a : -1.0;
      sidx[c] = c;
    }
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid; e < np2; e += JAC_NT) {
        const int q = e ^ stride;
        if (q > e) {
          const bool desc = ((e & size) == 0);
          const double ke = skey[e], kq = skey[q];
          const bool swap = desc ? (ke < kq || (ke == kq && sidx[e] > sidx[q]))
                                 : (ke > kq || (ke == kq && sidx[e] < sidx[q]));
          if (swap) {
            skey[e] = kq; skey[q] = ke;
            const int ti = sidx[e]; sidx[e] = sidx[q]; sidx[q] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < cols; e += JAC_NT) order[e] = sidx[e];
}
__global__ __launch_bounds__(JAC_NT) void col_norm_order(const c128* __restrict__ X, int rows, int cols,
                                                         int* __restrict__ order) {
  col_norm_order_body(X, rows, cols, order);
}

// Xp[:, c] = X[:, order[c]]
__device__ __forceinline__ void permute_cols_body(const c128* __restrict__ X, int rows, int cols,
                                                  const int* __restrict__ order, c128* __restrict__ Xp,
                                                  unsigned blk, unsigned nblk) {
  for (long e = blk * (long)blockDim.x + threadIdx.x; e < (long)rows * cols;
       e += (long)nblk * blockDim.x) {
    const int c = (int)(e / rows), r = (int)(e - (long)c * rows);
    Xp[e] = X[(long)order[c] * rows + r];
  }
}
__global__ void permute_cols(const c128* __restrict__ X, int rows, int cols,
                             const int* __restrict__ order, c128* __restrict__ Xp) {
  permute_cols_body(X, rows, cols, order, Xp, blockIdx.x, gridDim.x);
}

// Y = R^H (cols x kq, column-major) from the R on/above the diagonal of the
// QR-factored M (rows x cols): Y(c, k) = conj(R(k, c)), zero for k > c.
__device__ __forceinline__ void rh_from_qr_body(const c128* __restrict__ M, int rows, int cols, int kq,
                                                c128* __restrict__ Y, unsigned blk, unsigned nblk) {
  for (long e = blk * (long)blockDim.x + threadIdx.x; e < (long)cols * kq;
       e += (long)nblk * blockDim.x) {
    const int k = (int)(e / cols), c = (int)(e - (long)k * cols);
    const c128 v = (k <= c) ? M[(long)c * rows + k] : c128{0.0, 0.0};
    Y[e] = c128{v.re, -v.im};
  }
}
__global__ void rh_from_qr(const c128* __restrict__ M, int rows, int cols, int kq,
                           c128* __restrict__ Y) {
  rh_from_qr_body(M, rows, cols, kq, Y, blockIdx.x, gridDim.x);
}

// X[:, j] = |Y[:, j]| * Z[:, j], j < kq: the left singular vectors (columns of
// Z = Q J, orthonormal) scaled by the singular values, the form X J = U S the
// callers gather from.
__device__ __forceinline__ void scale_by_norms_body(const c128* __restrict__ Y, int yrows,
                                                    const c128* __restrict__ Z, int rows, int kq,
                                                    c128* __restrict__ X, unsigned blk, unsigned nblk) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = blk * (JAC_NT / 64) + wave; j < kq; j += nblk * (JAC_NT / 64)) {
    const double sg = sqrt(wsum(lane_norm2(Y + (long)j * yrows, yrows, lane)));
    for (int r = lane; r < rows; r += 64) {
      const c128 z = Z[(long)j * rows + r];
      X[(long)j * rows + r] = c128{sg * z.re, sg * z.im};
    }
  }
}
__global__ __launch_bounds__(JAC_NT) void scale_by_norms(const c128* __restrict__ Y, int yrows,
                                                         const c128* __restrict__ Z, int rows, int kq,
                                                         c128* __restrict__ X) {
  scale_by_norms_body(Y, yrows, Z, rows, kq, X, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------ grouped kernels
struct JacSmallArgs {
  c128* X;
  int rows, cols;
  double* sigma;
  int* perm;
  SvdResult* res;
  int chi_max;
  double eps;
  int rank_cap;
  double tol;
};
struct ColOrderArgs {
  const c128* X;
  int rows, cols;
  int* order;
  c128* Xp;          // permute_cols output
};
struct RhArgs {
  const c128* M;
  int rows, cols, kq;
  c128* Y;
};
struct ScaleArgs {
  const c128* Y;
  int yrows;
  const c128* Z;
  int rows, kq;
  c128* X;
};
// The block Jacobi of one matrix (gram_jacobi's launch arguments).
struct JacArgs {
  c128* Y;
  JacobiGeom g;
  c128* Z;
  JacobiGeom gz;
  JacobiCtrl* ctrl;
  int* rotated;
  c128* Gpart;
  c128* Jbuf;
  double tol;
};
struct FinArgs {
  const c128* X;
  int rows, cols;
  double* sigma;
  int* perm;
  SvdResult* res;
  int chi_max;
  double eps;
  int rank_cap;
  const JacobiCtrl* ctrl;
};

__global__ __launch_bounds__(JAC_NT) void jac_small_grouped(const GRec<JacSmallArgs>* __restrict__ recs, int n) {
  const JacSmallArgs a = recs[blockIdx.x].a;
  jac_small_body(gbl(a.X), a.rows, a.cols, gbl(a.sigma), gbl(a.perm), gbl(a.res), a.chi_max, a.eps, a.rank_cap,
                 a.tol);
}
__global__ __launch_bounds__(JAC_NT) void col_norm_order_grouped(const GRec<ColOrderArgs>* __restrict__ recs,
                                                                 int n) {
  const ColOrderArgs a = recs[blockIdx.x].a;
  col_norm_order_body(gbl(a.X), a.rows, a.cols, gbl(a.order));
}
__global__ void permute_cols_grouped(const GRec<ColOrderArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<ColOrderArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  permute_cols_body(gbl(r.a.X), r.a.rows, r.a.cols, gbl(r.a.order), gbl(r.a.Xp), bx, r.grid.x);
}
__global__ __launch_bounds__(QR_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4))) void qr_panel_grouped(const GRec<QrPanelArgs>* __restrict__ recs, int n) {
  const QrPanelArgs a = recs[blockIdx.x].a;
  qr_panel_body(gbl(a.M), a.rows, a.j0, a.pb, gbl(a.tau), gbl(a.V), gbl(a.T), gbl(a.VT));
}
__global__ void set_identity_grouped(const GRec<QrIdentArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<QrIdentArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  set_identity_body(gbl(r.a.Q), r.a.rows, r.a.cols, bx, r.grid.x);
}
__global__ void rh_from_qr_grouped(const GRec<RhArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<RhArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  rh_from_qr_body(gbl(r.a.M), r.a.rows, r.a.cols, r.a.kq, gbl(r.a.Y), bx, r.grid.x);
}
__global__ __launch_bounds__(JAC_NT) void scale_by_norms_grouped(const GRec<ScaleArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<ScaleArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  scale_by_norms_body(gbl(r.a.Y), r.a.yrows, gbl(r.a.Z), r.a.rows, r.a.kq, gbl(r.a.X), bx, r.grid.x);
}
__global__ __launch_bounds__(JAC_NT) void jac_init_grouped(const GRec<JacArgs>* __restrict__ recs, int n) {
  const JacArgs& a = recs[blockIdx.x].a;
  jac_init_body(gbl(a.Y), (long)a.g.rows * a.g.cols, gbl(a.ctrl));
}
// Rounds beyond a matrix's own tournament (nb - 1 rounds) are no-ops for it.
__global__ __launch_bounds__(JAC_GRAM_NT) void jac_gram_grouped(const GRec<JacArgs>* __restrict__ recs, int n,
                                                           int round) {
  unsigned bx, by, bz;
  const GRec<JacArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const JacobiGeom g = r.a.g;
  if (round >= g.nb - 1) return;
  jac_gram_body(gbl(r.a.Y), g, round, gbl(r.a.Gpart), gbl(r.a.ctrl), bx, by);
}
__global__ __launch_bounds__(JAC_NT) __attribute__((amdgpu_waves_per_eu(8))) void jac_inner_grouped(const GRec<JacArgs>* __restrict__ recs, int n,
                                                            int round, int sweep) {
  unsigned bx, by, bz;
  const GRec<JacArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const JacobiGeom g = r.a.g;
  if (round >= g.nb - 1) return;
  jac_inner_body(g, round, sweep, gbl(r.a.Gpart), gbl(r.a.Jbuf), gbl(r.a.rotated), r.a.tol, gbl(r.a.ctrl), bx);
}
__global__ __launch_bounds__(JAC_APPLY_NT) void jac_apply_grouped(const GRec<JacArgs>* __restrict__ recs, int n,
                                                            int round) {
  unsigned bx, by, bz;
  const GRec<JacArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const JacobiGeom g = r.a.g, gz = r.a.gz;
  if (round >= g.nb - 1) return;
  jac_apply_body(gbl(r.a.Y), g, gbl(r.a.Z), gz, round, gbl(r.a.Jbuf), gbl(r.a.rotated), gbl(r.a.ctrl), bx, by);
}
__global__ void jac_check_grouped(const GRec<JacArgs>* __restrict__ recs, int n, int sweep) {
  JacobiCtrl* ctrl = recs[blockIdx.x].a.ctrl;
  if (threadIdx.x != 0 || ctrl->done) return;
  ctrl->sweeps = sweep + 1;
  if (ctrl->rot[sweep] == 0) ctrl->done = 1;
}
// out[i] = done flag of matrix i (one thread per record)
__global__ void jac_done_gather(const GRec<JacArgs>* __restrict__ recs, int n, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = recs[i].a.ctrl->done;
}
__global__ __launch_bounds__(JAC_NT) void jac_finalize_grouped(const GRec<FinArgs>* __restrict__ recs, int n) {
  const FinArgs a = recs[blockIdx.x].a;
  const JacobiCtrl* ctrl = gbl(a.ctrl);
  finalize_body(gbl(a.X), a.rows, a.cols, gbl(a.sigma), gbl(a.perm), gbl(a.res), a.chi_max, a.eps, a.rank_cap,
                ctrl->sweeps, ctrl->done);
}

}  // namespace

JacobiLayout jacobi_layout(int rows, int cols) {
  JacobiLayout l;
  l.bs = cols <= JAC_LMAX ? std::max(1, (cols + 1) / 2) : JAC_BS;
  l.nb = std::max(2, ceil_div(cols, l.bs));
  if (l.nb & 1) ++l.nb;
  l.kspl = ceil_div(rows, JAC_RC);
  QCA_CHECK(l.kspl <= JAC_MAXKSPL, "Jacobi rows out of range (1..512)");
  const long L = 2L * l.bs, np = l.nb / 2;
  l.gpart_elems = np * l.kspl * L * L;
  l.j_elems = np * L * L;
  l.bytes = 256 + sizeof(int) * np + sizeof(c128) * (l.gpart_elems + l.j_elems);
  return l;
}

size_t jacobi_workspace_bytes(int max_rows, int max_cols) {
  return jacobi_layout(max_rows, max_cols).bytes + 1024;
}

namespace {

// Gram-based block Jacobi on the columns of Y (rows x cols), the same block
// rotations applied to the columns of Z (zrows x cols) when Z != null; leaves
// the control block in the workspace for the finalize.
JacobiCtrl* gram_jacobi(c128* Y, int rows, int cols, c128* Z, int zrows, hipStream_t s,
                        JacobiWork& w) {
  const JacobiLayout l = jacobi_layout(rows, cols);
  QCA_CHECK(l.bytes <= w.ws_bytes, "Jacobi workspace too small");
  static_assert(sizeof(JacobiCtrl) <= 256, "control block exceeds its slot");
  unsigned char* base = reinterpret_cast<unsigned char*>(w.ws);
  JacobiCtrl* ctrl = reinterpret_cast<JacobiCtrl*>(base);
  int* rotated = reinterpret_cast<int*>(base + 256);
  c128* Gpart = reinterpret_cast<c128*>(base + 256 + ((sizeof(int) * (l.nb / 2) + 15) / 16) * 16);
  c128* Jbuf = Gpart + l.gpart_elems;
  QCA_CHECK((reinterpret_cast<unsigned char*>(Jbuf + l.j_elems) - base) <= (long)w.ws_bytes,
            "Jacobi workspace too small");
  const JacobiGeom g{rows, cols, l.bs, l.nb, l.kspl};
  const JacobiGeom gz{zrows, cols, l.bs, l.nb, Z ? ceil_div(zrows, JAC_RC) : 0};
  const int np = l.nb / 2;
  const double tol = w.tol > 0 ? w.tol : 1.1102230246251565e-16 * sqrt((double)rows);
  hipLaunchKernelGGL(jac_init, dim3(1), dim3(JAC_NT), 0, s, Y, (long)rows * cols, ctrl);
  // launch the sweeps predicted from the previous call, then read the device
  // flag once per chunk (rounds after convergence are no-ops)
  int sweep = 0;
  int& pred = Z ? w.pred_sweeps_pre : w.pred_sweeps;
  int chunk = std::max(2, pred);
  while (sweep < JAC_MAX_SWEEPS) {
    for (int c = 0; c < chunk && sweep < JAC_MAX_SWEEPS; ++c, ++sweep) {
      for (int r = 0; r < l.nb - 1; ++r) {
        hipLaunchKernelGGL(jac_gram, dim3(np, l.kspl), dim3(JAC_GRAM_NT), 0, s, (const c128*)Y, g, r,
                           Gpart, (const JacobiCtrl*)ctrl);
        hipLaunchKernelGGL(jac_inner, dim3(np), dim3(JAC_NT), 0, s, g, r, sweep, (const c128*)Gpart,
                           Jbuf, rotated, tol, ctrl);
        hipLaunchKernelGGL(jac_apply, dim3(np, l.kspl + gz.kspl), dim3(JAC_APPLY_NT), 0, s, Y, g, Z, gz, r,
                           (const c128*)Jbuf, (const int*)rotated, (const JacobiCtrl*)ctrl);
      }
      hipLaunchKernelGGL(jac_check, dim3(1), dim3(64), 0, s, ctrl, sweep);
    }
    QCA_HIP(hipGetLastError());
    QCA_HIP(hipMemcpyAsync(w.host, &ctrl->done, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    QCA_HIP(hipStreamSynchronize(s));
    const int* hc = reinterpret_cast<const int*>(w.host);
    if (hc[0]) {
      pred = hc[1];
      break;
    }
    chunk = 2;
  }
  return ctrl;
}

}  // namespace

size_t jacobi_precond_elems(int max_rows, int max_cols) {
  const long kq = std::min(max_rows, max_cols);
  return (size_t)max_rows * max_cols       // column-permuted copy, then R / reflectors
         + (size_t)max_rows * kq           // Q, then Q J
         + (size_t)max_cols * kq           // Y = R^H
         + (size_t)kq + 64                 // tau
         + qr_workspace_elems(max_rows, max_cols, (int)kq);
}

void launch_jacobi_svd(c128* X, int rows, int cols, double* sigma, int* perm, SvdResult* res,
                       int chi_max, double eps, int rank_cap, hipStream_t s, JacobiWork& w) {
  QCA_CHECK(rows >= 1 && rows <= 512, "SVD rows out of range (1..512)");
  QCA_CHECK(cols >= 1 && cols <= JAC_MAXCOLS, "SVD cols out of range (1..512)");
  QCA_CHECK(w.ws != nullptr && w.host != nullptr, "Jacobi workspace not set");
  if (cols <= JAC_LMAX) {
    const double tol = w.tol > 0 ? w.tol : 1.1102230246251565e-16 * sqrt((double)rows);
    hipLaunchKernelGGL(jac_small, dim3(1), dim3(JAC_NT), 0, s, X, rows, cols, sigma, perm, res, chi_max,
                       eps, rank_cap, tol);
    QCA_HIP(hipGetLastError());
    return;
  }
  const int kq = std::min(rows, cols);
  if (w.pre != nullptr && kq > JAC_LMAX && w.pre_elems >= jacobi_precond_elems(rows, cols)) {
    // Preconditioned: X P = Q R with the columns by descending norm, one-sided
    // Jacobi on Y = R^H with the same rotations applied to Q.  Y J = U' S and
    // Q J = U, the left singular vectors of X, orthonormal by construction.
    c128* Xp = w.pre;
    c128* Q = Xp + (long)rows * cols;
    c128* Y = Q + (long)rows * kq;
    c128* tau = Y + (long)cols * kq;
    c128* qws = tau + kq + 64;
    const size_t qws_elems = w.pre_elems - (size_t)(qws - w.pre);
    int* order = perm;  // scratch until the finalize writes the permutation
    hipLaunchKernelGGL(col_norm_order, dim3(1), dim3(JAC_NT), 0, s, (const c128*)X, rows, cols, order);
    hipLaunchKernelGGL(permute_cols, dim3(256), dim3(256), 0, s, (const c128*)X, rows, cols,
                       (const int*)order, Xp);
    launch_qr(Xp, rows, cols, tau, Q, kq, qws, qws_elems, s);
    hipLaunchKernelGGL(rh_from_qr, dim3(256), dim3(256), 0, s, (const c128*)Xp, rows, cols, kq, Y);
    JacobiCtrl* ctrl = gram_jacobi(Y, cols, kq, Q, rows, s, w);
    hipLaunchKernelGGL(jac_finalize, dim3(1), dim3(JAC_NT), 0, s, (const c128*)Y, cols, kq, sigma,
                       perm, res, chi_max, eps, rank_cap, (const JacobiCtrl*)ctrl);
    hipLaunchKernelGGL(scale_by_norms, dim3(ceil_div(kq, JAC_NT / 64)), dim3(JAC_NT), 0, s,
                       (const c128*)Y, cols, (const c128*)Q, rows, kq, X);
    QCA_HIP(hipGetLastError());
    return;
  }
  JacobiCtrl* ctrl = gram_jacobi(X, rows, cols, nullptr, 0, s, w);
  hipLaunchKernelGGL(jac_finalize, dim3(1), dim3(JAC_NT), 0, s, (const c128*)X, rows, cols, sigma,
                     perm, res, chi_max, eps, rank_cap, (const JacobiCtrl*)ctrl);
  QCA_HIP(hipGetLastError());
}

size_t qr_workspace_elems(int max_rows, int max_cols, int max_qcols) {
  const long w = std::max(max_cols, max_qcols);
  const long panels = ceil_div(std::max(1, std::min(max_rows, max_cols)), QRB_PB);
  return 2L * max_rows * max_cols                    // V, V T
         + panels * QRB_PB * QRB_PB                  // T per panel
         + QRB_PB * w                                // W
         + 32L * QRB_PB * w + 4096;                  // GEMM split-K partials
}

// The blocked QR as a list of steps (panel factorisations, GEMMs, the
// identity start of Q), each with a slot number that orders it among the steps
// of any other matrix: the single-matrix launcher runs the steps in order, the
// grouped one runs every slot for all matrices that have it.
std::vector<QrStep> blocked_qr_steps(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols,
                                     c128* ws, size_t ws_elems) {
  QCA_CHECK(rows >= 1 && cols >= 1, "empty QR");
  QCA_CHECK(rows <= QRB_MAXROWS, "blocked QR rows above 512");
  const int kmax = std::min(rows, cols);
  const long panels = ceil_div(kmax, QRB_PB);
  const long wcols = std::max(cols, qcols);
  c128* V = ws;
  c128* VT = V + (long)rows * cols;
  c128* T = VT + (long)rows * cols;
  c128* W = T + panels * QRB_PB * QRB_PB;
  c128* gws = W + QRB_PB * wcols;
  const long used = gws - ws;
  QCA_CHECK((long)ws_elems > used + 1024, "QR workspace too small");
  const long gws_elems = (long)ws_elems - used;
  std::vector<QrStep> out;
  auto gemm_step = [&](int slot, const GemmDesc& d, const c128* A, const c128* B, c128* C) {
    QrStep st{};
    st.slot = slot;
    st.kind = 1;
    st.gemm = GemmJob<double>{d, A, B, C, gws, gws_elems};
    out.push_back(st);
  };
  // GEMM descriptors on column-major operands: X(r, c) at X + r + c * ld
  for (int p = 0; p < panels; ++p) {
    const int j0 = p * QRB_PB, pb = std::min(QRB_PB, kmax - j0);
    c128* Tp = T + (long)p * QRB_PB * QRB_PB;
    QrStep st{};
    st.slot = QR_SLOT_FACTOR + 4 * p;
    st.kind = 0;
    st.panel = QrPanelArgs{M, rows, j0, pb, tau, V, Tp, VT};
    out.push_back(st);
    const int nt = cols - (j0 + pb), nr = rows - j0;
    if (nt <= 0) continue;
    const c128* Vp = V + (long)j0 * rows + j0;
    const c128* VTp = VT + (long)j0 * rows + j0;
    c128* At = M + (long)(j0 + pb) * rows + j0;
    // W = (V T)^H A_trail = T^H V^H A_trail  (pb x nt)
    GemmDesc g{};
    g.M = pb; g.N = nt; g.K = nr; g.K2 = nr; g.NB1 = g.NB2 = 1; g.alpha_re = 1.0;
    g.a_m = rows; g.a_k2 = 1; g.conj_a = 1;
    g.b_k2 = 1; g.b_n = rows;
    g.c_m = 1; g.c_n = pb;
    gemm_step(QR_SLOT_FACTOR + 4 * p + 1, g, VTp, At, W);
    // A_trail -= V W
    GemmDesc u{};
    u.M = nr; u.N = nt; u.K = pb; u.K2 = pb; u.NB1 = u.NB2 = 1; u.alpha_re = -1.0; u.accumulate = 1;
    u.a_m = 1; u.a_k2 = rows;
    u.b_k2 = 1; u.b_n = pb;
    u.c_m = 1; u.c_n = rows;
    gemm_step(QR_SLOT_FACTOR + 4 * p + 3, u, Vp, W, At);
  }
  if (Q == nullptr || qcols == 0) return out;
  // Q = H_0 ... H_{kk-1} [e_0 .. e_{qcols-1}], block reflectors applied from
  // the last panel back: Q(j0:, j0:) <- (I - V T V^H) Q(j0:, j0:)  (zungqr)
  {
    QrStep st{};
    st.slot = QR_SLOT_IDENTITY;
    st.kind = 2;
    st.ident = QrIdentArgs{Q, rows, qcols};
    out.push_back(st);
  }
  const int kk = std::min(kmax, qcols);
  for (int p = (int)ceil_div(kk, QRB_PB) - 1; p >= 0; --p) {
    const int j0 = p * QRB_PB, pb = std::min(QRB_PB, kk - j0);
    const int nr = rows - j0, nq = qcols - j0;
    const c128* Vp = V + (long)j0 * rows + j0;
    const c128* VTp = VT + (long)j0 * rows + j0;
    c128* Qs = Q + (long)j0 * rows + j0;
    const int base = QR_SLOT_Q + 3 * (QR_MAX_PANELS - 1 - p);
    // W = V^H Q(j0:, j0:), then Q(j0:, j0:) -= (V T) W
    GemmDesc g{};
    g.M = pb; g.N = nq; g.K = nr; g.K2 = nr; g.NB1 = g.NB2 = 1; g.alpha_re = 1.0;
    g.a_m = rows; g.a_k2 = 1; g.conj_a = 1;
    g.b_k2 = 1; g.b_n = rows;
    g.c_m = 1; g.c_n = pb;
    gemm_step(base, g, Vp, Qs, W);
    GemmDesc u{};
    u.M = nr; u.N = nq; u.K = pb; u.K2 = pb; u.NB1 = u.NB2 = 1; u.alpha_re = -1.0; u.accumulate = 1;
    u.a_m = 1; u.a_k2 = rows;
    u.b_k2 = 1; u.b_n = pb;
    u.c_m = 1; u.c_n = rows;
    gemm_step(base + 2, u, VTp, W, Qs);
  }
  return out;
}

void launch_blocked_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols, c128* ws,
                       size_t ws_elems, hipStream_t s) {
  for (const QrStep& st : blocked_qr_steps(M, rows, cols, tau, Q, qcols, ws, ws_elems)) {
    if (st.kind == 0) {
      const QrPanelArgs& a = st.panel;
      hipLaunchKernelGGL(qr_panel_kernel, dim3(1), dim3(QR_THREADS), 0, s, a.M, a.rows, a.j0, a.pb, a.tau,
                         a.V, a.T, a.VT);
    } else if (st.kind == 1) {
      gemm_c128(st.gemm.d, st.gemm.A, st.gemm.B, st.gemm.C, s, st.gemm.ws, st.gemm.ws_elems);
    } else {
      hipLaunchKernelGGL(set_identity, dim3(256), dim3(256), 0, s, st.ident.Q, st.ident.rows, st.ident.cols);
    }
  }
  QCA_HIP(hipGetLastError());
}

void grouped_jacobi_svd(const std::vector<SvdJob>& jobs, RecordArena& arena, GroupFlags& flags,
                        hipStream_t s) {
  std::vector<GRec<JacSmallArgs>> small;
  std::vector<GRec<ColOrderArgs>> order1, perm;
  std::vector<GRec<RhArgs>> rh;
  std::vector<GRec<ScaleArgs>> scale;
  std::vector<GRec<JacArgs>> jac_one, jac_gram_r, jac_inner_r, jac_apply_r;
  std::vector<GRec<FinArgs>> fin;
  std::vector<QrStep> steps;
  int max_rounds = 0, pred = 2;
  for (const SvdJob& j : jobs) {
    QCA_CHECK(j.rows >= 1 && j.rows <= 512, "SVD rows out of range (1..512)");
    QCA_CHECK(j.cols >= 1 && j.cols <= JAC_MAXCOLS, "SVD cols out of range (1..512)");
    JacobiWork& w = *j.w;
    QCA_CHECK(w.ws != nullptr, "Jacobi workspace not set");
    if (j.cols <= JAC_LMAX) {
      const double tol = w.tol > 0 ? w.tol : 1.1102230246251565e-16 * sqrt((double)j.rows);
      small.push_back(grec(JacSmallArgs{j.X, j.rows, j.cols, j.sigma, j.perm, j.res, j.chi_max, j.eps,
                                        j.rank_cap, tol}, 1));
      continue;
    }
    const int kq = std::min(j.rows, j.cols);
    c128 *Y = j.X, *Z = nullptr;
    int yrows = j.rows, ycols = j.cols, zrows = 0;
    const bool pre = w.pre != nullptr && kq > JAC_LMAX && w.pre_elems >= jacobi_precond_elems(j.rows, j.cols);
    if (pre) {
      c128* Xp = w.pre;
      c128* Q = Xp + (long)j.rows * j.cols;
      c128* Yp = Q + (long)j.rows * kq;
      c128* tau = Yp + (long)j.cols * kq;
      c128* qws = tau + kq + 64;
      const size_t qws_elems = w.pre_elems - (size_t)(qws - w.pre);
      QCA_CHECK(std::min(j.rows, j.cols) > 2 * QRB_PB && qws_elems >= qr_workspace_elems(j.rows, j.cols, kq),
                "grouped SVD preconditioning needs the blocked QR");
      const ColOrderArgs co{j.X, j.rows, j.cols, j.perm, Xp};
      order1.push_back(grec(co, 1));
      perm.push_back(grec(co, 256));
      for (QrStep& st : blocked_qr_steps(Xp, j.rows, j.cols, tau, Q, kq, qws, qws_elems)) steps.push_back(st);
      rh.push_back(grec(RhArgs{Xp, j.rows, j.cols, kq, Yp}, 256));
      Y = Yp;
      Z = Q;
      yrows = j.cols;
      ycols = kq;
      zrows = j.rows;
      scale.push_back(grec(ScaleArgs{Yp, j.cols, Q, j.rows, kq, j.X}, ceil_div(kq, JAC_NT / 64)));
    }
    const JacobiLayout l = jacobi_layout(yrows, ycols);
    QCA_CHECK(l.bytes <= w.ws_bytes, "Jacobi workspace too small");
    unsigned char* base = reinterpret_cast<unsigned char*>(w.ws);
    JacArgs a{};
    a.ctrl = reinterpret_cast<JacobiCtrl*>(base);
    a.rotated = reinterpret_cast<int*>(base + 256);
    a.Gpart = reinterpret_cast<c128*>(base + 256 + ((sizeof(int) * (l.nb / 2) + 15) / 16) * 16);
    a.Jbuf = a.Gpart + l.gpart_elems;
    a.Y = Y;
    a.g = JacobiGeom{yrows, ycols, l.bs, l.nb, l.kspl};
    a.Z = Z;
    a.gz = JacobiGeom{zrows, ycols, l.bs, l.nb, Z ? ceil_div(zrows, JAC_RC) : 0};
    a.tol = w.tol > 0 ? w.tol : 1.1102230246251565e-16 * sqrt((double)yrows);
    const int np = l.nb / 2;
    jac_one.push_back(grec(a, 1));
    jac_gram_r.push_back(grec(a, np, l.kspl));
    jac_inner_r.push_back(grec(a, np));
    jac_apply_r.push_back(grec(a, np, l.kspl + a.gz.kspl));
    fin.push_back(grec(FinArgs{Y, yrows, ycols, j.sigma, j.perm, j.res, j.chi_max, j.eps, j.rank_cap, a.ctrl}, 1));
    max_rounds = std::max(max_rounds, l.nb - 1);
    pred = std::max(pred, pre ? w.pred_sweeps_pre : w.pred_sweeps);
  }
  KGroup<JacSmallArgs> g_small;
  KGroup<ColOrderArgs> g_order, g_perm;
  KGroup<RhArgs> g_rh;
  KGroup<ScaleArgs> g_scale;
  KGroup<JacArgs> g_one, g_gram, g_inner, g_apply;
  KGroup<FinArgs> g_fin;
  g_small.prepare(small, arena);
  g_order.prepare(order1, arena);
  g_perm.prepare(perm, arena);
  g_rh.prepare(rh, arena);
  g_scale.prepare(scale, arena);
  g_one.prepare(jac_one, arena);
  g_gram.prepare(jac_gram_r, arena);
  g_inner.prepare(jac_inner_r, arena);
  g_apply.prepare(jac_apply_r, arena);
  g_fin.prepare(fin, arena);
  // blocked QR steps of every preconditioned matrix, slot by slot
  std::sort(steps.begin(), steps.end(), [](const QrStep& x, const QrStep& y) { return x.slot < y.slot; });
  struct Slot {
    int kind;
    KGroup<QrPanelArgs> panel;
    KGroup<QrIdentArgs> ident;
    GemmGroup<double> gemm;
  };
  std::vector<Slot> slots;
  for (size_t i = 0; i < steps.size();) {
    size_t e = i;
    while (e < steps.size() && steps[e].slot == steps[i].slot) ++e;
    Slot sl{};
    sl.kind = steps[i].kind;
    if (sl.kind == 0) {
      std::vector<GRec<QrPanelArgs>> v;
      for (size_t k = i; k < e; ++k) v.push_back(grec(steps[k].panel, 1));
      sl.panel.prepare(v, arena);
    } else if (sl.kind == 1) {
      std::vector<GemmJob<double>> v;
      for (size_t k = i; k < e; ++k) v.push_back(steps[k].gemm);
      sl.gemm.prepare(v, arena);
    } else {
      std::vector<GRec<QrIdentArgs>> v;
      for (size_t k = i; k < e; ++k) v.push_back(grec(steps[k].ident, 256));
      sl.ident.prepare(v, arena);
    }
    slots.push_back(sl);
    i = e;
  }
  QCA_CHECK((int)jac_one.size() <= flags.cap, "convergence flag buffer too small");
  arena.upload(s);

  if (!g_small.empty())
    hipLaunchKernelGGL(jac_small_grouped, dim3(g_small.l.blocks), dim3(JAC_NT), 0, s, g_small.l.recs, g_small.l.n);
  if (!g_order.empty()) {
    hipLaunchKernelGGL(col_norm_order_grouped, dim3(g_order.l.blocks), dim3(JAC_NT), 0, s, g_order.l.recs,
                       g_order.l.n);
    hipLaunchKernelGGL(permute_cols_grouped, dim3(g_perm.l.blocks), dim3(256), 0, s, g_perm.l.recs, g_perm.l.n);
  }
  for (const Slot& sl : slots) {
    if (sl.kind == 0)
      hipLaunchKernelGGL(qr_panel_grouped, dim3(sl.panel.l.blocks), dim3(QR_THREADS), 0, s, sl.panel.l.recs,
                         sl.panel.l.n);
    else if (sl.kind == 1)
      sl.gemm.launch(s);
    else
      hipLaunchKernelGGL(set_identity_grouped, dim3(sl.ident.l.blocks), dim3(256), 0, s, sl.ident.l.recs,
                         sl.ident.l.n);
  }
  if (!g_rh.empty())
    hipLaunchKernelGGL(rh_from_qr_grouped, dim3(g_rh.l.blocks), dim3(256), 0, s, g_rh.l.recs, g_rh.l.n);
  QCA_HIP(hipGetLastError());
  if (!g_one.empty()) {
    hipLaunchKernelGGL(jac_init_grouped, dim3(g_one.l.blocks), dim3(JAC_NT), 0, s, g_one.l.recs, g_one.l.n);
    const int nj = g_one.l.n;
    int sweep = 0, chunk = std::max(2, pred);
    while (sweep < JAC_MAX_SWEEPS) {
      for (int c = 0; c < chunk && sweep < JAC_MAX_SWEEPS; ++c, ++sweep) {
        for (int r = 0; r < max_rounds; ++r) {
          hipLaunchKernelGGL(jac_gram_grouped, dim3(g_gram.l.blocks), dim3(JAC_GRAM_NT), 0, s, g_gram.l.recs,
                             g_gram.l.n, r);
          hipLaunchKernelGGL(jac_inner_grouped, dim3(g_inner.l.blocks), dim3(JAC_NT), 0, s, g_inner.l.recs,
                             g_inner.l.n, r, sweep);
          hipLaunchKernelGGL(jac_apply_grouped, dim3(g_apply.l.blocks), dim3(JAC_APPLY_NT), 0, s, g_apply.l.recs,
                             g_apply.l.n, r);
        }
        hipLaunchKernelGGL(jac_check_grouped, dim3(nj), dim3(64), 0, s, g_one.l.recs, nj, sweep);
      }
      hipLaunchKernelGGL(jac_done_gather, dim3(ceil_div(nj, 256)), dim3(256), 0, s, g_one.l.recs, nj, flags.dev);
      QCA_HIP(hipGetLastError());
      QCA_HIP(hipMemcpyAsync(flags.host, flags.dev, sizeof(int) * nj, hipMemcpyDeviceToHost, s));
      QCA_HIP(hipStreamSynchronize(s));
      bool all = true;
      for (int i = 0; i < nj; ++i) all = all && flags.host[i] != 0;
      if (all) break;
      chunk = 2;
    }
    hipLaunchKernelGGL(jac_finalize_grouped, dim3(g_fin.l.blocks), dim3(JAC_NT), 0, s, g_fin.l.recs, g_fin.l.n);
    if (!g_scale.empty())
      hipLaunchKernelGGL(scale_by_norms_grouped, dim3(g_scale.l.blocks), dim3(JAC_NT), 0, s, g_scale.l.recs,
                         g_scale.l.n);
  }
  QCA_HIP(hipGetLastError());
}

void launch_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols, c128* ws,
               size_t ws_elems, hipStream_t s) {
  // one workgroup is quicker while the panels would be few
  if (ws != nullptr && std::min(rows, cols) > 2 * QRB_PB &&
      ws_elems >= qr_workspace_elems(rows, cols, qcols))
    launch_blocked_qr(M, rows, cols, tau, Q, qcols, ws, ws_elems, s);
  else
    launch_householder_qr(M, rows, cols, tau, Q, qcols, s);
}

void launch_householder_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols,
                           hipStream_t s) {
  QCA_CHECK(rows >= 1 && cols >= 1, "empty QR");
  hipLaunchKernelGGL(householder_qr_kernel, dim3(1), dim3(QR_THREADS), 0, s, M, rows, cols, tau, Q,
                     qcols);
  QCA_HIP(hipGetLastError());
}

}  // namespace qca
