// Non-GEMM device kernels of the TDVP path (gfx950).
//
//  * wmix      -- sparse MPO mixing between the two GEMMs of every contraction
//  * permute   -- Theta <-> SVD matrix layouts, gathers of singular vectors
//  * lanczos   -- FP64 Krylov exp(-i tau H) v with CGS2 reorthogonalisation;
//                 the tridiagonal exponential runs on one wave (Taylor with
//                 scaling) so no host round trip is needed per iteration
//  * measure   -- <n_i> and the single-site density matrix from the transfer
//  * exact     -- closed-form 2^N rule-Hamiltonian application
#include "grouped.hpp"
#include "wave.hpp"

#include <cstring>

namespace qca {
namespace {

// ------------------------------------------------------------------ helpers
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // conj(a) * b
  return cplx<T>{a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}

__device__ __forceinline__ double wave_sum(double v) { return wave_sum64(v); }

// Block-wide sum of a double; blockDim.x must be a multiple of 64, <= 1024.
__device__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double r = 0;
  if (threadIdx.x < 64) {
    r = (l < nw) ? sh[l] : 0.0;
    r = wave_sum(r);
    if (l == 0) sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ wmix
// U[q][xc][v] = sum_{i} Wsp[(q,v), i] * Tin[i][xc]
// A block owns WMIX_XC consecutive xc and runs WMIX_WAVES groups of WMIX_XC
// threads: the input channels of those xc are staged through LDS once (each
// element read from memory once), the groups share the nq*Dout output
// channels, and the output rows [q][xc0..xc0+WMIX_XC)[0..Dout) are written as
// contiguous runs.  32 xc per block (two groups per wave, two blocks' worth of
// LDS per CU at chi = 128) beat 64: 51.5 vs 52.5 us per 2-site H_eff.
constexpr int WMIX_XC = 32;
constexpr int WMIX_WAVES = 16;
constexpr int WMIX_DMAX = 16;

template <typename T>
__device__ __forceinline__ void wmix_body(const cplx<T>* __restrict__ Tin, long in_stride, long nxc,
                                          int nin, cplx<T>* __restrict__ U, int nq, int Dout,
                                          const int* __restrict__ rowptr, const int* __restrict__ col,
                                          const cplx<T>* __restrict__ val, const int* __restrict__ skip,
                                          unsigned blk, unsigned char* smem_raw) {
  if (skip && *skip) return;
  cplx<T>* sin = reinterpret_cast<cplx<T>*>(smem_raw);        // [nin][64]
  cplx<T>* sout = sin + (long)nin * WMIX_XC;                   // [nq][WMIX_XC][Dout]
  const int t = threadIdx.x & (WMIX_XC - 1), wv = threadIdx.x / WMIX_XC;
  const long xc0 = (long)blk * WMIX_XC;
  const long xc = xc0 + t;
  const bool live = xc < nxc;
  for (int i = wv; i < nin; i += WMIX_WAVES)
    sin[i * WMIX_XC + t] = live ? Tin[(long)i * in_stride + xc] : cplx<T>{T(0), T(0)};
  __syncthreads();
  for (int o = wv; o < nq * Dout; o += WMIX_WAVES) {
    const int q = o / Dout, v = o - q * Dout;
    T ar = 0, ai = 0;
    for (int z = rowptr[o]; z < rowptr[o + 1]; ++z) {
      const cplx<T> x = sin[col[z] * WMIX_XC + t];
      const cplx<T> w = val[z];
      ar += w.re * x.re - w.im * x.im;
      ai += w.re * x.im + w.im * x.re;
    }
    sout[((long)q * WMIX_XC + t) * Dout + v] = cplx<T>{ar, ai};
  }
  __syncthreads();
  const int nrun = (int)min<long>(WMIX_XC, nxc - xc0);
  for (int q = 0; q < nq; ++q) {
    cplx<T>* dst = U + ((long)q * nxc + xc0) * Dout;
    const cplx<T>* src = sout + (long)q * WMIX_XC * Dout;
    for (int e = threadIdx.x; e < nrun * Dout; e += WMIX_XC * WMIX_WAVES) dst[e] = src[e];
  }
}

template <typename T>
__global__ __launch_bounds__(WMIX_XC * WMIX_WAVES) void wmix_kernel(
    const cplx<T>* __restrict__ Tin, long in_stride, long nxc, int nin, cplx<T>* __restrict__ U,
    int nq, int Dout, const int* __restrict__ rowptr, const int* __restrict__ col,
    const cplx<T>* __restrict__ val, const int* __restrict__ skip) {
  extern __shared__ unsigned char smem_raw[];
  wmix_body<T>(Tin, in_stride, nxc, nin, U, nq, Dout, rowptr, col, val, skip, blockIdx.x, smem_raw);
}

template <typename T>
__global__ __launch_bounds__(WMIX_XC * WMIX_WAVES) void wmix_grouped(const GRec<WmixArgs<T>>* __restrict__ recs,
                                                                     int n) {
  extern __shared__ unsigned char smem_raw[];
  unsigned bx, by, bz;
  const GRec<WmixArgs<T>>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const WmixArgs<T> a = r.a;
  wmix_body<T>(gbl(a.Tin), a.in_stride, a.nxc, a.nin, gbl(a.U), a.nq, a.Dout, gbl(a.rowptr), gbl(a.col),
               gbl(a.val), gbl(a.skip), bx, smem_raw);
}

// ------------------------------------------------------------------ conversions
template <typename S, typename D>
__device__ __forceinline__ void convert_body(const cplx<S>* __restrict__ in, cplx<D>* __restrict__ out,
                                             long n, const int* __restrict__ skip, unsigned blk,
                                             unsigned nblk) {
  if (skip && *skip) return;
  for (long i = blk * (long)blockDim.x + threadIdx.x; i < n; i += (long)nblk * blockDim.x)
    out[i] = cplx<D>{D(in[i].re), D(in[i].im)};
}
template <typename S, typename D>
__global__ void convert_kernel(const cplx<S>* __restrict__ in, cplx<D>* __restrict__ out, long n,
                               const int* __restrict__ skip) {
  convert_body<S, D>(in, out, n, skip, blockIdx.x, gridDim.x);
}
__global__ void convert_grouped(const GRec<ConvArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<ConvArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  convert_body<double, float>(gbl(r.a.in), gbl(r.a.out), r.a.n, gbl(r.a.skip), bx, r.grid.x);
}

// ------------------------------------------------------------------ permutes
// Theta[p1][p2][a][c] -> column-major X of the SVD matrix M[(p1,a),(p2,c)].
// mode 0: X[j][r] = M[r][j] (columns of M); mode 1: X[r][j] = conj(M[r][j])
// (columns of M^dagger).
__device__ __forceinline__ void theta_to_svd_body(const c128* __restrict__ th, c128* __restrict__ X,
                                                  int chil, int chir, int mode, unsigned blk,
                                                  unsigned nblk) {
  const long m = 2L * chil, n = 2L * chir;
  for (long idx = blk * (long)blockDim.x + threadIdx.x; idx < m * n;
       idx += (long)nblk * blockDim.x) {
    // idx enumerates Theta in memory order: p1, p2, a, c
    const long c = idx % chir;
    long t = idx / chir;
    const long a = t % chil;
    t /= chil;
    const long p2 = t % 2, p1 = t / 2;
    const long r = p1 * chil + a, j = p2 * chir + c;
    c128 v = th[idx];
    if (mode == 0) {
      X[j * m + r] = v;
    } else {
      v.im = -v.im;
      X[r * n + j] = v;
    }
  }
}
__global__ void theta_to_svd(const c128* __restrict__ th, c128* __restrict__ X, int chil, int chir,
                             int mode) {
  theta_to_svd_body(th, X, chil, chir, mode, blockIdx.x, gridDim.x);
}
__global__ void theta_to_svd_grouped(const GRec<ThetaSvdArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<ThetaSvdArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  theta_to_svd_body(gbl(r.a.th), gbl(r.a.X), r.a.chil, r.a.chir, r.a.mode, bx, r.grid.x);
}

// A[p][a][k] = X[perm[k]][(p,a)] / sigma[k]   (mode 0, left isometry from U)
// A[p][k][c] = conj(X[perm[k]][(p,c)]) / sigma[k]  (mode 1, right isometry from V)
__device__ __forceinline__ void gather_isometry_body(const c128* __restrict__ X, const int* __restrict__ perm,
                                                     const double* __restrict__ sigma, c128* __restrict__ A,
                                                     int rows, int chi_bond, int k, int mode, unsigned blk,
                                                     unsigned nblk) {
  // rows = 2*chi_bond (length of each singular vector)
  const long total = (long)rows * k;
  for (long idx = blk * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)nblk * blockDim.x) {
    const int kk = (int)(idx % k);
    const int r = (int)(idx / k);
    const double s = sigma[kk];
    const double inv = s > 0 ? 1.0 / s : 0.0;
    c128 v = X[(long)perm[kk] * rows + r];
    const int p = r / chi_bond, a = r - p * chi_bond;
    if (mode == 0) {
      A[((long)p * chi_bond + a) * k + kk] = c128{v.re * inv, v.im * inv};
    } else {
      A[((long)p * k + kk) * chi_bond + a] = c128{v.re * inv, -v.im * inv};
    }
  }
}
__global__ void gather_isometry(const c128* __restrict__ X, const int* __restrict__ perm,
                                const double* __restrict__ sigma, c128* __restrict__ A, int rows,
                                int chi_bond, int k, int mode) {
  gather_isometry_body(X, perm, sigma, A, rows, chi_bond, k, mode, blockIdx.x, gridDim.x);
}
__global__ void gather_isometry_grouped(const GRec<GatherIsoArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<GatherIsoArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const GatherIsoArgs a = r.a;
  gather_isometry_body(gbl(a.X), gbl(a.perm), gbl(a.sigma), gbl(a.A), a.rows, a.chi_bond, a.k, a.mode, bx,
                       r.grid.x);
}

// Row-major (rows x cols) -> column-major copy: out[c*rows + r] = in[r*cols + c]
__global__ void to_colmajor(const c128* __restrict__ in, c128* __restrict__ out, int rows, int cols) {
  const long n = (long)rows * cols;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long r = e / cols, c = e - r * cols;
    out[c * rows + r] = in[e];
  }
}

// A[p][a][b] -> column-major X of A^T reshaped: X[a][(p,b)], rows = 2*chir
__global__ void site_to_rightqr(const c128* __restrict__ A, c128* __restrict__ X, int chil, int chir) {
  const long n = 2L * chil * chir;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long b = e % chir, t = e / chir;
    const long a = t % chil, p = t / chil;
    X[a * (2L * chir) + p * chir + b] = A[e];
  }
}

// Left QR results: Anew[(p,a)][k] = Q[k][(p,a)] (k < qcols, else 0) for k < target;
// C[k][b] = R[k][b] (upper triangle of X, zero below / beyond min(rows, cols)).
__global__ void qr_gather_left(const c128* __restrict__ Q, const c128* __restrict__ X, c128* __restrict__ Anew,
                               c128* __restrict__ C, int rows, int cols, int qcols, int target) {
  const long na = (long)rows * target, nc = (long)target * cols;
  const int kmin = min(rows, cols);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < na + nc; e += (long)gridDim.x * blockDim.x) {
    if (e < na) {
      const long r = e / target, k = e - r * target;
      Anew[e] = (k < qcols) ? Q[k * rows + r] : c128{0.0, 0.0};
    } else {
      const long f = e - na;
      const long k = f / cols, b = f - k * cols;
      C[f] = (k <= b && k < kmin) ? X[b * rows + k] : c128{0.0, 0.0};
    }
  }
}

// Right QR results: Anew[p][k][b] = Q[k][(p,b)] (k < qcols, else 0), rows = 2*chir;
// C[a][k] = R[k][a] for k < target.
__global__ void qr_gather_right(const c128* __restrict__ Q, const c128* __restrict__ X, c128* __restrict__ Anew,
                                c128* __restrict__ C, int chir, int cols, int qcols, int target) {
  const int rows = 2 * chir;
  const long na = 2L * target * chir, nc = (long)cols * target;
  const int kmin = min(rows, cols);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < na + nc; e += (long)gridDim.x * blockDim.x) {
    if (e < na) {
      const long b = e % chir, t = e / chir;
      const long k = t % target, p = t / target;
      Anew[e] = (k < qcols) ? Q[k * rows + p * chir + b] : c128{0.0, 0.0};
    } else {
      const long f = e - na;
      const long a = f / target, k = f - a * target;
      C[f] = (k <= a && k < kmin) ? X[a * rows + k] : c128{0.0, 0.0};
    }
  }
}

// ------------------------------------------------------------------ lanczos
constexpr int LZ_BLOCK = 256;
constexpr int LZ_DIRECT_BLOCKS = 4096;   // in-kernel partial reductions up to this many blocks
constexpr int LZ_EPT = 1;                     // vector elements per thread
constexpr int LZ_ELEMS = LZ_BLOCK * LZ_EPT;   // elements per block in the partial reductions

__device__ __forceinline__ void lz_norm_partials_body(const c128* __restrict__ x, long n,
                                                      double* __restrict__ part, unsigned blk) {
  __shared__ double sh[16];
  double acc = 0;
  const long base = (long)blk * LZ_ELEMS;
  for (long i = base + threadIdx.x; i < min(n, base + LZ_ELEMS); i += blockDim.x)
    acc += x[i].re * x[i].re + x[i].im * x[i].im;
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blk] = acc;
}
__global__ void lz_norm_partials(const c128* __restrict__ x, long n, double* __restrict__ part) {
  lz_norm_partials_body(x, n, part, blockIdx.x);
}

__global__ void lz_init_finalize(LanczosCtrl* ctrl, const double* __restrict__ part, int nblk) {
  if (threadIdx.x != 0) return;
  double s = 0;
  for (int b = 0; b < nblk; ++b) s += part[b];
  ctrl->beta0 = sqrt(s);
  ctrl->done = (s == 0.0) ? 1 : 0;
  ctrl->m = 0;
  ctrl->err = 0;
  ctrl->zero = (s == 0.0) ? 1 : 0;
  ctrl->hit_limit = 0;
}

__device__ __forceinline__ void lz_scale_first_body(const c128* __restrict__ x, c128* __restrict__ v0,
                                                    long n, const LanczosCtrl* ctrl,
                                                    c64* __restrict__ v32, unsigned blk, unsigned nblk) {
  const double b = ctrl->beta0;
  const double inv = b > 0 ? 1.0 / b : 0.0;
  for (long i = blk * (long)blockDim.x + threadIdx.x; i < n; i += (long)nblk * blockDim.x) {
    const c128 v{x[i].re * inv, x[i].im * inv};
    v0[i] = v;
    if (v32) v32[i] = c64{float(v.re), float(v.im)};
  }
}
__global__ void lz_scale_first(const c128* __restrict__ x, c128* __restrict__ v0, long n,
                               const LanczosCtrl* ctrl, c64* __restrict__ v32) {
  lz_scale_first_body(x, v0, n, ctrl, v32, blockIdx.x, gridDim.x);
}

// This block's partial projections part[blk][i] = sum_e conj(V_i[e]) x[e],
// i = 0..j, with x held in registers (LZ_EPT elements per thread).  Each
// projection is reduced inside its wave; one barrier then combines the waves.
__device__ void block_projections(const c128* __restrict__ V, long n, int j, long base,
                                  const double* xr, const double* xi,
                                  c128* __restrict__ part_blk) {
  constexpr int NW = LZ_BLOCK / 64;
  __shared__ double sh[NW][LZ_MAXK][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = 0; i <= j; ++i) {
    const c128* vi = V + (long)i * n;
    double ar = 0, ai = 0;
#pragma unroll
    for (int r = 0; r < LZ_EPT; ++r) {
      const long e = base + (long)r * LZ_BLOCK;
      if (e < n) {
        const c128 a = vi[e];
        ar += a.re * xr[r] + a.im * xi[r];
        ai += a.re * xi[r] - a.im * xr[r];
      }
    }
    ar = wave_sum(ar);
    ai = wave_sum(ai);
    if (lane == 0) {
      sh[wv][i][0] = ar;
      sh[wv][i][1] = ai;
    }
  }
  __syncthreads();
  if (threadIdx.x <= j) {
    double ar = 0, ai = 0;
    for (int w = 0; w < NW; ++w) {
      ar += sh[w][threadIdx.x][0];
      ai += sh[w][threadIdx.x][1];
    }
    part_blk[threadIdx.x] = c128{ar, ai};
  }
}

// Partial projections c_i = <V_i, w> over this block's chunk.  Each thread
// keeps its LZ_EPT elements of w in registers across the i loop.
template <typename TW>
__device__ __forceinline__ void lz_project_body(const LanczosCtrl* ctrl, const cplx<TW>* __restrict__ w,
                                                const c128* __restrict__ V, long n, int j,
                                                c128* __restrict__ part, unsigned blk) {
  const int done = ctrl->done;  // tested once w's loads are in flight
  const long base = (long)blk * LZ_ELEMS + threadIdx.x;
  double wr[LZ_EPT], wi[LZ_EPT];
#pragma unroll
  for (int r = 0; r < LZ_EPT; ++r) {
    const long e = base + (long)r * LZ_BLOCK;
    wr[r] = e < n ? double(w[e].re) : 0.0;
    wi[r] = e < n ? double(w[e].im) : 0.0;
  }
  if (done) return;
  block_projections(V, n, j, base, wr, wi, part + (long)blk * LZ_MAXK);
}
template <typename TW>
__global__ __launch_bounds__(LZ_BLOCK) void lz_project(const LanczosCtrl* ctrl,
                                                       const cplx<TW>* __restrict__ w,
                                                       const c128* __restrict__ V, long n, int j,
                                                       c128* __restrict__ part) {
  lz_project_body<TW>(ctrl, w, V, n, j, part, blockIdx.x);
}

// coef_i = sum over blocks of part_in[b][i] (i = 0..j), reduced by every block
// in one fixed order (one wave per coefficient, lanes striding over the
// blocks, then a butterfly: deterministic), so no separate reduction launch is
// needed; block 0 also stores them in coef_out.  Then wout = w - sum_i coef_i V_i (promoting w to
// FP64), and the partial projections of wout (pass 1) or its partial squared
// norm (pass 2) go to part_out, a different buffer from part_in.
template <typename TW>
__device__ __forceinline__ void lz_subtract_body(const LanczosCtrl* ctrl, const cplx<TW>* __restrict__ w,
                                                 c128* __restrict__ wout, const c128* __restrict__ V,
                                                 long n, int j, const c128* __restrict__ part_in,
                                                 int nblk, c128* __restrict__ coef_out,
                                                 c128* __restrict__ part_out, int pass, unsigned blk,
                                                 const c128* __restrict__ coef_in = nullptr) {
  const int done = ctrl->done;  // tested after the coefficient loads
  __shared__ double sh[16];
  __shared__ c128 cs[LZ_MAXK];
  if (coef_in) {
    // coefficients reduced beforehand by lz_reduce (long vectors)
    if (threadIdx.x <= j) cs[threadIdx.x] = coef_in[threadIdx.x];
  } else {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = wv; i <= j; i += LZ_BLOCK / 64) {
      double ar = 0, ai = 0;
      for (int b = lane; b < nblk; b += 64) {
        ar += part_in[(long)b * LZ_MAXK + i].re;
        ai += part_in[(long)b * LZ_MAXK + i].im;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) {
        cs[i] = c128{ar, ai};
        if (blk == 0 && !done) coef_out[i] = c128{ar, ai};
      }
    }
  }
  __syncthreads();
  if (done) return;
  const long base = (long)blk * LZ_ELEMS + threadIdx.x;
  double xr[LZ_EPT], xi[LZ_EPT];
#pragma unroll
  for (int r = 0; r < LZ_EPT; ++r) {
    const long e = base + (long)r * LZ_BLOCK;
    xr[r] = e < n ? double(w[e].re) : 0.0;
    xi[r] = e < n ? double(w[e].im) : 0.0;
  }
  for (int i = 0; i <= j; ++i) {
    const c128 c = cs[i];
#pragma unroll
    for (int r = 0; r < LZ_EPT; ++r) {
      const long e = base + (long)r * LZ_BLOCK;
      if (e < n) {
        const c128 a = V[(long)i * n + e];
        xr[r] -= c.re * a.re - c.im * a.im;
        xi[r] -= c.re * a.im + c.im * a.re;
      }
    }
  }
  double nrm = 0;
#pragma unroll
  for (int r = 0; r < LZ_EPT; ++r) {
    const long e = base + (long)r * LZ_BLOCK;
    if (e < n) {
      wout[e] = c128{xr[r], xi[r]};
      nrm += xr[r] * xr[r] + xi[r] * xi[r];
    }
  }
  if (pass == 2) {
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) part_out[(long)blk * LZ_MAXK] = c128{nrm, 0};
    return;
  }
  block_projections(V, n, j, base, xr, xi, part_out + (long)blk * LZ_MAXK);
}
template <typename TW>
__global__ __launch_bounds__(LZ_BLOCK) void lz_subtract(const LanczosCtrl* ctrl,
                                                        const cplx<TW>* __restrict__ w,
                                                        c128* __restrict__ wout,
                                                        const c128* __restrict__ V, long n, int j,
                                                        const c128* __restrict__ part_in, int nblk,
                                                        c128* __restrict__ coef_out,
                                                        c128* __restrict__ part_out, int pass,
                                                        const c128* __restrict__ coef_in) {
  lz_subtract_body<TW>(ctrl, w, wout, V, n, j, part_in, nblk, coef_out, part_out, pass, blockIdx.x,
                       coef_in);
}

// Two-level reduction for long vectors: out[i] = sum_b part[b][i] for the
// coefficients i = blockIdx.x (a fixed order: lanes stride over the blocks,
// then the waves are combined in order).  With norm != 0, only the squared
// norm partials part[b][0].re are summed, into ctrl->wnorm2.
__global__ __launch_bounds__(256) void lz_reduce(LanczosCtrl* ctrl, const c128* __restrict__ part, int nblk,
                                                 c128* __restrict__ out, int norm) {
  if (ctrl->done) return;
  __shared__ double sh[16];
  const int i = blockIdx.x;
  double ar = 0, ai = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    const c128 v = part[(long)b * LZ_MAXK + i];
    ar += v.re;
    ai += v.im;
  }
  ar = block_sum(ar, sh);
  if (norm) {
    if (threadIdx.x == 0) ctrl->wnorm2 = ar;
    return;
  }
  ai = block_sum(ai, sh);
  if (threadIdx.x == 0) out[i] = c128{ar, ai};
}

// beta_j = |w| from the per-block partial norms (one wave, fixed order).
__device__ __forceinline__ double lz_beta(const c128* __restrict__ part_norm, int nblk) {
  double s = 0;
  for (int b = threadIdx.x & 63; b < nblk; b += 64) s += part_norm[(long)b * LZ_MAXK].re;
  return sqrt(wave_sum(s));
}

// One wave: close iteration j.  alpha_j = Re(c_j + c'_j), beta_j = |w|;
// z = exp(-i tau T_{j+1}) e_0 by scaled Taylor steps on the tridiagonal
// (lane = row); converged when beta0*beta_j*|z_j| < tol*beta0.
__device__ void lz_close_iteration(LanczosCtrl* ctrl, double beta, const c128* __restrict__ coef1,
                                   const c128* __restrict__ coef2, int j, double tau, double tol,
                                   int max_m) {
  const int l = threadIdx.x & 63;
  const double alpha_j = coef1[j].re + coef2[j].re;
  if (l == 0) {
    ctrl->alpha[j] = alpha_j;
    ctrl->beta[j] = beta;
  }
  const int m = j + 1;
  // alpha[0..j-1] and beta[0..j-1] were written by earlier launches
  const double al = (l < j) ? ctrl->alpha[l] : ((l == j) ? alpha_j : 0.0);
  const double bup = (l + 1 < m) ? ctrl->beta[l] : 0.0;     // T[l][l+1]
  const double bdn = (l >= 1 && l < m) ? ctrl->beta[l - 1] : 0.0;  // T[l][l-1]
  double rown = fabs(al) + fabs(bup) + fabs(bdn);
  for (int o = 32; o > 0; o >>= 1) rown = fmax(rown, __shfl_xor(rown, o, 64));
  const double h = fabs(tau) * rown;
  // bounded even for non-finite input (h = NaN/inf gives one substep and a
  // non-finite error estimate, which the host reports as non-convergence)
  const int nsub = (h < 2048.0) ? max(1, (int)ceil(h / 0.5)) : 1;
  const double dt = tau / nsub;
  // z <- exp(-i dt T) z, nsub times;  multiplying by -i dt: (x+iy)*(-i dt) = dt*y - i dt*x
  double zr = (l == 0) ? 1.0 : 0.0, zi = 0.0;
  for (int sstep = 0; sstep < nsub; ++sstep) {
    double ar = zr, ai = zi, tr = zr, ti = zi;
    for (int k = 1; k < 60; ++k) {
      const double upr = __shfl_down(tr, 1, 64), upi = __shfl_down(ti, 1, 64);
      const double dnr = __shfl_up(tr, 1, 64), dni = __shfl_up(ti, 1, 64);
      double yr = al * tr + bup * upr + bdn * dnr;
      double yi = al * ti + bup * upi + bdn * dni;
      if (l >= m) { yr = 0; yi = 0; }
      const double nr = dt * yi / k, ni = -dt * yr / k;
      tr = nr;
      ti = ni;
      ar += tr;
      ai += ti;
      // stop once every row's term is below 1e-18 (a NaN keeps going, bounded)
      if (__ballot(!(fabs(tr) + fabs(ti) < 1e-18)) == 0) break;
    }
    zr = ar;
    zi = ai;
  }
  const double zlr = __shfl(zr, j, 64), zli = __shfl(zi, j, 64);
  const double err = beta * sqrt(zlr * zlr + zli * zli);
  const bool breakdown = beta <= 1e-13 * fmax(1.0, fabs(alpha_j));
  const bool conv = breakdown || err < tol || m >= max_m - 1;
  if (l < m) {
    ctrl->z[l] = c128{zr, zi};
  }
  if (l == 0) {
    ctrl->err = err;
    if (conv) {
      ctrl->m = m;
      ctrl->done = 1;
      ctrl->hit_limit = (!breakdown && err >= tol) ? 1 : 0;
    }
  }
}

// Closes iteration j and forms the next basis vector in one launch: every
// block reduces beta_j itself (same order everywhere), the first wave of block
// 0 closes the iteration, and all blocks write V_{j+1} = w / beta_j, plus its
// complex64 copy for the mixed path's next matvec when v32 is given.  Blocks
// that start after convergence skip the write; V_{j+1} is not read then.
__device__ __forceinline__ void lz_step_scale_body(LanczosCtrl* ctrl, const c128* __restrict__ part_norm,
                                                   int nblk, const c128* __restrict__ coef1,
                                                   const c128* __restrict__ coef2, int j, double tau,
                                                   double tol, int max_m, const c128* __restrict__ w,
                                                   c128* __restrict__ vnext, c64* __restrict__ v32,
                                                   long n, unsigned blk, unsigned ngrid,
                                                   int beta_reduced = 0) {
  const int done = ctrl->done;  // tested after beta's loads
  __shared__ double sbeta;
  if (threadIdx.x < 64) {
    const double b = beta_reduced ? sqrt(ctrl->wnorm2) : lz_beta(part_norm, nblk);
    if (threadIdx.x == 0) sbeta = b;
  }
  __syncthreads();
  if (done) return;
  const double beta = sbeta;
  if (blk == 0 && threadIdx.x < 64) lz_close_iteration(ctrl, beta, coef1, coef2, j, tau, tol, max_m);
  const double inv = 1.0 / beta;
  for (long i = blk * (long)blockDim.x + threadIdx.x; i < n; i += (long)ngrid * blockDim.x) {
    const c128 v{w[i].re * inv, w[i].im * inv};
    vnext[i] = v;
    if (v32) v32[i] = c64{float(v.re), float(v.im)};
  }
}
__global__ __launch_bounds__(256) void lz_step_scale(LanczosCtrl* ctrl,
                                                     const c128* __restrict__ part_norm, int nblk,
                                                     const c128* __restrict__ coef1,
                                                     const c128* __restrict__ coef2, int j,
                                                     double tau, double tol, int max_m,
                                                     const c128* __restrict__ w,
                                                     c128* __restrict__ vnext,
                                                     c64* __restrict__ v32, long n, int beta_reduced) {
  lz_step_scale_body(ctrl, part_norm, nblk, coef1, coef2, j, tau, tol, max_m, w, vnext, v32, n,
                     blockIdx.x, gridDim.x, beta_reduced);
}

// y = beta0 * sum_{i<m} z_i V_i   (a zero input stays zero)
__device__ __forceinline__ void lz_combine_body(const LanczosCtrl* ctrl, const c128* __restrict__ V,
                                                long n, c128* __restrict__ y, unsigned blk,
                                                unsigned ngrid) {
  __shared__ c128 zs[LZ_MAXK];
  const int m = ctrl->m;
  const double b0 = ctrl->beta0;
  if (threadIdx.x < m) zs[threadIdx.x] = ctrl->z[threadIdx.x];
  __syncthreads();
  for (long e = blk * (long)blockDim.x + threadIdx.x; e < n; e += (long)ngrid * blockDim.x) {
    double r = 0, im = 0;
    for (int i = 0; i < m; ++i) {
      const c128 a = V[(long)i * n + e];
      r += zs[i].re * a.re - zs[i].im * a.im;
      im += zs[i].re * a.im + zs[i].im * a.re;
    }
    y[e] = c128{b0 * r, b0 * im};
  }
}
__global__ void lz_combine(const LanczosCtrl* ctrl, const c128* __restrict__ V, long n,
                           c128* __restrict__ y) {
  lz_combine_body(ctrl, V, n, y, blockIdx.x, gridDim.x);
}

// ---- grouped Lanczos: one record per instance (LzArgs), the iteration index
// j shared by all instances of the launch (they iterate in lockstep; an
// instance that has converged skips every kernel through its done flag).
__global__ void lz_norm_partials_grouped(const GRec<LzArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<LzArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  lz_norm_partials_body(gbl(r.a.x), r.a.n, gbl(r.a.part_norm), bx);
}
__global__ void lz_init_finalize_grouped(const GRec<LzArgs>* __restrict__ recs, int n) {
  if (threadIdx.x != 0) return;
  const LzArgs& a = recs[blockIdx.x].a;
  double sum = 0;
  for (int b = 0; b < a.nblk; ++b) sum += a.part_norm[b];
  LanczosCtrl* ctrl = a.ctrl;
  ctrl->beta0 = sqrt(sum);
  ctrl->done = (sum == 0.0) ? 1 : 0;
  ctrl->m = 0;
  ctrl->err = 0;
  ctrl->zero = (sum == 0.0) ? 1 : 0;
  ctrl->hit_limit = 0;
}
__global__ void lz_scale_first_grouped(const GRec<LzArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<LzArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  lz_scale_first_body(gbl(r.a.x), gbl(r.a.V), r.a.n, gbl(r.a.ctrl), gbl(r.a.v32), bx, r.grid.x);
}
template <typename TW>
__global__ __launch_bounds__(LZ_BLOCK) void lz_project_grouped(const GRec<LzArgs>* __restrict__ recs, int n,
                                                               int j) {
  unsigned bx, by, bz;
  const GRec<LzArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  lz_project_body<TW>(gbl(r.a.ctrl), gbl(reinterpret_cast<const cplx<TW>*>(r.a.w)), gbl(r.a.V), r.a.n, j,
                      gbl(r.a.part), bx);
}
template <typename TW>
__global__ __launch_bounds__(LZ_BLOCK) void lz_subtract_grouped(const GRec<LzArgs>* __restrict__ recs, int n,
                                                                int j, int pass) {
  unsigned bx, by, bz;
  const GRec<LzArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const LzArgs a = r.a;
  if (pass == 1)
    lz_subtract_body<TW>(gbl(a.ctrl), gbl(reinterpret_cast<const cplx<TW>*>(a.w)), gbl(a.w64), gbl(a.V), a.n,
                         j, gbl(a.part), a.nblk, gbl(a.coef1), gbl(a.part2), 1, bx);
  else
    lz_subtract_body<double>(gbl(a.ctrl), gbl(a.w64), gbl(a.w64), gbl(a.V), a.n, j, gbl(a.part2), a.nblk,
                             gbl(a.coef2), gbl(a.part), 2, bx);
}
__global__ __launch_bounds__(256) void lz_step_scale_grouped(const GRec<LzArgs>* __restrict__ recs, int n,
                                                             int j) {
  unsigned bx, by, bz;
  const GRec<LzArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const LzArgs a = r.a;
  lz_step_scale_body(gbl(a.ctrl), gbl(a.part), a.nblk, gbl(a.coef1), gbl(a.coef2), j, a.tau, a.tol, a.max_m,
                     gbl(a.w64), gbl(a.V + (long)(j + 1) * a.n), gbl(a.v32), a.n, bx, r.grid.x);
}
__global__ void lz_combine_grouped(const GRec<LzArgs>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<LzArgs>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  lz_combine_body(gbl(r.a.ctrl), gbl(r.a.V), r.a.n, gbl(r.a.y), bx, r.grid.x);
}

// ------------------------------------------------------------------ measurement
// rho[p][q] = sum_{x,b} A[p][x][b] * T[q][x][b], T[q][x][b] = sum_y E[x][y] conj(A[q][y][b])
// One block; writes rho (2x2 complex) for one site.
__global__ void rho_kernel(const c128* __restrict__ A, const c128* __restrict__ T, long nxb,
                           c128* __restrict__ rho) {
  __shared__ double sh[16];
  for (int p = 0; p < 2; ++p)
    for (int q = 0; q < 2; ++q) {
      double ar = 0, ai = 0;
      for (long e = threadIdx.x; e < nxb; e += blockDim.x) {
        const c128 a = A[(long)p * nxb + e], t = T[(long)q * nxb + e];
        ar += a.re * t.re - a.im * t.im;
        ai += a.re * t.im + a.im * t.re;
      }
      ar = block_sum(ar, sh);
      ai = block_sum(ai, sh);
      if (threadIdx.x == 0) rho[p * 2 + q] = c128{ar, ai};
    }
}

// Observables of the reference's MPS.measure (mps.py:122-140) from per-site
// 2x2 density matrices: <n_i> = Re rho[1][1] (not normalised), its rounding
// (half to even, as numpy.round), and -Tr rho log2 rho from the eigenvalues of
// the Hermitian part, (a+d)/2 +- sqrt(((a-d)/2)^2 + |b|^2), with 0 log 0 = 0.
// out = [population(n) | d_population(n) | single_site_entropy(n)].
__global__ void observables_kernel(const c128* __restrict__ rho, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const c128 r00 = rho[4 * i], r01 = rho[4 * i + 1], r10 = rho[4 * i + 2], r11 = rho[4 * i + 3];
  const double a = r00.re, d = r11.re;
  const double br = 0.5 * (r01.re + r10.re), bi = 0.5 * (r01.im - r10.im);
  const double mean = 0.5 * (a + d), half = 0.5 * (a - d);
  const double rad = sqrt(half * half + br * br + bi * bi);
  const double lam[2] = {mean + rad, mean - rad};
  double sse = 0.0;
  for (int k = 0; k < 2; ++k)
    if (lam[k] > 0.0) sse -= lam[k] * log2(lam[k]);
  out[i] = d;
  out[n + i] = rint(d);
  out[2 * n + i] = sse;
}

// ------------------------------------------------------------------ exact 2^N
// (H psi)[x] = sum_i [popcount(x & nbr_i) in [lo, hi)] psi[x ^ bit_i]
// A workgroup owns one tile of 2^EX_TILE_BITS consecutive basis states (the
// whole vector when it is shorter), staged in LDS: partners across the low
// bits come from the tile, only flips of the higher bits gather from memory.
constexpr int EX_TILE_BITS = 10;
constexpr int EX_THREADS = 256;

template <typename T>
__global__ __launch_bounds__(EX_THREADS) void exact_apply_kernel(const cplx<T>* __restrict__ psi,
                                                                 cplx<T>* __restrict__ out,
                                                                 int ncells, int distance, int lo,
                                                                 int hi) {
  __shared__ cplx<T> tile[1 << EX_TILE_BITS];
  const int tbits = min(ncells, EX_TILE_BITS);
  const long tsize = 1L << tbits;
  const long base = (long)blockIdx.x << tbits;
  for (long t = threadIdx.x; t < tsize; t += EX_THREADS) tile[t] = psi[base + t];
  __syncthreads();
  for (long t = threadIdx.x; t < tsize; t += EX_THREADS) {
    const long x = base + t;
    T ar = 0, ai = 0;
    for (int i = 0; i < ncells; ++i) {
      const int jlo = max(0, i - distance), jhi = min(ncells - 1, i + distance);
      // neighbours j in [jlo, jhi], j != i; cell j is bit (N-1-j)
      const int bhi = ncells - 1 - jlo, blo = ncells - 1 - jhi;
      const unsigned long long span =
          ((bhi - blo + 1) >= 64) ? ~0ULL : (((1ULL << (bhi - blo + 1)) - 1ULL) << blo);
      const unsigned long long self = 1ULL << (ncells - 1 - i);
      const int cnt = __popcll((unsigned long long)x & span & ~self);
      if (cnt >= lo && cnt < hi) {
        const cplx<T> v = ((long)self < tsize) ? tile[t ^ (long)self] : psi[x ^ (long)self];
        ar += v.re;
        ai += v.im;
      }
    }
    out[x] = cplx<T>{ar, ai};
  }
}

// Single-site reduced density matrices of a 2^N state, two-stage reduction.
// rho_i[p][q] = sum_{x: x_i = p} psi_x conj(psi_{x with x_i = q}); site i is bit N-1-i.
// Stage 1: block (bx, i) writes partial (rho00, rho11, Re rho01, Im rho01).
__global__ void exact_rho_partials(const c128* __restrict__ psi, int ncells, double* __restrict__ part) {
  __shared__ double sh[16];
  const int i = blockIdx.y;
  const long dim = 1L << ncells;
  const long bit = 1L << (ncells - 1 - i);
  double d0 = 0, d1 = 0, orr = 0, oi = 0;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < dim; x += (long)gridDim.x * blockDim.x) {
    const c128 a = psi[x];
    const double n2 = a.re * a.re + a.im * a.im;
    if (x & bit) {
      d1 += n2;
    } else {
      d0 += n2;
      const c128 b = psi[x | bit];  // conj(b) * a summed into rho01 = sum a conj(b)
      orr += a.re * b.re + a.im * b.im;
      oi += a.im * b.re - a.re * b.im;
    }
  }
  d0 = block_sum(d0, sh);
  d1 = block_sum(d1, sh);
  orr = block_sum(orr, sh);
  oi = block_sum(oi, sh);
  if (threadIdx.x == 0) {
    double* p = part + ((long)i * gridDim.x + blockIdx.x) * 4;
    p[0] = d0; p[1] = d1; p[2] = orr; p[3] = oi;
  }
}

__global__ void exact_rho_finalize(const double* __restrict__ part, int nbx, int ncells, c128* __restrict__ rho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells) return;
  double d0 = 0, d1 = 0, orr = 0, oi = 0;
  for (int b = 0; b < nbx; ++b) {
    const double* p = part + ((long)i * nbx + b) * 4;
    d0 += p[0]; d1 += p[1]; orr += p[2]; oi += p[3];
  }
  rho[4L * i + 0] = c128{d0, 0.0};
  rho[4L * i + 1] = c128{orr, oi};
  rho[4L * i + 2] = c128{orr, -oi};
  rho[4L * i + 3] = c128{d1, 0.0};
}

int grid_for(long n, int block = 256, int maxb = 8192) {
  return (int)std::max<long>(1, std::min<long>((n + block - 1) / block, maxb));
}

}  // namespace

// ------------------------------------------------------------------ word gather
__global__ void gather_words_kernel(const void* const* __restrict__ src, int words, int n,
                                    unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long* p = reinterpret_cast<const unsigned long long*>(src[i]);
  for (int w = 0; w < words; ++w) out[(long)i * words + w] = p[w];
}

void launch_gather_words(const void* const* src, int words, int n, unsigned long long* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(gather_words_kernel, dim3(ceil_div(n, 64)), dim3(64), 0, s, src, words, n, out);
  QCA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ record arena
void RecordArena::reserve(size_t bytes) {
  if (bytes <= cap_) return;
  QCA_CHECK(used_ == 0, "record arena grown while holding records");
  release();
  QCA_HIP(hipHostMalloc((void**)&host_, bytes));
  QCA_HIP(hipMalloc((void**)&dev_, bytes));
  cap_ = bytes;
}

void RecordArena::release() {
  if (host_) (void)hipHostFree(host_);
  if (dev_) (void)hipFree(dev_);
  host_ = dev_ = nullptr;
  cap_ = used_ = uploaded_ = 0;
}

const void* RecordArena::push_bytes(const void* src, size_t bytes) {
  const size_t off = (used_ + 63) / 64 * 64;
  QCA_CHECK(off + bytes <= cap_, "record arena full");
  std::memcpy(host_ + off, src, bytes);
  used_ = off + bytes;
  return dev_ + off;
}

void RecordArena::upload(hipStream_t s) {
  if (used_ > uploaded_) {
    QCA_HIP(hipMemcpyAsync(dev_ + uploaded_, host_ + uploaded_, used_ - uploaded_, hipMemcpyHostToDevice, s));
    uploaded_ = used_;
  }
}

// ------------------------------------------------------------------ grouped launchers
GRec<ConvArgs> conv_rec(const c128* in, c64* out, long n, const int* skip) {
  return grec(ConvArgs{in, out, n, skip}, grid_for(n));
}
GRec<ThetaSvdArgs> theta_svd_rec(const c128* th, c128* X, int chil, int chir, int mode) {
  return grec(ThetaSvdArgs{th, X, chil, chir, mode}, grid_for(4L * chil * chir));
}
GRec<GatherIsoArgs> gather_iso_rec(const c128* X, const int* perm, const double* sigma, c128* A,
                                   int rows, int chi_bond, int k, int mode) {
  return grec(GatherIsoArgs{X, perm, sigma, A, rows, chi_bond, k, mode}, grid_for((long)rows * k));
}
template <typename T>
size_t wmix_lds_bytes(int nin, int nq, int Dout) {
  return ((size_t)nin * WMIX_XC + (size_t)nq * WMIX_XC * Dout) * sizeof(cplx<T>);
}
template <typename T>
GRec<WmixArgs<T>> wmix_rec(const cplx<T>* Tin, long in_stride, long nxc, int nin, cplx<T>* U, int nq,
                           int Dout, const WSparseDev<T>& w, const int* skip) {
  QCA_CHECK(Dout <= WMIX_DMAX, "MPO bond dimension above 16");
  QCA_CHECK(nin <= 64, "too many input channels in MPO mixing");
  QCA_CHECK(wmix_lds_bytes<T>(nin, nq, Dout) <= 150 * 1024, "MPO mixing stage exceeds LDS");
  return grec(WmixArgs<T>{Tin, in_stride, nxc, nin, U, nq, Dout, w.rowptr, w.col, w.val, skip},
              (unsigned)ceil_div(nxc, WMIX_XC));
}
template GRec<WmixArgs<double>> wmix_rec<double>(const c128*, long, long, int, c128*, int, int,
                                                 const WSparseDev<double>&, const int*);
template GRec<WmixArgs<float>> wmix_rec<float>(const c64*, long, long, int, c64*, int, int,
                                               const WSparseDev<float>&, const int*);
template size_t wmix_lds_bytes<double>(int, int, int);
template size_t wmix_lds_bytes<float>(int, int, int);

void launch_conv_grouped(const KGroup<ConvArgs>& g, hipStream_t s) {
  if (g.empty()) return;
  hipLaunchKernelGGL(convert_grouped, dim3(g.l.blocks), dim3(256), 0, s, g.l.recs, g.l.n);
  QCA_HIP(hipGetLastError());
}
void launch_theta_svd_grouped(const KGroup<ThetaSvdArgs>& g, hipStream_t s) {
  if (g.empty()) return;
  hipLaunchKernelGGL(theta_to_svd_grouped, dim3(g.l.blocks), dim3(256), 0, s, g.l.recs, g.l.n);
  QCA_HIP(hipGetLastError());
}
void launch_gather_iso_grouped(const KGroup<GatherIsoArgs>& g, hipStream_t s) {
  if (g.empty()) return;
  hipLaunchKernelGGL(gather_isometry_grouped, dim3(g.l.blocks), dim3(256), 0, s, g.l.recs, g.l.n);
  QCA_HIP(hipGetLastError());
}
template <typename T>
void launch_wmix_grouped(const KGroup<WmixArgs<T>>& g, size_t lds, hipStream_t s) {
  if (g.empty()) return;
  ensure_dynamic_lds((const void*)wmix_grouped<T>, lds);
  hipLaunchKernelGGL(wmix_grouped<T>, dim3(g.l.blocks), dim3(WMIX_XC * WMIX_WAVES), lds, s, g.l.recs, g.l.n);
  QCA_HIP(hipGetLastError());
}
template void launch_wmix_grouped<double>(const KGroup<WmixArgs<double>>&, size_t, hipStream_t);
template void launch_wmix_grouped<float>(const KGroup<WmixArgs<float>>&, size_t, hipStream_t);

void LzGroup::prepare(const std::vector<LzArgs>& jobs, RecordArena& arena) {
  std::vector<GRec<LzArgs>> a, b, c;
  for (const LzArgs& j : jobs) {
    QCA_CHECK(j.nblk == lanczos_nblk(j.n), "Lanczos block count does not match the vector");
    a.push_back(grec(j, j.nblk));
    b.push_back(grec(j, grid_for(j.n)));
    c.push_back(grec(j, 1));
  }
  nb.prepare(a, arena);
  gf.prepare(b, arena);
  one.prepare(c, arena);
}
void LzGroup::begin(hipStream_t s) const {
  if (empty()) return;
  hipLaunchKernelGGL(lz_norm_partials_grouped, dim3(nb.l.blocks), dim3(LZ_BLOCK), 0, s, nb.l.recs, nb.l.n);
  hipLaunchKernelGGL(lz_init_finalize_grouped, dim3(one.l.blocks), dim3(64), 0, s, one.l.recs, one.l.n);
  hipLaunchKernelGGL(lz_scale_first_grouped, dim3(gf.l.blocks), dim3(256), 0, s, gf.l.recs, gf.l.n);
  QCA_HIP(hipGetLastError());
}
template <typename TW>
void LzGroup::tail(int j, hipStream_t s) const {
  if (empty()) return;
  QCA_CHECK(j + 1 < LZ_MAXK, "Krylov dimension limit");
  hipLaunchKernelGGL(lz_project_grouped<TW>, dim3(nb.l.blocks), dim3(LZ_BLOCK), 0, s, nb.l.recs, nb.l.n, j);
  hipLaunchKernelGGL(lz_subtract_grouped<TW>, dim3(nb.l.blocks), dim3(LZ_BLOCK), 0, s, nb.l.recs, nb.l.n, j, 1);
  hipLaunchKernelGGL(lz_subtract_grouped<TW>, dim3(nb.l.blocks), dim3(LZ_BLOCK), 0, s, nb.l.recs, nb.l.n, j, 2);
  hipLaunchKernelGGL(lz_step_scale_grouped, dim3(gf.l.blocks), dim3(256), 0, s, gf.l.recs, gf.l.n, j);
  QCA_HIP(hipGetLastError());
}
template void LzGroup::tail<double>(int, hipStream_t) const;
template void LzGroup::tail<float>(int, hipStream_t) const;
void LzGroup::finish(hipStream_t s) const {
  if (empty()) return;
  hipLaunchKernelGGL(lz_combine_grouped, dim3(gf.l.blocks), dim3(256), 0, s, gf.l.recs, gf.l.n);
  QCA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ launchers
template <typename T>
void launch_wmix(const cplx<T>* Tin, long in_stride, long nxc, int nin, cplx<T>* U, int nq,
                 int Dout, const WSparseDev<T>& w, hipStream_t s, const int* skip) {
  QCA_CHECK(Dout <= WMIX_DMAX, "MPO bond dimension above 16");
  QCA_CHECK(nin <= 64, "too many input channels in MPO mixing");
  if (nxc == 0) return;
  const size_t shm = ((size_t)nin * WMIX_XC + (size_t)nq * WMIX_XC * Dout) * sizeof(cplx<T>);
  QCA_CHECK(shm <= 150 * 1024, "MPO mixing stage exceeds LDS");
  ensure_dynamic_lds((const void*)wmix_kernel<T>, shm);
  hipLaunchKernelGGL(wmix_kernel<T>, dim3(ceil_div(nxc, WMIX_XC)), dim3(WMIX_XC * WMIX_WAVES), shm, s, Tin,
                     in_stride, nxc, nin, U, nq, Dout, w.rowptr, w.col, w.val, skip);
  QCA_HIP(hipGetLastError());
}
template void launch_wmix<double>(const c128*, long, long, int, c128*, int, int,
                                  const WSparseDev<double>&, hipStream_t, const int*);
template void launch_wmix<float>(const c64*, long, long, int, c64*, int, int,
                                 const WSparseDev<float>&, hipStream_t, const int*);

void launch_convert_c128_c64(const c128* in, c64* out, long n, hipStream_t s, const int* skip) {
  if (n == 0) return;
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for(n)), dim3(256), 0, s, in, out, n, skip);
  QCA_HIP(hipGetLastError());
}
void launch_convert_c64_c128(const c64* in, c128* out, long n, hipStream_t s, const int* skip) {
  if (n == 0) return;
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for(n)), dim3(256), 0, s, in, out, n, skip);
  QCA_HIP(hipGetLastError());
}

void launch_to_colmajor(const c128* in, c128* out, int rows, int cols, hipStream_t s) {
  hipLaunchKernelGGL(to_colmajor, dim3(grid_for((long)rows * cols)), dim3(256), 0, s, in, out, rows, cols);
  QCA_HIP(hipGetLastError());
}
void launch_site_to_rightqr(const c128* A, c128* X, int chil, int chir, hipStream_t s) {
  hipLaunchKernelGGL(site_to_rightqr, dim3(grid_for(2L * chil * chir)), dim3(256), 0, s, A, X, chil, chir);
  QCA_HIP(hipGetLastError());
}
void launch_qr_gather_left(const c128* Q, const c128* X, c128* Anew, c128* C, int rows, int cols,
                           int qcols, int target, hipStream_t s) {
  const long n = (long)rows * target + (long)target * cols;
  hipLaunchKernelGGL(qr_gather_left, dim3(grid_for(n)), dim3(256), 0, s, Q, X, Anew, C, rows, cols, qcols, target);
  QCA_HIP(hipGetLastError());
}
void launch_qr_gather_right(const c128* Q, const c128* X, c128* Anew, c128* C, int chir, int cols,
                            int qcols, int target, hipStream_t s) {
  const long n = 2L * target * chir + (long)cols * target;
  hipLaunchKernelGGL(qr_gather_right, dim3(grid_for(n)), dim3(256), 0, s, Q, X, Anew, C, chir, cols, qcols, target);
  QCA_HIP(hipGetLastError());
}

void launch_theta_to_svd(const c128* th, c128* X, int chil, int chir, int mode, hipStream_t s) {
  const long n = 4L * chil * chir;
  hipLaunchKernelGGL(theta_to_svd, dim3(grid_for(n)), dim3(256), 0, s, th, X, chil, chir, mode);
  QCA_HIP(hipGetLastError());
}

void launch_gather_isometry(const c128* X, const int* perm, const double* sigma, c128* A, int rows,
                            int chi_bond, int k, int mode, hipStream_t s) {
  const long n = (long)rows * k;
  if (n == 0) return;
  hipLaunchKernelGGL(gather_isometry, dim3(grid_for(n)), dim3(256), 0, s, X, perm, sigma, A, rows,
                     chi_bond, k, mode);
  QCA_HIP(hipGetLastError());
}

int lanczos_nblk(long n) { return (int)std::max<long>(1, (n + LZ_ELEMS - 1) / LZ_ELEMS); }

const int* lanczos_done_flag(const LanczosWork& lw) { return &lw.ctrl->done; }

void lanczos_begin(const c128* x, long n, LanczosWork& lw, hipStream_t s, c64* v32) {
  const int nb = lanczos_nblk(n);
  QCA_CHECK(nb <= lw.max_blocks, "Lanczos vector longer than the workspace");
  hipLaunchKernelGGL(lz_norm_partials, dim3(nb), dim3(LZ_BLOCK), 0, s, x, n, lw.part_norm);
  hipLaunchKernelGGL(lz_init_finalize, dim3(1), dim3(64), 0, s, lw.ctrl, lw.part_norm, nb);
  hipLaunchKernelGGL(lz_scale_first, dim3(grid_for(n)), dim3(256), 0, s, x, lw.V, n,
                     (const LanczosCtrl*)lw.ctrl, v32);
  QCA_HIP(hipGetLastError());
}

template <typename TW>
void lanczos_iteration_tail(const cplx<TW>* w, long n, int j, double tau, double tol,
                            LanczosWork& lw, hipStream_t s, c64* v32) {
  QCA_CHECK(j + 1 < lw.max_m && lw.max_m <= LZ_MAXK, "Krylov dimension limit");
  const int nb = lanczos_nblk(n);
  const LanczosCtrl* cc = lw.ctrl;
  // Up to LZ_DIRECT_BLOCKS blocks every block reduces the previous partials
  // itself (no extra launch); beyond that (state vectors of the exact path)
  // that would be quadratic in the block count, so one launch per reduction
  // sums them first.
  const bool two_level = nb > LZ_DIRECT_BLOCKS;
  hipLaunchKernelGGL(lz_project<TW>, dim3(nb), dim3(LZ_BLOCK), 0, s, cc, w, lw.V, n, j, lw.part);
  if (two_level)
    hipLaunchKernelGGL(lz_reduce, dim3(j + 1), dim3(256), 0, s, lw.ctrl, (const c128*)lw.part, nb, lw.coef1, 0);
  hipLaunchKernelGGL(lz_subtract<TW>, dim3(nb), dim3(LZ_BLOCK), 0, s, cc, w, lw.w64, lw.V, n, j,
                     (const c128*)lw.part, nb, lw.coef1, lw.part2, 1,
                     two_level ? (const c128*)lw.coef1 : nullptr);
  if (two_level)
    hipLaunchKernelGGL(lz_reduce, dim3(j + 1), dim3(256), 0, s, lw.ctrl, (const c128*)lw.part2, nb, lw.coef2, 0);
  hipLaunchKernelGGL(lz_subtract<double>, dim3(nb), dim3(LZ_BLOCK), 0, s, cc, lw.w64, lw.w64, lw.V,
                     n, j, (const c128*)lw.part2, nb, lw.coef2, lw.part, 2,
                     two_level ? (const c128*)lw.coef2 : nullptr);
  if (two_level)
    hipLaunchKernelGGL(lz_reduce, dim3(1), dim3(256), 0, s, lw.ctrl, (const c128*)lw.part, nb, lw.coef2, 1);
  hipLaunchKernelGGL(lz_step_scale, dim3(grid_for(n)), dim3(256), 0, s, lw.ctrl,
                     (const c128*)lw.part, nb, (const c128*)lw.coef1, (const c128*)lw.coef2, j, tau,
                     tol, lw.max_m, (const c128*)lw.w64, lw.V + (long)(j + 1) * n, v32, n,
                     two_level ? 1 : 0);
  QCA_HIP(hipGetLastError());
}
template void lanczos_iteration_tail<double>(const c128*, long, int, double, double, LanczosWork&,
                                             hipStream_t, c64*);
template void lanczos_iteration_tail<float>(const c64*, long, int, double, double, LanczosWork&,
                                            hipStream_t, c64*);

void lanczos_finish(long n, c128* y, LanczosWork& lw, hipStream_t s) {
  hipLaunchKernelGGL(lz_combine, dim3(grid_for(n)), dim3(256), 0, s, (const LanczosCtrl*)lw.ctrl,
                     lw.V, n, y);
  QCA_HIP(hipGetLastError());
}

void launch_observables(const c128* rho, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(observables_kernel, dim3(ceil_div(n, 64)), dim3(64), 0, s, rho, n, out);
  QCA_HIP(hipGetLastError());
}

void launch_rho(const c128* A, const c128* T, long nxb, c128* rho, hipStream_t s) {
  hipLaunchKernelGGL(rho_kernel, dim3(1), dim3(256), 0, s, A, T, nxb, rho);
  QCA_HIP(hipGetLastError());
}

template <typename T>
void launch_exact_apply(const cplx<T>* psi, cplx<T>* out, int ncells, int distance, int lo, int hi,
                        hipStream_t s) {
  QCA_CHECK(ncells >= 1 && ncells <= 34, "ncells out of range for the state vector");
  const long tiles = 1L << std::max(0, ncells - EX_TILE_BITS);
  QCA_CHECK(tiles <= 0x7fffffffL, "too many tiles for one launch");
  hipLaunchKernelGGL(exact_apply_kernel<T>, dim3((unsigned)tiles), dim3(EX_THREADS), 0, s, psi, out,
                     ncells, distance, lo, hi);
  QCA_HIP(hipGetLastError());
}
template void launch_exact_apply<double>(const c128*, c128*, int, int, int, int, hipStream_t);

int exact_rho_blocks(int ncells) {
  const long dim = 1L << ncells;
  return (int)std::max<long>(1, std::min<long>((dim + 255) / 256, 512));
}

void launch_exact_rho(const c128* psi, int ncells, double* part, c128* rho, hipStream_t s) {
  const int nbx = exact_rho_blocks(ncells);
  hipLaunchKernelGGL(exact_rho_partials, dim3(nbx, ncells), dim3(256), 0, s, psi, ncells, part);
  hipLaunchKernelGGL(exact_rho_finalize, dim3((ncells + 63) / 64), dim3(64), 0, s, part, nbx, ncells, rho);
  QCA_HIP(hipGetLastError());
}
template void launch_exact_apply<float>(const c64*, c64*, int, int, int, int, hipStream_t);

}  // namespace qca
