// One-launch Krylov exponential for small blocks (gfx950).
//
// y = exp(-i tau H_eff) x for the 2-site, 1-site and bond effective
// Hamiltonians of reference algorithms/tdvp.py:298-326 (applied matrix-free)
// with the Lanczos scheme of lautils.py:44-81, entirely inside one workgroup:
//
//   * H_eff: L-contraction -> sparse MPO mixing -> R-contraction, the three
//     stages separated by workgroup barriers, in the environments' type
//     (complex128, or complex64 on FP32 for the mixed path, as on the
//     multi-launch path); the result lands, in FP64, in the registers of the
//     threads that own its elements;
//   * recurrence: FP64, classical Gram-Schmidt applied twice against the whole
//     basis (the multi-launch path's scheme), the residual in registers
//     (element e belongs to thread e % 1024);
//   * tridiagonal exponential and the stopping test on wave 0, as in the
//     multi-launch path (Taylor with scaling, lane = row of T).
//
// The multi-launch path needs ~11 launches and a host round trip per Krylov
// iteration; at small bond dimension those dominate. This kernel needs one
// launch and no host round trip per exponential; it works on one CU, so it is
// used only below a size limit chosen by the engine.  Its iterations are bound
// by memory latency, so the launcher places the operands in LDS as far as
// they fit (environments, sparse MPO, both intermediates, then leading basis
// vectors) and the kernel reaches each one through a generic pointer.
#include "grouped.hpp"
#include "wave.hpp"

namespace qca {
namespace {

constexpr int FK_THREADS = 1024;
constexpr int FK_WAVES = FK_THREADS / 64;
constexpr int FK_EPT = FUSED_KRYLOV_MAX_N / FK_THREADS;  // residual elements per thread
// Dynamic LDS of one launch (the static part, FkShared, adds ~3 KB; 160 KB per CU).
constexpr long FK_LDS_BYTES = 152 * 1024;

__device__ __forceinline__ double fk_wave_sum(double v) { return wave_sum64(v); }

// Element i of p in the compute type TC.
template <typename TC, typename TS>
__device__ __forceinline__ cplx<TC> ldc(const cplx<TS>* __restrict__ p, unsigned i) {
  const cplx<TS> v = p[i];
  return cplx<TC>{TC(v.re), TC(v.im)};
}

struct FkShared {
  c128 coef[LZ_MAXK];
  double alpha[LZ_MAXK];
  double beta[LZ_MAXK];
  c128 z[LZ_MAXK];
  double red[FK_WAVES];
  int flag;
};

// Operands of one launch, each in LDS or global memory.
template <typename TE>
struct FkCtx {
  const cplx<TE>* L;
  const cplx<TE>* R;
  const int* rowptr;
  const int* col;
  const cplx<TE>* val;
  cplx<TE>* T1;
  cplx<TE>* U;
  c128* Vl;  // leading vcap basis vectors (LDS)
  c128* Vg;  // the others (global)
  c128* wbuf;  // residual staging for the projections (LDS, n elements)
  int vcap;
  long n;
  __device__ __forceinline__ c128* basis(int i) const {
    return i < vcap ? Vl + (long)i * n : Vg + (long)i * n;
  }
};

template <typename T>
__device__ T* place(unsigned char* arena, long off, T* global) {
  return off >= 0 ? reinterpret_cast<T*>(arena + off) : global;
}

// Copies `bytes` (a multiple of 8: complex elements) from global memory into LDS.
__device__ void stage(void* dst, const void* src, long bytes) {
  const uint2* s = reinterpret_cast<const uint2*>(src);
  uint2* d = reinterpret_cast<uint2*>(dst);
  for (long i = threadIdx.x; i < bytes / 8; i += FK_THREADS) d[i] = s[i];
}

// Sum of v over the workgroup, returned to every thread.
__device__ double fk_block_sum(double v, FkShared& sh) {
  v = fk_wave_sum(v);
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (l == 0) sh.red[w] = v;
  __syncthreads();
  double r = 0;
#pragma unroll
  for (int i = 0; i < FK_WAVES; ++i) r += sh.red[i];
  __syncthreads();
  return r;
}

// sh.coef[i] = <V_i, w> for i = 0..j; w is this thread's residual elements.
// The residual goes to LDS once; wave v then reduces the coefficients
// i = v, v + 16, ... over the whole vector, so the j+1 cross-lane reductions
// run side by side instead of one after another (fixed order: deterministic).
template <typename TE>
__device__ void fk_project(const FkCtx<TE>& f, int j, const double* wr, const double* wi,
                           FkShared& sh) {
  const long n = f.n;
#pragma unroll
  for (int r = 0; r < FK_EPT; ++r) {
    const long e = threadIdx.x + (long)r * FK_THREADS;
    if (e < n) f.wbuf[e] = c128{wr[r], wi[r]};
  }
  __syncthreads();
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int i = wv; i <= j; i += FK_WAVES) {
    const c128* vi = f.basis(i);
    double ar = 0, ai = 0;
    for (long e = l; e < n; e += 64) {
      const c128 a = vi[e], x = f.wbuf[e];
      ar += a.re * x.re + a.im * x.im;
      ai += a.re * x.im - a.im * x.re;
    }
    ar = fk_wave_sum(ar);
    ai = fk_wave_sum(ai);
    if (l == 0) sh.coef[i] = c128{ar, ai};
  }
  __syncthreads();
}

// w -= sum_{i<=j} coef_i V_i
template <typename TE>
__device__ void fk_subtract(const FkCtx<TE>& f, int j, double* wr, double* wi, const FkShared& sh) {
  const long n = f.n;
  for (int i = 0; i <= j; ++i) {
    const c128 c = sh.coef[i];
    const c128* vi = f.basis(i);
#pragma unroll
    for (int r = 0; r < FK_EPT; ++r) {
      const long e = threadIdx.x + (long)r * FK_THREADS;
      if (e < n) {
        const c128 a = vi[e];
        wr[r] -= c.re * a.re - c.im * a.im;
        wi[r] -= c.re * a.im + c.im * a.re;
      }
    }
  }
}

// Sum over the S lanes of a group (S a power of two); every lane gets it.
template <typename T>
__device__ __forceinline__ void fk_segment_sum(T& re, T& im, unsigned S) {
  for (int d = (int)S >> 1; d > 0; d >>= 1) {
    re += __shfl_xor(re, d, 64);
    im += __shfl_xor(im, d, 64);
  }
}

// The H_eff result was written to wbuf by output; each thread takes its own
// elements into registers.
template <typename TE>
__device__ void fk_collect(const FkCtx<TE>& f, double* yr, double* yi) {
  __syncthreads();
#pragma unroll
  for (int r = 0; r < FK_EPT; ++r) {
    const long e = threadIdx.x + (long)r * FK_THREADS;
    const c128 y = e < f.n ? f.wbuf[e] : c128{0.0, 0.0};
    yr[r] = y.re;
    yi[r] = y.im;
  }
}

// y = H_eff v (v in LDS or global memory); y is left in this thread's registers.
// C(m, n) = sum_{k1 < K1, k2 < K2} A(m, k) B(k, n) over the workgroup, with
// A(m, k) = A[m*am + k1*ak1 + k2*ak2], B(k, n) = B[k1*bk1 + k2*bk2 + n*bn],
// C(m, n) = C[b*cb + m*cm + n] for batch b < NB (A, B shifted by ab, bb).
// Each group of S lanes owns a TM x TN output tile (register blocking: TM + TN
// loads per TM*TN complex multiply-adds) and splits its K range S ways.
struct FkGemm {
  unsigned M, N, K1, K2, NB;
  unsigned am, ak1, ak2, ab;
  unsigned bk1, bk2, bn, bb;
  unsigned cm, cb;
};

template <int TM, int TN, typename TC, typename TA, typename TB, typename TO>
__device__ void fk_gemm_tiled(const FkGemm& g, const cplx<TA>* A, const cplx<TB>* B, cplx<TO>* C,
                              unsigned S) {
  const unsigned tm = (g.M + TM - 1) / TM, tn = (g.N + TN - 1) / TN;
  const unsigned per_b = tm * tn, tiles = per_b * g.NB;
  const unsigned s = threadIdx.x & (S - 1);
  const unsigned K = g.K1 * g.K2;
  for (unsigned t = threadIdx.x / S; t < tiles; t += FK_THREADS / S) {
    const unsigned b = t / per_b, r = t - b * per_b;
    const unsigned im = r / tn, in = r - im * tn;
    const unsigned m0 = im * TM, n0 = in * TN;
    unsigned aoff[TM], boff[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) aoff[i] = b * g.ab + min(m0 + i, g.M - 1) * g.am;
#pragma unroll
    for (int j = 0; j < TN; ++j) boff[j] = b * g.bb + min(n0 + j, g.N - 1) * g.bn;
    TC cr[TM][TN], ci[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) cr[i][j] = ci[i][j] = TC(0);
    unsigned k1 = 0, k2 = s;
    while (k2 >= g.K2 && k1 < g.K1) {
      k2 -= g.K2;
      ++k1;
    }
    for (unsigned k = s; k < K; k += S) {
      const unsigned ka = k1 * g.ak1 + k2 * g.ak2, kb = k1 * g.bk1 + k2 * g.bk2;
      cplx<TC> av[TM], bv[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) av[i] = ldc<TC>(A, aoff[i] + ka);
#pragma unroll
      for (int j = 0; j < TN; ++j) bv[j] = ldc<TC>(B, boff[j] + kb);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] += av[i].re * bv[j].re - av[i].im * bv[j].im;
          ci[i][j] += av[i].re * bv[j].im + av[i].im * bv[j].re;
        }
      k2 += S;
      while (k2 >= g.K2) {
        k2 -= g.K2;
        ++k1;
      }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) fk_segment_sum(cr[i][j], ci[i][j], S);
    if (s == 0) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          if (m0 + i < g.M && n0 + j < g.N)
            C[b * g.cb + (m0 + i) * g.cm + n0 + j] = cplx<TO>{TO(cr[i][j]), TO(ci[i][j])};
    }
  }
}

// Picks the tile (4x2, 2x2 or 1x1: the largest that still gives most of the
// workgroup a tile; 4x4 would not fit the 128 registers of a 1024-thread
// workgroup) and the K split, then runs the GEMM.
template <typename TC, typename TA, typename TB, typename TO>
__device__ void fk_gemm(const FkGemm& g, const cplx<TA>* A, const cplx<TB>* B, cplx<TO>* C) {
  auto tiles = [&](unsigned tm, unsigned tn) {
    return ((g.M + tm - 1) / tm) * ((g.N + tn - 1) / tn) * g.NB;
  };
  auto split = [&](unsigned n) {
    unsigned S = 1;
    while (S < 16 && n * S * 2 <= FK_THREADS) S *= 2;
    return S;
  };
  if (tiles(4, 2) >= FK_THREADS / 2) fk_gemm_tiled<4, 2, TC>(g, A, B, C, 1);
  else if (tiles(2, 2) >= FK_THREADS / 4) fk_gemm_tiled<2, 2, TC>(g, A, B, C, split(tiles(2, 2)));
  else fk_gemm_tiled<1, 1, TC>(g, A, B, C, split(tiles(1, 1)));
}

template <typename TE>
__device__ void fk_heff(const FusedKrylovArgs<TE>& a, const FkCtx<TE>& f, const c128* v, double* yr,
                        double* yi) {
  // all index arithmetic in 32 bits: every extent here is below 2^20
  const unsigned chil = a.chil, chir = a.chir, Din = a.Din, Dout = a.Dout;
  const unsigned lr = chil * chir;
  const unsigned dq = a.kind == 0 ? 4 : (a.kind == 1 ? 2 : 1);
  // T1[q][(w,x)][c] = sum_a L[a][(w,x)] v[q][a][c]
  {
    FkGemm g;
    g.M = Din * chil; g.N = chir; g.K1 = 1; g.K2 = chil; g.NB = dq;
    g.am = 1; g.ak1 = 0; g.ak2 = Din * chil; g.ab = 0;
    g.bk1 = 0; g.bk2 = chir; g.bn = 1; g.bb = lr;
    g.cm = chir; g.cb = Din * lr;
    fk_gemm<TE>(g, f.L, v, f.T1);
  }
  __syncthreads();
  if (a.kind == 2) {
    // y[x][yy] = sum_{w,c} T1[w][x][c] R[c][w][yy]
    FkGemm g;
    g.M = chil; g.N = chir; g.K1 = Din; g.K2 = chir; g.NB = 1;
    g.am = chir; g.ak1 = lr; g.ak2 = 1; g.ab = 0;
    g.bk1 = chir; g.bk2 = Din * chir; g.bn = 1; g.bb = 0;
    g.cm = chir; g.cb = 0;
    fk_gemm<TE>(g, f.T1, f.R, f.wbuf);
    fk_collect(f, yr, yi);
    return;
  }
  // U[q'][xc][v] = sum over the sparse row (q',v) of W * T1[(q,w)][xc]
  const unsigned e2 = dq * Dout * lr;
  for (unsigned e = threadIdx.x; e < e2; e += FK_THREADS) {
    const unsigned o = e / lr, xc = e - o * lr;
    TE re = 0, im = 0;
    const int z1 = f.rowptr[o + 1];
    for (int zz = f.rowptr[o]; zz < z1; ++zz) {
      const cplx<TE> wv = f.val[zz];
      const cplx<TE> t = f.T1[(unsigned)f.col[zz] * lr + xc];
      re += wv.re * t.re - wv.im * t.im;
      im += wv.re * t.im + wv.im * t.re;
    }
    const unsigned qp = o / Dout, vv = o - qp * Dout;
    f.U[(qp * lr + xc) * Dout + vv] = cplx<TE>{re, im};
  }
  __syncthreads();
  // y[(q',x)][yy] = sum_{(c,v)} U[(q',x)][(c,v)] R[(c,v)][yy]
  {
    FkGemm g;
    g.M = dq * chil; g.N = chir; g.K1 = 1; g.K2 = chir * Dout; g.NB = 1;
    g.am = chir * Dout; g.ak1 = 0; g.ak2 = 1; g.ab = 0;
    g.bk1 = 0; g.bk2 = chir; g.bn = 1; g.bb = 0;
    g.cm = chir; g.cb = 0;
    fk_gemm<TE>(g, f.U, f.R, f.wbuf);
  }
  fk_collect(f, yr, yi);
}

// Wave 0: z = exp(-i tau T_m) e_0 for the m = j+1 leading rows of the
// tridiagonal T (alpha on the diagonal, beta off it), by Taylor steps with
// scaling (lane = row); returns |beta_j z_j| on every lane.
__device__ double fk_tridiag_exp(const FkShared& sh, int j, double beta, double tau, double& zr,
                                 double& zi) {
  const int l = threadIdx.x & 63;
  const int m = j + 1;
  const double al = (l < m) ? sh.alpha[l] : 0.0;
  const double bup = (l + 1 < m) ? sh.beta[l] : 0.0;
  const double bdn = (l >= 1 && l < m) ? sh.beta[l - 1] : 0.0;
  double rown = fabs(al) + fabs(bup) + fabs(bdn);
  for (int o = 32; o > 0; o >>= 1) rown = fmax(rown, __shfl_xor(rown, o, 64));
  const double h = fabs(tau) * rown;
  const int nsub = (h < 2048.0) ? max(1, (int)ceil(h / 0.5)) : 1;
  const double dt = tau / nsub;
  zr = (l == 0) ? 1.0 : 0.0;
  zi = 0.0;
  for (int sstep = 0; sstep < nsub; ++sstep) {
    double ar = zr, ai = zi, tr = zr, ti = zi;
    for (int k = 1; k < 60; ++k) {
      const double upr = __shfl_down(tr, 1, 64), upi = __shfl_down(ti, 1, 64);
      const double dnr = __shfl_up(tr, 1, 64), dni = __shfl_up(ti, 1, 64);
      double yr = al * tr + bup * upr + bdn * dnr;
      double yi = al * ti + bup * upi + bdn * dni;
      if (l >= m) {
        yr = 0;
        yi = 0;
      }
      const double ck = dt / k;
      const double nr = ck * yi, ni = -ck * yr;
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
  return beta * sqrt(zlr * zlr + zli * zli);
}

template <typename TE>
__device__ __forceinline__ void fused_krylov_body(const FusedKrylovArgs<TE>& a, unsigned char* arena) {
  __shared__ FkShared sh;
  const long n = a.n;
  const unsigned long long t0 = wall_clock64();
  FkCtx<TE> f;
  f.L = place(arena, a.offL, a.L);
  f.R = place(arena, a.offR, a.R);
  f.rowptr = place(arena, a.offRow, a.rowptr);
  f.col = place(arena, a.offCol, a.col);
  f.val = place(arena, a.offVal, a.val);
  f.T1 = place(arena, a.offT1, a.T1);
  f.U = place(arena, a.offU, a.U);
  f.Vl = place(arena, a.offV, a.V);
  f.Vg = a.V;
  f.wbuf = reinterpret_cast<c128*>(arena + a.offW);
  f.vcap = a.vcap;
  f.n = n;
  const int dq = a.kind == 0 ? 4 : (a.kind == 1 ? 2 : 1);
  if (a.offL >= 0) stage(arena + a.offL, a.L, (long)a.chil * a.Din * a.chil * sizeof(cplx<TE>));
  if (a.offR >= 0) stage(arena + a.offR, a.R, (long)a.chir * a.Dout * a.chir * sizeof(cplx<TE>));
  if (a.offVal >= 0) stage(arena + a.offVal, a.val, (long)a.nnz * sizeof(cplx<TE>));
  if (a.offRow >= 0) {
    int* rp = reinterpret_cast<int*>(arena + a.offRow);
    for (int i = threadIdx.x; i <= dq * a.Dout; i += FK_THREADS) rp[i] = a.rowptr[i];
  }
  if (a.offCol >= 0) {
    int* cp = reinterpret_cast<int*>(arena + a.offCol);
    for (int i = threadIdx.x; i < a.nnz; i += FK_THREADS) cp[i] = a.col[i];
  }
  double wr[FK_EPT], wi[FK_EPT];
#pragma unroll
  for (int r = 0; r < FK_EPT; ++r) {
    const long e = threadIdx.x + (long)r * FK_THREADS;
    wr[r] = e < n ? a.x[e].re : 0.0;
    wi[r] = e < n ? a.x[e].im : 0.0;
  }
  double nrm = 0;
#pragma unroll
  for (int r = 0; r < FK_EPT; ++r) nrm += wr[r] * wr[r] + wi[r] * wi[r];
  const double beta0 = sqrt(fk_block_sum(nrm, sh));
  int m = 0, hit_limit = 0;
  double err = 0;
  if (beta0 > 0.0) {
    const double inv0 = 1.0 / beta0;
#pragma unroll
    for (int r = 0; r < FK_EPT; ++r) {
      const long e = threadIdx.x + (long)r * FK_THREADS;
      if (e < n) f.basis(0)[e] = c128{wr[r] * inv0, wi[r] * inv0};
    }
    __syncthreads();
    for (int j = 0;; ++j) {
      fk_heff<TE>(a, f, f.basis(j), wr, wi);
      fk_project(f, j, wr, wi, sh);
      const double c1 = sh.coef[j].re;
      fk_subtract(f, j, wr, wi, sh);
      fk_project(f, j, wr, wi, sh);
      const double c2 = sh.coef[j].re;
      fk_subtract(f, j, wr, wi, sh);
      double s = 0;
#pragma unroll
      for (int r = 0; r < FK_EPT; ++r) s += wr[r] * wr[r] + wi[r] * wi[r];
      const double beta = sqrt(fk_block_sum(s, sh));
      const double alpha_j = c1 + c2;
      if (threadIdx.x < 64) {
        if (threadIdx.x == 0) {
          sh.alpha[j] = alpha_j;
          sh.beta[j] = beta;
        }
        // wave 0 reads its own writes: wave-local LDS order holds
        __builtin_amdgcn_wave_barrier();
        double zr, zi;
        const double e = fk_tridiag_exp(sh, j, beta, a.tau, zr, zi);
        const bool breakdown = beta <= 1e-13 * fmax(1.0, fabs(alpha_j));
        const bool conv = breakdown || e < a.tol || j + 1 >= a.max_m - 1;
        if ((int)threadIdx.x <= j) sh.z[threadIdx.x] = c128{zr, zi};
        if (threadIdx.x == 0) {
          sh.flag = conv ? 1 : 0;
          sh.red[0] = e;
          sh.red[1] = (!breakdown && e >= a.tol) ? 1.0 : 0.0;
        }
      }
      __syncthreads();
      const bool done = sh.flag != 0;
      err = sh.red[0];
      hit_limit = sh.red[1] != 0.0;
      __syncthreads();
      if (done) {
        m = j + 1;
        break;
      }
      const double inv = 1.0 / beta;
      c128* vn = f.basis(j + 1);
#pragma unroll
      for (int r = 0; r < FK_EPT; ++r) {
        const long e = threadIdx.x + (long)r * FK_THREADS;
        if (e < n) vn[e] = c128{wr[r] * inv, wi[r] * inv};
      }
      __syncthreads();
    }
  }
  // x = beta0 * sum_{i<m} z_i V_i  (a zero input stays zero)
#pragma unroll
  for (int r = 0; r < FK_EPT; ++r) {
    const long e = threadIdx.x + (long)r * FK_THREADS;
    if (e >= n) continue;
    double re = 0, im = 0;
    for (int i = 0; i < m; ++i) {
      const c128 zc = sh.z[i];
      const c128 v = f.basis(i)[e];
      re += zc.re * v.re - zc.im * v.im;
      im += zc.re * v.im + zc.im * v.re;
    }
    a.x[e] = c128{beta0 * re, beta0 * im};
  }
  if (threadIdx.x == 0) {
    if (a.ctrl) {
      a.ctrl->beta0 = beta0;
      a.ctrl->err = err;
      a.ctrl->m = m;
      a.ctrl->done = 1;
      a.ctrl->zero = beta0 == 0.0;
      a.ctrl->hit_limit = hit_limit;
    }
    if (a.tally) {
      // stream-ordered per engine: plain read-modify-write is race free
      a.tally->matvecs[a.kind] += m;
      a.tally->limit_hits += hit_limit;
      a.tally->flops += m * a.flops_per_matvec;
      a.tally->ticks += wall_clock64() - t0;
      a.tally->calls += 1;
    }
  }
}

template <typename TE>
__global__ __launch_bounds__(FK_THREADS) void fused_krylov_kernel(FusedKrylovArgs<TE> a) {
  extern __shared__ __align__(16) unsigned char arena[];
  fused_krylov_body<TE>(a, arena);
}

// One exponential per workgroup, each with its own arguments and LDS placement
// (the dynamic LDS of the launch is the largest placement).
template <typename TE>
__global__ __launch_bounds__(FK_THREADS) void fused_krylov_grouped(const GRec<FusedKrylovArgs<TE>>* __restrict__ recs,
                                                                   int n) {
  extern __shared__ __align__(16) unsigned char arena[];
  FusedKrylovArgs<TE> a = recs[blockIdx.x].a;
  a.L = gbl(a.L);
  a.R = gbl(a.R);
  a.rowptr = gbl(a.rowptr);
  a.col = gbl(a.col);
  a.val = gbl(a.val);
  a.x = gbl(a.x);
  a.V = gbl(a.V);
  a.T1 = gbl(a.T1);
  a.U = gbl(a.U);
  a.ctrl = gbl(a.ctrl);
  a.tally = gbl(a.tally);
  fused_krylov_body<TE>(a, arena);
}

}  // namespace

template <typename TE>
size_t fused_krylov_place(const FusedKrylovArgs<TE>& a, FusedKrylovArgs<TE>& b) {
  QCA_CHECK(a.n >= 1 && a.n <= FUSED_KRYLOV_MAX_N, "fused Krylov block above its size limit");
  QCA_CHECK(a.kind >= 0 && a.kind <= 2, "unknown effective-Hamiltonian kind");
  const int dq = a.kind == 0 ? 4 : (a.kind == 1 ? 2 : 1);
  QCA_CHECK(a.n == (long)dq * a.chil * a.chir, "Krylov vector length does not match the block");
  QCA_CHECK(a.max_m >= 2 && a.max_m <= LZ_MAXK, "Krylov dimension limit out of range");
  QCA_CHECK(a.kind == 2 || (a.rowptr && a.col && a.val), "sparse MPO missing");
  QCA_CHECK(a.kind != 2 || a.Din == a.Dout, "bond block needs equal MPO bonds");
  QCA_CHECK(a.kind == 2 || a.nnz >= 0, "negative MPO entry count");
  QCA_CHECK(a.kind == 2 || a.Dout <= 16, "MPO bond dimension above 16");
  // LDS placement, in order of use per Krylov iteration
  b = a;
  const long lr = (long)a.chil * a.chir;
  long off = a.n * (long)sizeof(c128);  // residual staging first: always in LDS
  b.offW = 0;
  auto take = [&](long bytes) -> long {
    bytes = (bytes + 15) / 16 * 16;
    if (off + bytes > FK_LDS_BYTES) return -1;
    const long o = off;
    off += bytes;
    return o;
  };
  b.offL = take((long)a.chil * a.Din * a.chil * sizeof(cplx<TE>));
  b.offR = take((long)a.chir * a.Dout * a.chir * sizeof(cplx<TE>));
  b.offRow = b.offCol = b.offVal = b.offU = -1;
  if (a.kind != 2) {
    b.offRow = take((long)(dq * a.Dout + 1) * sizeof(int));
    b.offCol = take((long)a.nnz * sizeof(int));
    b.offVal = take((long)a.nnz * sizeof(cplx<TE>));
    b.offU = take((long)dq * a.Dout * lr * sizeof(cplx<TE>));
  }
  b.offT1 = take((long)dq * a.Din * lr * sizeof(cplx<TE>));
  const long vbytes = a.n * (long)sizeof(c128);
  b.vcap = (int)std::min<long>(a.max_m - 1, (FK_LDS_BYTES - off) / vbytes);
  b.offV = b.vcap > 0 ? off : -1;
  if (b.vcap > 0) off += b.vcap * vbytes;
  return (size_t)off;
}
template size_t fused_krylov_place<double>(const FusedKrylovArgs<double>&, FusedKrylovArgs<double>&);
template size_t fused_krylov_place<float>(const FusedKrylovArgs<float>&, FusedKrylovArgs<float>&);

template <typename TE>
void launch_fused_krylov(const FusedKrylovArgs<TE>& a, hipStream_t s) {
  FusedKrylovArgs<TE> b;
  const size_t off = fused_krylov_place(a, b);
  ensure_dynamic_lds((const void*)fused_krylov_kernel<TE>, off);
  hipLaunchKernelGGL(fused_krylov_kernel<TE>, dim3(1), dim3(FK_THREADS), off, s, b);
  QCA_HIP(hipGetLastError());
}
template void launch_fused_krylov<double>(const FusedKrylovArgs<double>&, hipStream_t);
template void launch_fused_krylov<float>(const FusedKrylovArgs<float>&, hipStream_t);

template <typename TE>
void launch_fused_krylov_grouped(const KGroup<FusedKrylovArgs<TE>>& g, size_t lds, hipStream_t s) {
  if (g.empty()) return;
  ensure_dynamic_lds((const void*)fused_krylov_grouped<TE>, lds);
  hipLaunchKernelGGL(fused_krylov_grouped<TE>, dim3(g.l.blocks), dim3(FK_THREADS), lds, s, g.l.recs, g.l.n);
  QCA_HIP(hipGetLastError());
}
template void launch_fused_krylov_grouped<double>(const KGroup<FusedKrylovArgs<double>>&, size_t, hipStream_t);
template void launch_fused_krylov_grouped<float>(const KGroup<FusedKrylovArgs<float>>&, size_t, hipStream_t);

}  // namespace qca
