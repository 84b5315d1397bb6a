// Strided, batched complex GEMM on CDNA4 matrix cores (gfx950).
//
// One kernel body serves the plain path (complex128 on v_mfma_f64_16x16x4f64)
// and the mixed path (complex64 on v_mfma_f32_16x16x4f32).  A complex product
// is four real MFMAs on planar (re, im) LDS tiles:
//   Cre += Are*Bre - Aim*Bim,   Cim += Are*Bim + Aim*Bre.
// Tile 64x64x16, 256 threads = 4 wave64s in a 2x2 grid, each wave owning a
// 32x32 block = 2x2 MFMA 16x16 tiles.  Operands are read with generic two-level
// strides (see GemmDesc) so that every TDVP contraction is a single launch with
// no permutation copies; the load loop walks the unit-stride index so global
// reads stay coalesced; the next K-tile is loaded into registers while the
// current one is multiplied.  Launches with too few output tiles to fill 256
// CUs split K and reduce the partial tiles in a second, deterministic pass.
#include "grouped.hpp"

namespace qca {
namespace {

constexpr int BM = 64, BN = 64, BK = 16, PAD = 1;
constexpr int LOADS = BM * BK / 256;  // operand elements per thread per K-tile

template <typename T>
struct Mfma;
// Both instructions take A[i][k] from lane i + 16k and B[k][j] from lane
// j + 16k, and return column j = lane % 16, but they differ in which row the
// four accumulator registers hold (measured on gfx950):
//   f64 16x16x4: register r of lane l holds row (l / 16) + 4 r
//   f32 16x16x4: register r of lane l holds row 4 (l / 16) + r
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// KG K-groups of 256 threads share a workgroup: group g takes the K-tiles
// g, g + KG, ... of the workgroup's range into its own LDS buffers, and the
// groups' partial tiles are summed through LDS at the end.  With KG = 2 a CU
// holding one workgroup runs two waves per SIMD, so one group's staging and
// barrier waits overlap the other's MFMAs.
// One output tile (bx: N tile, by: M tile, bz: batch * splits + split) of the
// GEMM d; the kernels below supply the block index (direct launch) or a
// record's virtual one (grouped launch over many instances).
template <typename T, bool A_MCONT, bool B_NCONT, int KG>
__device__ __forceinline__ void gemm_body(const GemmDesc& d, const cplx<T>* __restrict__ A,
                                          const cplx<T>* __restrict__ B, cplx<T>* __restrict__ C,
                                          cplx<T>* __restrict__ ws, int splits, int k_per_split,
                                          unsigned bx, unsigned by, unsigned bz) {
  using acc_t = typename Mfma<T>::acc_t;
  static_assert(BM == BN, "one LDS stride for both operands");
  if (d.skip && *d.skip) return;
  constexpr int LS = BM + PAD;
  // [4 planes: A re, A im, B re, B im][KG][BK][LS]
  __shared__ T smem[4 * KG * BK * LS];
  static_assert(KG == 1 || 4 * KG * BK * LS >= 4 * 32 * 64, "partial-tile exchange fits the operand LDS");
  auto lds = [&](int plane, int g, int k, int x) -> T& {
    return smem[((plane * KG + g) * BK + k) * LS + x];
  };

  const int tid = threadIdx.x & 255;  // thread within its K-group
  const int grp = threadIdx.x >> 8;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32;
  const int wn = (wave & 1) * 32;
  const int m0 = by * BM;
  const int n0 = bx * BN;
  const int zb = bz / splits;
  const int split = bz - zb * splits;
  const int b1 = zb / d.NB2;
  const int b2 = zb - b1 * d.NB2;

  const cplx<T>* Ab = A + b1 * d.a_b1 + b2 * d.a_b2;
  const cplx<T>* Bb = B + b1 * d.b_b1 + b2 * d.b_b2;

  const int k_begin = split * k_per_split;
  const int k_end = min(d.K, k_begin + k_per_split);

  acc_t acc_re[2][2], acc_im[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_re[i][j] = acc_t{0, 0, 0, 0};
      acc_im[i][j] = acc_t{0, 0, 0, 0};
    }

  const T sa = d.conj_a ? T(-1) : T(1);
  const T sb = d.conj_b ? T(-1) : T(1);

  // Operand elements of one K-tile, loaded into registers one tile ahead so
  // that the global loads of tile k0+BK overlap the MFMAs of tile k0.
  const bool flatK = d.K2 >= d.K;  // single-level K: no k1/k2 split
  cplx<T> pa[LOADS], pb[LOADS];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const int e = tid + 256 * r;
      const int ml = A_MCONT ? (e % BM) : (e / BK);
      const int kl = A_MCONT ? (e / BM) : (e % BK);
      const int m = m0 + ml, k = k0 + kl;
      cplx<T> v{T(0), T(0)};
      if (m < d.M && k < k_end) {
        if (flatK) {
          v = Ab[m * d.a_m + k * d.a_k2];
        } else {
          const int k1 = k / d.K2, k2 = k - k1 * d.K2;
          v = Ab[m * d.a_m + k1 * d.a_k1 + k2 * d.a_k2];
        }
      }
      pa[r] = v;
    }
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const int e = tid + 256 * r;
      const int nl = B_NCONT ? (e % BN) : (e / BK);
      const int kl = B_NCONT ? (e / BN) : (e % BK);
      const int n = n0 + nl, k = k0 + kl;
      cplx<T> v{T(0), T(0)};
      if (n < d.N && k < k_end) {
        if (flatK) {
          v = Bb[k * d.b_k2 + n * d.b_n];
        } else {
          const int k1 = k / d.K2, k2 = k - k1 * d.K2;
          v = Bb[k1 * d.b_k1 + k2 * d.b_k2 + n * d.b_n];
        }
      }
      pb[r] = v;
    }
  };

  // tiles beyond k_end load as zeros (no memory access), so every group can
  // run the same number of iterations
  load_tile(k_begin + grp * BK);
  for (int k0 = k_begin; k0 < k_end; k0 += KG * BK) {
    const int kt = k0 + grp * BK;
    // ---- registers -> LDS (re/im planes), then prefetch this group's next tile
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const int e = tid + 256 * r;
      const int ml = A_MCONT ? (e % BM) : (e / BK);
      const int kl = A_MCONT ? (e / BM) : (e % BK);
      lds(0, grp, kl, ml) = pa[r].re;
      lds(1, grp, kl, ml) = sa * pa[r].im;
    }
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const int e = tid + 256 * r;
      const int nl = B_NCONT ? (e % BN) : (e / BK);
      const int kl = B_NCONT ? (e / BN) : (e % BK);
      lds(2, grp, kl, nl) = pb[r].re;
      lds(3, grp, kl, nl) = sb * pb[r].im;
    }
    __syncthreads();
    if (kt + KG * BK < k_end) load_tile(kt + KG * BK);
    else if (k0 + KG * BK < k_end) load_tile(k_end);  // this group idles in the last round: zeros

    // a group without a tile in the last round skips its MFMAs (its registers
    // hold zeros; the barriers stay outside)
    if (kt < k_end)
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      const int c = lane & 15;
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = lds(0, grp, kr, wm + 16 * i + c);
        ai[i] = lds(1, grp, kr, wm + 16 * i + c);
        br[i] = lds(2, grp, kr, wn + 16 * i + c);
        bi[i] = lds(3, grp, kr, wn + 16 * i + c);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc_re[i][j] = Mfma<T>::mac(ar[i], br[j], acc_re[i][j]);
          acc_im[i][j] = Mfma<T>::mac(ar[i], bi[j], acc_im[i][j]);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc_re[i][j] = Mfma<T>::mac(-ai[i], bi[j], acc_re[i][j]);
          acc_im[i][j] = Mfma<T>::mac(ai[i], br[j], acc_im[i][j]);
        }
    }
    __syncthreads();
  }

  if constexpr (KG == 2) {
    // group 1 hands its partial tile to group 0 through the (now free) operand LDS
    T* xs = smem + (wave * 32) * 64 + lane;  // [wave][value][lane]
    if (grp == 1) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            xs[(((i * 2 + j) * 4 + r) * 2 + 0) * 64] = acc_re[i][j][r];
            xs[(((i * 2 + j) * 4 + r) * 2 + 1) * 64] = acc_im[i][j][r];
          }
    }
    __syncthreads();
    if (grp == 1) return;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          acc_re[i][j][r] += xs[(((i * 2 + j) * 4 + r) * 2 + 0) * 64];
          acc_im[i][j][r] += xs[(((i * 2 + j) * 4 + r) * 2 + 1) * 64];
        }
  }
  // ---- epilogue: column lane%16, rows per Mfma<T>::row of each 16x16 tile
  const int col = lane & 15;
  const T al_re = T(d.alpha_re), al_im = T(d.alpha_im);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + Mfma<T>::row(lane, r);
        const int n = n0 + wn + 16 * j + col;
        if (m >= d.M || n >= d.N) continue;
        const T xr = acc_re[i][j][r], xi = acc_im[i][j][r];
        if (splits > 1) {
          ws[((long)bz * d.M + m) * d.N + n] = cplx<T>{xr, xi};
        } else {
          cplx<T>* cp = C + b1 * d.c_b1 + b2 * d.c_b2 + m * d.c_m + n * d.c_n;
          cplx<T> out{al_re * xr - al_im * xi, al_re * xi + al_im * xr};
          if (d.accumulate) {
            out.re += cp->re;
            out.im += cp->im;
          }
          *cp = out;
        }
      }
}

template <typename T, bool A_MCONT, bool B_NCONT, int KG>
__global__ __launch_bounds__(256 * KG) void gemm_kernel(GemmDesc d, const cplx<T>* __restrict__ A,
                                                        const cplx<T>* __restrict__ B,
                                                        cplx<T>* __restrict__ C,
                                                        cplx<T>* __restrict__ ws, int splits,
                                                        int k_per_split) {
  gemm_body<T, A_MCONT, B_NCONT, KG>(d, A, B, C, ws, splits, k_per_split, blockIdx.x, blockIdx.y,
                                     blockIdx.z);
}

// Sum the split-K partial tiles in split order (deterministic) and apply
// alpha/beta while scattering into C's strides; block `blk` of `nblk`, 256
// threads each.
template <typename T>
__device__ __forceinline__ void splitk_reduce_body(const GemmDesc& d, const cplx<T>* __restrict__ ws,
                                                   cplx<T>* __restrict__ C, int splits, int batch,
                                                   unsigned blk, unsigned nblk) {
  if (d.skip && *d.skip) return;
  const long mn = (long)d.M * d.N;
  const long total = mn * batch;
  for (long idx = blk * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)nblk * blockDim.x) {
    const int zb = (int)(idx / mn);
    const long r = idx - zb * mn;
    const int m = (int)(r / d.N), n = (int)(r - (long)m * d.N);
    T xr = 0, xi = 0;
    // unrolled so that several partial loads are in flight; the sum stays in
    // split order
#pragma unroll 8
    for (int s = 0; s < splits; ++s) {
      const cplx<T> v = ws[((long)(zb * splits + s) * d.M + m) * d.N + n];
      xr += v.re;
      xi += v.im;
    }
    const int b1 = zb / d.NB2, b2 = zb - b1 * d.NB2;
    cplx<T>* cp = C + b1 * d.c_b1 + b2 * d.c_b2 + m * d.c_m + n * d.c_n;
    const T al_re = T(d.alpha_re), al_im = T(d.alpha_im);
    cplx<T> out{al_re * xr - al_im * xi, al_re * xi + al_im * xr};
    if (d.accumulate) {
      out.re += cp->re;
      out.im += cp->im;
    }
    *cp = out;
  }
}

template <typename T>
__global__ void splitk_reduce(GemmDesc d, const cplx<T>* __restrict__ ws, cplx<T>* __restrict__ C,
                              int splits, int batch) {
  splitk_reduce_body<T>(d, ws, C, splits, batch, blockIdx.x, gridDim.x);
}

// Launch decisions of one GEMM, a function of its own descriptor and
// workspace only (so a GEMM computes the same whether launched alone or
// grouped with others).
struct GemmPlan {
  bool empty = false;
  int tm = 0, tn = 0, batch = 1, splits = 1, kps = 0;
  bool am = false, bn = false, kg2 = false;
  int reduce_blocks = 0;
};

GemmPlan gemm_plan(const GemmDesc& d, long ws_elems) {
  QCA_CHECK(d.M >= 0 && d.N >= 0 && d.K >= 0, "negative GEMM extent");
  QCA_CHECK(d.K2 >= 1 && (d.K == 0 || d.K % d.K2 == 0), "K2 must divide K");
  QCA_CHECK(d.NB1 >= 1 && d.NB2 >= 1, "empty batch");
  GemmPlan p;
  p.batch = d.NB1 * d.NB2;
  if (d.M == 0 || d.N == 0) {
    p.empty = true;
    return p;
  }
  p.tm = ceil_div(d.M, BM);
  p.tn = ceil_div(d.N, BN);
  const long tiles = (long)p.tm * p.tn * p.batch;
  // split K when the output tiles alone cannot occupy the chip
  int splits = 1;
  if (tiles < 192 && d.K >= 4 * BK) {
    splits = (int)std::min<long>((256 + tiles - 1) / tiles, d.K / (2 * BK));
    while (splits > 1 && (long)splits * p.batch * d.M * d.N > ws_elems) --splits;
    if (splits < 1) splits = 1;
  }
  int kps = d.K;
  if (splits > 1) {
    kps = ceil_div(ceil_div(d.K, splits), BK) * BK;
    splits = ceil_div(d.K, kps);
  }
  if (d.K == 0) {
    // C = beta*C: only zeroing is needed when not accumulating
    splits = 1;
    kps = 0;
  }
  QCA_CHECK((long)tiles * splits < (1L << 31), "GEMM grid too large");
  QCA_CHECK(p.batch * splits <= 65535, "GEMM batch*splits exceeds grid z limit");
  p.splits = splits;
  p.kps = kps;
  p.am = (d.a_m == 1);
  p.bn = (d.b_n == 1);
  // two K-groups per workgroup when each has at least two K-tiles to take
  p.kg2 = kps >= 4 * BK;
  if (splits > 1) {
    const long total = (long)d.M * d.N * p.batch;
    p.reduce_blocks = (int)std::min<long>(ceil_div(total, 256), 4096);
  }
  return p;
}

template <typename T>
void gemm_launch(const GemmDesc& d, const cplx<T>* A, const cplx<T>* B, cplx<T>* C,
                 hipStream_t s, cplx<T>* ws, long ws_elems) {
  const GemmPlan p = gemm_plan(d, ws_elems);
  if (p.empty) return;
  const int splits = p.splits, kps = p.kps;
  dim3 grid(p.tn, p.tm, p.batch * splits);
  const bool am = p.am, bn = p.bn, kg2 = p.kg2;
#define QCA_GEMM_L(AM, BNC)                                                                     \
  do {                                                                                         \
    if (kg2)                                                                                   \
      hipLaunchKernelGGL((gemm_kernel<T, AM, BNC, 2>), grid, dim3(512), 0, s, d, A, B, C, ws,  \
                         splits, kps);                                                         \
    else                                                                                       \
      hipLaunchKernelGGL((gemm_kernel<T, AM, BNC, 1>), grid, dim3(256), 0, s, d, A, B, C, ws,  \
                         splits, kps);                                                         \
  } while (0)
  if (am && bn) QCA_GEMM_L(true, true);
  else if (am) QCA_GEMM_L(true, false);
  else if (bn) QCA_GEMM_L(false, true);
  else QCA_GEMM_L(false, false);
#undef QCA_GEMM_L
  QCA_HIP(hipGetLastError());
  if (splits > 1) {
    hipLaunchKernelGGL(splitk_reduce<T>, dim3(p.reduce_blocks), dim3(256), 0, s, d, ws, C, splits,
                       p.batch);
    QCA_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------------------ grouped
template <typename T, bool A_MCONT, bool B_NCONT, int KG>
__global__ __launch_bounds__(256 * KG) void gemm_grouped_kernel(const GRec<GemmArgs<T>>* __restrict__ recs,
                                                                int n, int j) {
  unsigned bx, by, bz;
  const GRec<GemmArgs<T>>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const GemmDesc d = r.a.d;
  gemm_body<T, A_MCONT, B_NCONT, KG>(d, r.a.A, r.a.B + (long)j * r.a.b_jstride, r.a.C, r.a.ws, r.a.splits,
                                     r.a.kps, bx, by, bz);
}

template <typename T>
__global__ void splitk_reduce_grouped(const GRec<GemmArgs<T>>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<GemmArgs<T>>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const GemmDesc d = r.a.d;
  splitk_reduce_body<T>(d, r.a.ws, r.a.C, r.a.splits, d.NB1 * d.NB2, bx, r.grid.x);
}

}  // namespace

template <typename T>
void GemmGroup<T>::prepare(const std::vector<GemmJob<T>>& jobs, RecordArena& arena) {
  std::vector<GRec<GemmArgs<T>>> var[8], red;
  for (const auto& j : jobs) {
    const GemmPlan p = gemm_plan(j.d, j.ws_elems);
    if (p.empty) continue;
    GRec<GemmArgs<T>> r{};
    r.a.d = j.d;
    r.a.A = j.A;
    r.a.B = j.B;
    r.a.C = j.C;
    r.a.ws = j.ws;
    r.a.b_jstride = j.b_jstride;
    r.a.splits = p.splits;
    r.a.kps = p.kps;
    r.grid.x = p.tn;
    r.grid.y = p.tm;
    r.grid.z = p.batch * p.splits;
    var[(p.am ? 4 : 0) + (p.bn ? 2 : 0) + (p.kg2 ? 1 : 0)].push_back(r);
    if (p.splits > 1) {
      r.grid = VGrid{};
      r.grid.x = p.reduce_blocks;
      red.push_back(r);
    }
  }
  for (int v = 0; v < 8; ++v) {
    launches[v].n = (int)var[v].size();
    launches[v].blocks = grec_layout(var[v]);
    launches[v].recs = var[v].empty() ? nullptr : arena.push(var[v]);
  }
  reduce.n = (int)red.size();
  reduce.blocks = grec_layout(red);
  reduce.recs = red.empty() ? nullptr : arena.push(red);
}

template <typename T>
void GemmGroup<T>::launch(hipStream_t s, int j) const {
#define QCA_GG_L(V, AM, BNC, KG)                                                               \
  if (launches[V].n)                                                                          \
    hipLaunchKernelGGL((gemm_grouped_kernel<T, AM, BNC, KG>), dim3(launches[V].blocks),        \
                       dim3(256 * KG), 0, s, launches[V].recs, launches[V].n, j);
  QCA_GG_L(0, false, false, 1)
  QCA_GG_L(1, false, false, 2)
  QCA_GG_L(2, false, true, 1)
  QCA_GG_L(3, false, true, 2)
  QCA_GG_L(4, true, false, 1)
  QCA_GG_L(5, true, false, 2)
  QCA_GG_L(6, true, true, 1)
  QCA_GG_L(7, true, true, 2)
#undef QCA_GG_L
  QCA_HIP(hipGetLastError());
  if (reduce.n) {
    hipLaunchKernelGGL(splitk_reduce_grouped<T>, dim3(reduce.blocks), dim3(256), 0, s, reduce.recs,
                       reduce.n);
    QCA_HIP(hipGetLastError());
  }
}

template struct GemmGroup<double>;
template struct GemmGroup<float>;

void gemm_c128(const GemmDesc& d, const c128* A, const c128* B, c128* C, hipStream_t s,
               c128* ws, long ws_elems) {
  gemm_launch<double>(d, A, B, C, s, ws, ws_elems);
}
void gemm_c64(const GemmDesc& d, const c64* A, const c64* B, c64* C, hipStream_t s, c64* ws,
              long ws_elems) {
  gemm_launch<float>(d, A, B, C, s, ws, ws_elems);
}

}  // namespace qca
