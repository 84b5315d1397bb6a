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
// current one is multiplied.  The K range is walked as segments, which also
// lets the MPO mixing of the effective Hamiltonian be gathered into the
// operand loads (GemmDesc's gather form) instead of a separate pass.
// Launches with too few output tiles to fill 256 CUs split K and reduce the
// partial tiles in a second, deterministic pass.
#include <cstdlib>

#include "grouped.hpp"

namespace qca {
namespace {

constexpr int BM = 64, BN = 64, BK = 16;
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

// One LDS plane (real or imaginary part of one operand tile).  An operand
// whose tile rows (m or n) are contiguous in memory is staged as [k][mn], one
// whose k is contiguous as [mn][k].  The row strides keep both the stores of
// the coalesced loads and the MFMA operand reads (16 consecutive mn at each of
// 4 k per wave) free of bank conflicts: LDS serves a wave64 b32 access 32
// lanes per cycle and a b64 access 16 lanes per cycle over 32 banks.
template <typename T, bool MNCONT>
struct Plane {
  static constexpr int LS = MNCONT ? (sizeof(T) == 4 ? BM + 16 : BM) : (sizeof(T) == 4 ? BK + 2 : BK + 1);
  static constexpr int SIZE = MNCONT ? BK * LS : BM * LS;
  static __device__ __forceinline__ int at(int k, int mn) { return MNCONT ? k * LS + mn : mn * LS + k; }
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}


// One output tile (bx: N tile, by: M tile, bz: batch * splits + split) of the
// GEMM d; the kernels below supply the block index (direct launch) or a
// record's virtual one (grouped launch over many instances).
//
// The K range of the batch is a list of segments (see GemmDesc) cut into
// K-tiles of BK (a segment's last tile is masked); split `split` takes tiles
// [split * tiles_per_split, +tiles_per_split).  KG K-groups of 256 threads
// share the workgroup: group g takes every KG-th tile into its own LDS
// stages, and the groups' partial tiles are summed through LDS at the end.
// With KG = 2 a CU runs two waves per SIMD, so one group's staging overlaps
// the other's MFMAs.  Each group double-buffers its LDS: the tile loaded into
// registers during iteration i-1 is stored into stage i%2 while stage
// (i-1)%2 may still be read, so one barrier per K-tile suffices, and the
// global loads of tile i+1 are in flight during the MFMAs of tile i.
template <typename T, bool A_MCONT, bool B_NCONT, int KG, bool FUSED>
__device__ __forceinline__ void gemm_body(const GemmDesc& d, const cplx<T>* __restrict__ A,
                                          const cplx<T>* __restrict__ B, cplx<T>* __restrict__ C,
                                          cplx<T>* __restrict__ ws, int splits, int tiles_per_split,
                                          unsigned bx, unsigned by, unsigned bz) {
  constexpr bool fused = FUSED;
  using acc_t = typename Mfma<T>::acc_t;
  using PA = Plane<T, A_MCONT>;
  using PB = Plane<T, B_NCONT>;
  constexpr int STAGE = 2 * PA::SIZE + 2 * PB::SIZE;  // [A re, A im, B re, B im]
  __shared__ T smem[KG * 2 * STAGE];
  static_assert(KG == 1 || KG * 2 * STAGE >= 4 * 32 * 64, "partial-tile exchange fits the operand LDS");
  // the skip word is read here and tested once the first tile's loads are in
  // flight: the two memory latencies overlap (a skipped launch only loads)
  const int skipped = d.skip ? *d.skip : 0;

  const int tid = threadIdx.x & 255;  // thread within its K-group
  // the K-group is wave-uniform: keeping it (and so the tile schedule, the
  // segment lookups and the tile base addresses) in scalar registers
  const int grp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8);
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 32;
  const int wn = (wave & 1) * 32;
  const int m0 = by * BM;
  const int n0 = bx * BN;
  // fused: this workgroup takes every split of its output tile in turn (bz is
  // the batch) and sums them in split order itself; otherwise bz = batch *
  // splits + split and the partial tile goes to the split-K workspace
  const int zb = fused ? (int)bz : (int)bz / splits;
  const int split = fused ? 0 : (int)bz - zb * splits;
  const int b1 = zb / d.NB2;
  const int b2 = zb - b1 * d.NB2;

  const cplx<T>* Ab = A + b1 * d.a_b1 + b2 * d.a_b2 + (long)m0 * d.a_m;
  const cplx<T>* Bb = B + b1 * d.b_b1 + b2 * d.b_b2 + (long)n0 * d.b_n;

  const bool gather = d.seg_rows > 0;
  int e0 = 0, nseg;
  if (gather && d.seg_starts) {
    // from the arguments (a select chain, no indexed copy of the descriptor)
    int lo = d.seg_start[0], hi = d.seg_start[1];
#pragma unroll
    for (int g = 1; g < SEG_STARTS_MAX; ++g)
      if (b2 == g) {
        lo = d.seg_start[g];
        hi = d.seg_start[g + 1];
      }
    e0 = lo;
    nseg = hi - lo;
  } else if (gather) {
    e0 = d.seg_rowptr[b2 * d.seg_rows];
    nseg = d.seg_rowptr[(b2 + 1) * d.seg_rows] - e0;
  } else {
    nseg = d.K / d.K2;
  }
  const int tps = (d.K2 + BK - 1) / BK;
  const int ntiles = nseg * tps;
  const int ch_begin = fused ? 0 : split, ch_end = fused ? splits : split + 1;
  // K-group iterations of split (chunk) ch, and the tile of iteration li (-1: none)
  auto chunk_iters = [&](int ch) {
    const int tb = ch * tiles_per_split, te = min(ntiles, tb + tiles_per_split);
    return te > tb ? (te - tb + KG - 1) / KG : 0;
  };
  auto tile_at = [&](int ch, int li) {
    const int tb = ch * tiles_per_split, te = min(ntiles, tb + tiles_per_split);
    const int t = tb + li * KG + grp;
    return t < te ? t : -1;
  };

  // element offsets of this thread's loads within a tile (strides are
  // non-negative and a tile spans less than 2^31 elements: checked on the host)
  unsigned offa[LOADS], offb[LOADS];
  int kla[LOADS], klb[LOADS];
  bool oka[LOADS], okb[LOADS];
#pragma unroll
  for (int r = 0; r < LOADS; ++r) {
    const int e = tid + 256 * r;
    const int ml = A_MCONT ? (e % BM) : (e / BK);
    const int ka = A_MCONT ? (e / BM) : (e % BK);
    const int nl = B_NCONT ? (e % BN) : (e / BK);
    const int kb = B_NCONT ? (e / BN) : (e % BK);
    oka[r] = m0 + ml < d.M;
    okb[r] = n0 + nl < d.N;
    kla[r] = ka;
    klb[r] = kb;
    offa[r] = (unsigned)(ml * d.a_m + ka * d.a_k2);
    offb[r] = (unsigned)(nl * d.b_n + kb * d.b_k2);
  }

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

  // operand elements of one K-tile (and the tile's gather weight), loaded into
  // registers one iteration ahead of their LDS store
  cplx<T> pa[LOADS], pb[LOADS];
  cplx<T> wgt{T(1), T(0)};
  // the gather form's segment record, looked up again only when a tile starts
  // a new segment (one lookup per K2/BK tiles instead of one per tile)
  int cached_seg = -1;
  long cached_soa = 0, cached_sob = 0;
  cplx<T> cached_wgt{T(1), T(0)};
  auto load_tile = [&](int t) {
    if (t < 0) {  // uniform per K-group: no memory access
#pragma unroll
      for (int r = 0; r < LOADS; ++r) pa[r] = pb[r] = cplx<T>{T(0), T(0)};
      return;
    }
    const int seg = t / tps;
    const int k0 = (t - seg * tps) * BK;
    long soa, sob;
    if (gather) {
      if (seg != cached_seg) {
        const int e = e0 + seg;
        cached_seg = seg;
        cached_soa = (long)d.seg_col[e] * d.a_k1;
        cached_sob = (long)d.seg_outv[e] * d.b_k1;
        cached_wgt = reinterpret_cast<const cplx<T>*>(d.seg_val)[e];
      }
      soa = cached_soa;
      sob = cached_sob;
      wgt = cached_wgt;
    } else {
      soa = (long)seg * d.a_k1;
      sob = (long)seg * d.b_k1;
    }
    const cplx<T>* ta = Ab + soa + (long)k0 * d.a_k2;
    const cplx<T>* tb = Bb + sob + (long)k0 * d.b_k2;
    const int klim = d.K2 - k0;
    // an element outside the operand reads the tile's first element (always
    // valid) and is replaced by zero
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const bool ok = oka[r] && kla[r] < klim;
      const cplx<T> v = ta[ok ? offa[r] : 0u];
      pa[r] = ok ? v : cplx<T>{T(0), T(0)};
    }
#pragma unroll
    for (int r = 0; r < LOADS; ++r) {
      const bool ok = okb[r] && klb[r] < klim;
      const cplx<T> v = tb[ok ? offb[r] : 0u];
      pb[r] = ok ? v : cplx<T>{T(0), T(0)};
    }
  };

  auto stage_ptr = [&](int stage) { return smem + (grp * 2 + stage) * STAGE; };

  acc_t tot_re[2][2], tot_im[2][2];   // fused: sum of the finished splits
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      tot_re[i][j] = acc_t{0, 0, 0, 0};
      tot_im[i][j] = acc_t{0, 0, 0, 0};
    }
  T* const xs = smem + (wave * 32) * 64 + lane;  // partial-tile exchange: [wave][value][lane]

  load_tile(ch_begin < ch_end ? tile_at(ch_begin, 0) : -1);
  if (skipped) return;
  int it = 0;  // iterations so far: the LDS stage parity
  for (int ch = ch_begin; ch < ch_end; ++ch) {
    const int nit = chunk_iters(ch);
    for (int li = 0; li < nit; ++li, ++it) {
      const int t = tile_at(ch, li);
      T* st = stage_ptr(it & 1);
      T* a_re = st;
      T* a_im = st + PA::SIZE;
      T* b_re = st + 2 * PA::SIZE;
      T* b_im = st + 2 * PA::SIZE + PB::SIZE;
      // ---- registers -> LDS (re/im planes), then start this group's next tile
      const cplx<T> w = wgt;
#pragma unroll
      for (int r = 0; r < LOADS; ++r) {
        const int e = tid + 256 * r;
        const int ml = A_MCONT ? (e % BM) : (e / BK);
        const int kl = A_MCONT ? (e / BM) : (e % BK);
        const cplx<T> v = gather ? cmul(pa[r], w) : pa[r];
        a_re[PA::at(kl, ml)] = v.re;
        a_im[PA::at(kl, ml)] = sa * v.im;
      }
#pragma unroll
      for (int r = 0; r < LOADS; ++r) {
        const int e = tid + 256 * r;
        const int nl = B_NCONT ? (e % BN) : (e / BK);
        const int kl = B_NCONT ? (e / BN) : (e % BK);
        b_re[PB::at(kl, nl)] = pb[r].re;
        b_im[PB::at(kl, nl)] = sb * pb[r].im;
      }
      // the next tile of this K-group: this split's next iteration, else the
      // first iteration of the next split that has tiles
      int tn = -1;
      if (li + 1 < nit) {
        tn = tile_at(ch, li + 1);
      } else {
        for (int c2 = ch + 1; c2 < ch_end; ++c2)
          if (chunk_iters(c2) > 0) {
            tn = tile_at(c2, 0);
            break;
          }
      }
      load_tile(tn);
      __syncthreads();

      // a group without a tile in the last round skips its MFMAs (the barrier
      // stays outside)
      if (t >= 0) {
        const int c = lane & 15;
        const int q = lane >> 4;
        T ar[4][2], ai[4][2], br[4][2], bi[4][2];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            ar[kk][i] = a_re[PA::at(4 * kk + q, wm + 16 * i + c)];
            ai[kk][i] = a_im[PA::at(4 * kk + q, wm + 16 * i + c)];
            br[kk][i] = b_re[PB::at(4 * kk + q, wn + 16 * i + c)];
            bi[kk][i] = b_im[PB::at(4 * kk + q, wn + 16 * i + c)];
          }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              acc_re[i][j] = Mfma<T>::mac(ar[kk][i], br[kk][j], acc_re[i][j]);
              acc_im[i][j] = Mfma<T>::mac(ar[kk][i], bi[kk][j], acc_im[i][j]);
            }
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              acc_re[i][j] = Mfma<T>::mac(-ai[kk][i], bi[kk][j], acc_re[i][j]);
              acc_im[i][j] = Mfma<T>::mac(ai[kk][i], br[kk][j], acc_im[i][j]);
            }
        }
      }
    }
    // end of the split: with KG = 2 group 1 hands its partial tile to group 0
    // through the operand LDS (once every wave is done reading it; the last
    // barrier keeps the next tile's stores off the exchange area until group 0
    // has read it)
    if constexpr (KG == 2) {
      __syncthreads();
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
      if (grp == 0) {
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
      __syncthreads();
    }
    if (fused) {
      // the split's partial tile joins the sum in split order, as splitk_reduce adds it
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            tot_re[i][j][r] += acc_re[i][j][r];
            tot_im[i][j][r] += acc_im[i][j][r];
          }
          acc_re[i][j] = acc_t{0, 0, 0, 0};
          acc_im[i][j] = acc_t{0, 0, 0, 0};
        }
    }
  }

  const T al_re = T(d.alpha_re), al_im = T(d.alpha_im);
  auto store_out = [&](int m, int n, T xr, T xi) {
    cplx<T>* cp = C + b1 * d.c_b1 + b2 * d.c_b2 + m * d.c_m + n * d.c_n;
    cplx<T> out{al_re * xr - al_im * xi, al_re * xi + al_im * xr};
    if (d.accumulate) {
      out.re += cp->re;
      out.im += cp->im;
    }
    *cp = out;
  };
  // ---- epilogue: column lane%16, rows per Mfma<T>::row of each 16x16 tile
  if (grp == 0) {
    const int col = lane & 15;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + wm + 16 * i + Mfma<T>::row(lane, r);
          const int n = n0 + wn + 16 * j + col;
          if (m >= d.M || n >= d.N) continue;
          if (fused) {
            store_out(m, n, tot_re[i][j][r], tot_im[i][j][r]);
          } else {
            const T xr = acc_re[i][j][r], xi = acc_im[i][j][r];
            if (splits > 1) ws[((long)bz * d.M + m) * d.N + n] = cplx<T>{xr, xi};
            else store_out(m, n, xr, xi);
          }
        }
  }
}

template <typename T, bool A_MCONT, bool B_NCONT, int KG>
__global__ __launch_bounds__(256 * KG) void gemm_kernel(GemmDesc d, const cplx<T>* __restrict__ A,
                                                        const cplx<T>* __restrict__ B,
                                                        cplx<T>* __restrict__ C,
                                                        cplx<T>* __restrict__ ws, int splits,
                                                        int tiles_per_split) {
  gemm_body<T, A_MCONT, B_NCONT, KG, false>(d, A, B, C, ws, splits, tiles_per_split, blockIdx.x, blockIdx.y,
                                            blockIdx.z);
}

// Sum the split-K partial tiles in split order (deterministic) and apply
// alpha/beta while scattering into C's strides; block `blk` of `nblk`, 256
// threads each.
template <typename T>
__device__ __forceinline__ void splitk_reduce_body(const GemmDesc& d, const cplx<T>* __restrict__ ws,
                                                   cplx<T>* __restrict__ C, int splits, int batch,
                                                   unsigned blk, unsigned nblk) {
  const int skipped = d.skip ? *d.skip : 0;  // tested after the first partial loads
  const long mn = (long)d.M * d.N;
  const long total = mn * batch;
  for (long idx = blk * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)nblk * blockDim.x) {
    const int zb = (int)(idx / mn);
    const long r = idx - zb * mn;
    const int m = (int)(r / d.N), n = (int)(r - (long)m * d.N);
    T xr = 0, xi = 0;
    // 16 partial loads in flight per block of splits (one memory latency for
    // the 15 splits of the two-site chain): slots past the last split load it
    // again and add nothing; the sum stays in split order
    const cplx<T>* wp = ws + ((long)zb * splits * d.M + m) * d.N + n;
    const long sstride = (long)d.M * d.N;
    for (int s0 = 0; s0 < splits; s0 += 16) {
      cplx<T> v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = wp[min(s0 + i, splits - 1) * sstride];
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (s0 + i < splits) {
          xr += v[i].re;
          xi += v[i].im;
        }
    }
    if (skipped) return;
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
// Output tiles of a grouped GEMM launch from which split GEMMs run fused (one
// workgroup per output tile over all splits): four per CU.
constexpr long GEMM_GROUP_FUSE_TILES = 1024;

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
  if (d.seg_rows > 0) {
    QCA_CHECK(d.NB1 == 1 && d.seg_rowptr && d.seg_col && d.seg_outv && d.seg_val && !d.conj_a,
              "gathered GEMM needs one batch level, the sparse operand and no conjugation of A");
    QCA_CHECK(d.seg_max >= 0 && d.K == d.seg_max * d.K2, "gathered GEMM: K must be seg_max * K2");
    QCA_CHECK(!d.seg_starts || d.NB2 <= SEG_STARTS_MAX,
              "gathered GEMM: segment starts in the descriptor cover at most SEG_STARTS_MAX batches");
  }
  // per-thread element offsets within a tile are 32-bit and non-negative
  QCA_CHECK(d.a_m >= 0 && d.a_k2 >= 0 && d.b_n >= 0 && d.b_k2 >= 0 && d.a_k1 >= 0 && d.b_k1 >= 0,
            "GEMM strides must be non-negative");
  QCA_CHECK((BM - 1) * d.a_m + (BK - 1) * d.a_k2 < (1L << 31) && (BN - 1) * d.b_n + (BK - 1) * d.b_k2 < (1L << 31),
            "GEMM tile spans more than 2^31 elements");
  GemmPlan p;
  p.batch = d.NB1 * d.NB2;
  if (d.M == 0 || d.N == 0) {
    p.empty = true;
    return p;
  }
  p.tm = ceil_div(d.M, BM);
  p.tn = ceil_div(d.N, BN);
  const long tiles = (long)p.tm * p.tn * p.batch;
  // K-tiles of the longest batch: segments of K2, each cut into BK tiles
  const long ktiles = (long)(d.K / d.K2) * ceil_div(d.K2, BK);
  QCA_CHECK(ktiles < (1L << 30), "GEMM K too long");
  // split K when the output tiles alone cannot occupy the chip
  int splits = 1;
  if (tiles < 192 && ktiles >= 4) {
    splits = (int)std::min<long>((256 + tiles - 1) / tiles, ktiles / 2);
    while (splits > 1 && (long)splits * p.batch * d.M * d.N > ws_elems) --splits;
    if (splits < 1) splits = 1;
  }
  int kps = (int)ktiles;  // K-tiles per split
  if (splits > 1) {
    kps = ceil_div(ktiles, splits);
    splits = ceil_div(ktiles, kps);
  }
  QCA_CHECK((long)tiles * splits < (1L << 31), "GEMM grid too large");
  QCA_CHECK(p.batch * splits <= 65535, "GEMM batch*splits exceeds grid z limit");
  p.splits = splits;
  p.kps = kps;
  p.am = (d.a_m == 1);
  p.bn = (d.b_n == 1);
  // two K-groups per workgroup when each has at least two K-tiles to take
  p.kg2 = kps >= 4;
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
// The descriptor's pointers in the global address space (see gbl).
__device__ __forceinline__ GemmDesc global_desc(GemmDesc d) {
  d.skip = gbl(d.skip);
  d.seg_rowptr = gbl(d.seg_rowptr);
  d.seg_col = gbl(d.seg_col);
  d.seg_outv = gbl(d.seg_outv);
  d.seg_val = gbl(d.seg_val);
  return d;
}

template <typename T, bool A_MCONT, bool B_NCONT, int KG, bool FUSED>
__global__ __launch_bounds__(256 * KG) void gemm_grouped_kernel(const GRec<GemmArgs<T>>* __restrict__ recs,
                                                                int n, int j) {
  unsigned bx, by, bz;
  const GRec<GemmArgs<T>>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const GemmDesc d = global_desc(r.a.d);
  gemm_body<T, A_MCONT, B_NCONT, KG, FUSED>(d, gbl(r.a.A + (long)j * r.a.a_jstride),
                                            gbl(r.a.B + (long)j * r.a.b_jstride), gbl(r.a.C),
                                            gbl(r.a.ws), r.a.splits, r.a.kps, bx, by, bz);
}

template <typename T>
__global__ void splitk_reduce_grouped(const GRec<GemmArgs<T>>* __restrict__ recs, int n) {
  unsigned bx, by, bz;
  const GRec<GemmArgs<T>>& r = grec_find(recs, n, blockIdx.x, bx, by, bz);
  const GemmDesc d = global_desc(r.a.d);
  splitk_reduce_body<T>(d, gbl(r.a.ws), gbl(r.a.C), r.a.splits, d.NB1 * d.NB2, bx, r.grid.x);
}

}  // namespace

template <typename T>
void GemmGroup<T>::prepare(const std::vector<GemmJob<T>>& jobs, RecordArena& arena) {
  std::vector<GRec<GemmArgs<T>>> var[16], red;
  // A GEMM's splits (and K-groups) are fixed by its own plan, so that it sums
  // in the same order whether launched alone or grouped.  When the group's
  // output tiles alone fill the GPU, each split GEMM runs one workgroup per
  // output tile that walks all its splits and adds their partial tiles in
  // split order (bit-identical to the split-K launch plus splitk_reduce),
  // without the partial tiles' round trip through memory and the reduce.
  long out_tiles = 0;
  for (const auto& j : jobs) {
    const GemmPlan p = gemm_plan(j.d, j.ws_elems);
    if (!p.empty) out_tiles += (long)p.tm * p.tn * p.batch;
  }
  // QCA_GEMM_GROUP_FUSE_TILES overrides the threshold (0: always), so that the
  // fused form can be tested against single launches on small groups
  long fuse_tiles = GEMM_GROUP_FUSE_TILES;
  if (const char* e = std::getenv("QCA_GEMM_GROUP_FUSE_TILES")) fuse_tiles = std::atol(e);
  const bool fuse = out_tiles >= fuse_tiles;
  for (const auto& j : jobs) {
    const GemmPlan p = gemm_plan(j.d, j.ws_elems);
    if (p.empty) continue;
    GRec<GemmArgs<T>> r{};
    r.a.d = j.d;
    r.a.A = j.A;
    r.a.B = j.B;
    r.a.C = j.C;
    r.a.ws = j.ws;
    r.a.a_jstride = j.a_jstride;
    r.a.b_jstride = j.b_jstride;
    r.a.splits = p.splits;
    r.a.kps = p.kps;
    r.a.fused = (fuse && p.splits > 1) ? 1 : 0;
    r.grid.x = p.tn;
    r.grid.y = p.tm;
    r.grid.z = r.a.fused ? p.batch : p.batch * p.splits;
    var[(p.am ? 8 : 0) + (p.bn ? 4 : 0) + (p.kg2 ? 2 : 0) + r.a.fused].push_back(r);
    if (p.splits > 1 && !r.a.fused) {
      r.grid = VGrid{};
      r.grid.x = p.reduce_blocks;
      red.push_back(r);
    }
  }
  for (int v = 0; v < 16; ++v) {
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
#define QCA_GG_L(V, AM, BNC, KG, F)                                                            \
  if (launches[V].n)                                                                          \
    hipLaunchKernelGGL((gemm_grouped_kernel<T, AM, BNC, KG, F>), dim3(launches[V].blocks),     \
                       dim3(256 * KG), 0, s, launches[V].recs, launches[V].n, j);
  QCA_GG_L(0, false, false, 1, false)
  QCA_GG_L(1, false, false, 1, true)
  QCA_GG_L(2, false, false, 2, false)
  QCA_GG_L(3, false, false, 2, true)
  QCA_GG_L(4, false, true, 1, false)
  QCA_GG_L(5, false, true, 1, true)
  QCA_GG_L(6, false, true, 2, false)
  QCA_GG_L(7, false, true, 2, true)
  QCA_GG_L(8, true, false, 1, false)
  QCA_GG_L(9, true, false, 1, true)
  QCA_GG_L(10, true, false, 2, false)
  QCA_GG_L(11, true, false, 2, true)
  QCA_GG_L(12, true, true, 1, false)
  QCA_GG_L(13, true, true, 1, true)
  QCA_GG_L(14, true, true, 2, false)
  QCA_GG_L(15, true, true, 2, true)
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
