// FP64 factorisations of the TDVP split (gfx950).
//
// jacobi_svd: block one-sided (Hestenes) Jacobi on the columns of a
// column-major matrix in one cooperative launch.  The columns are cut into
// blocks sized so that a pair of blocks fits in 128 KiB of LDS; in every round
// each workgroup loads one block pair, applies a full inner sweep of column
// rotations in LDS (round-robin pairs, one pair per wave64, wave-shuffle
// reductions for the 2x2 Gram entries), and writes it back; rounds follow the
// round-robin tournament over blocks and are separated by grid barriers.
// Jacobi keeps singular values accurate to working precision relative to the
// largest one, which the 1e-10 truncation tail test needs; columns below
// 1e-15 of the matrix norm are left alone (their contribution to any tail is
// < 1e-14).  The kernel also sorts the singular values and evaluates the
// truncation rank (first i with ||s[i:]|| < eps, capped; reference
// algorithms/tdvp.py:289-292).
//
// householder_qr: zgeqr2 + zung2r semantics (reference tensor_networks/mps.py:86
// calls numpy.linalg.qr -> LAPACK), including LAPACK's zlarfg sign convention
// beta = -sign(Re alpha)*|x|, so R has a real diagonal and Q matches LAPACK's
// complete-QR complement columns.
#include "kernels.hpp"

namespace qca {
namespace {

constexpr int JAC_THREADS = 1024;
constexpr int JAC_WAVES = JAC_THREADS / 64;
constexpr int JAC_MAXCOLS = 512;
constexpr int JAC_MAXR = 8;               // rows per lane (rows <= 512)
constexpr int JAC_LDS_BYTES = 128 * 1024;  // one block pair of columns

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Position t of round r in the round-robin tournament over n (even) players.
__device__ __forceinline__ int rr_player(int t, int r, int n) {
  return (t == 0) ? 0 : ((t - 1 + r) % (n - 1)) + 1;
}

// One Hestenes rotation of the column pair (xi, xj) held in LDS (rows long).
// Returns 1 when a rotation was applied.
__device__ int jacobi_pair(c128* xi, c128* xj, int rows, int lane, double tol, double negl) {
  double ir[JAC_MAXR], ii[JAC_MAXR], jr[JAC_MAXR], ji[JAC_MAXR];
  double a = 0, b = 0, gr = 0, gi = 0;
#pragma unroll
  for (int q = 0; q < JAC_MAXR; ++q) {
    const int r = lane + 64 * q;
    if (r < rows) {
      const c128 u = xi[r], v = xj[r];
      ir[q] = u.re; ii[q] = u.im; jr[q] = v.re; ji[q] = v.im;
    } else {
      ir[q] = ii[q] = jr[q] = ji[q] = 0.0;
    }
    a += ir[q] * ir[q] + ii[q] * ii[q];
    b += jr[q] * jr[q] + ji[q] * ji[q];
    gr += ir[q] * jr[q] + ii[q] * ji[q];  // g = x_i^H x_j
    gi += ir[q] * ji[q] - ii[q] * jr[q];
  }
  a = wsum(a); b = wsum(b); gr = wsum(gr); gi = wsum(gi);
  const double ag = sqrt(gr * gr + gi * gi);
  // converged pair, or both columns below the rounding level of the matrix
  if (ag <= tol * sqrt(a) * sqrt(b) || ag < 1e-280 || fmax(a, b) < negl) return 0;
  const double zeta = (b - a) / (2.0 * ag);
  double t;
  if (fabs(zeta) > 1e100) t = 0.5 / zeta;
  else t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
  const double er = gr / ag, ei = gi / ag;
  // x_i' = c x_i - s conj(e) x_j ;  x_j' = s e x_i + c x_j
#pragma unroll
  for (int q = 0; q < JAC_MAXR; ++q) {
    const int r = lane + 64 * q;
    if (r < rows) {
      const double ce_r = er * jr[q] + ei * ji[q], ce_i = er * ji[q] - ei * jr[q];
      const double e_r = er * ir[q] - ei * ii[q], e_i = er * ii[q] + ei * ir[q];
      xi[r] = c128{c * ir[q] - s * ce_r, c * ii[q] - s * ce_i};
      xj[r] = c128{s * e_r + c * jr[q], s * e_i + c * ji[q]};
    }
  }
  return 1;
}

// Grid barrier for a cooperative launch (all workgroups co-resident): every
// workgroup adds 1 to bar[0] and waits until the count reaches the target of
// this barrier instance (k-th barrier: k * gridDim.x).  Every thread fences at
// device scope on both sides, so global stores before the barrier are visible
// to, and stale cache lines are not read by, any thread after it.
__device__ void grid_barrier(unsigned int* bar, unsigned int& target) {
  target += gridDim.x;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(bar, 1u);
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __syncthreads();
  __threadfence();
}

__global__ void zero_words(unsigned int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Block one-sided Jacobi.  gridDim.x = nb/2 workgroups, each sweeping one pair
// of column blocks (bs columns each) in LDS per round; rounds follow the
// round-robin tournament over the nb blocks, separated by grid barriers.  The
// loop ends after a sweep without rotations (all workgroups read the same
// counter after a barrier) or after max_sweeps.  Workgroup 0 then sorts the
// column norms and evaluates the truncation rank.
__global__ __launch_bounds__(JAC_THREADS) void block_jacobi_kernel(
    c128* __restrict__ X, int rows, int cols, int bs, int nb, int max_sweeps, double tol,
    unsigned int* __restrict__ bar, int* __restrict__ rotc, double* __restrict__ fro,
    double* __restrict__ sigma, int* __restrict__ perm, SvdResult* __restrict__ res, int chi_max,
    double eps, int rank_cap) {
  extern __shared__ unsigned char jac_smem[];
  c128* lds = reinterpret_cast<c128*>(jac_smem);  // [2 bs][rows]
  __shared__ int local_rot;
  __shared__ double red[JAC_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wg = blockIdx.x, nwg = gridDim.x;
  unsigned int target = 0;

  // Frobenius norm^2 of X (for the negligible-column threshold)
  double part = 0;
  for (long e = (long)wg * JAC_THREADS + tid; e < (long)rows * cols; e += (long)nwg * JAC_THREADS)
    part += X[e].re * X[e].re + X[e].im * X[e].im;
  part = wsum(part);
  if (lane == 0) red[wave] = part;
  __syncthreads();
  if (tid == 0) {
    double sacc = 0;
    for (int w = 0; w < JAC_WAVES; ++w) sacc += red[w];
    atomicAdd(fro, sacc);
  }
  grid_barrier(bar, target);
  const double negl = 1e-30 * __hip_atomic_load(fro, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

  const int lc = 2 * bs;  // local columns
  int sweep = 0, converged = 0;
  for (sweep = 0; sweep < max_sweeps; ++sweep) {
    if (tid == 0) local_rot = 0;
    for (int round = 0; round < nb - 1; ++round) {
      const int bi = rr_player(wg, round, nb), bj = rr_player(nb - 1 - wg, round, nb);
      // stage the two blocks (zero columns beyond cols)
      for (int e = tid; e < lc * rows; e += JAC_THREADS) {
        const int c = e / rows, r = e - c * rows;
        const int gc = (c < bs ? bi * bs + c : bj * bs + (c - bs));
        lds[e] = (gc < cols) ? X[(long)gc * rows + r] : c128{0.0, 0.0};
      }
      __syncthreads();
      // one inner sweep over the 2 bs local columns
      for (int ir = 0; ir < lc - 1; ++ir) {
        for (int pr = wave; pr < bs; pr += JAC_WAVES) {
          const int ci = rr_player(pr, ir, lc), cj = rr_player(lc - 1 - pr, ir, lc);
          const int n = jacobi_pair(lds + (long)ci * rows, lds + (long)cj * rows, rows, lane, tol, negl);
          if (n && lane == 0) atomicAdd(&local_rot, 1);
        }
        __syncthreads();
      }
      for (int e = tid; e < lc * rows; e += JAC_THREADS) {
        const int c = e / rows, r = e - c * rows;
        const int gc = (c < bs ? bi * bs + c : bj * bs + (c - bs));
        if (gc < cols) X[(long)gc * rows + r] = lds[e];
      }
      grid_barrier(bar, target);
    }
    if (tid == 0) atomicAdd(&rotc[sweep & 1], local_rot);
    grid_barrier(bar, target);
    const int total = __hip_atomic_load(&rotc[sweep & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (wg == 0 && tid == 0) rotc[(sweep + 1) & 1] = 0;
    grid_barrier(bar, target);
    if (total == 0) {
      converged = 1;
      break;
    }
  }
  if (wg != 0) return;

  // ---- workgroup 0: column norms, descending sort, truncation rank
  double* skey = reinterpret_cast<double*>(jac_smem);            // reuse LDS
  int* sidx = reinterpret_cast<int*>(skey + JAC_MAXCOLS);
  int np2 = 1;
  while (np2 < cols) np2 <<= 1;
  for (int c = wave; c < np2; c += JAC_WAVES) {
    double a = 0;
    if (c < cols)
      for (int r = lane; r < rows; r += 64) {
        const c128 v = X[(long)c * rows + r];
        a += v.re * v.re + v.im * v.im;
      }
    a = wsum(a);
    if (lane == 0) {
      skey[c] = (c < cols) ? sqrt(a) : -1.0;
      sidx[c] = c;
    }
  }
  __syncthreads();
  for (int size = 2; size <= np2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid; e < np2; e += JAC_THREADS) {
        const int p = e ^ stride;
        if (p > e) {
          const bool desc = ((e & size) == 0);
          const double ke = skey[e], kp = skey[p];
          const bool swap = desc ? (ke < kp || (ke == kp && sidx[e] > sidx[p]))
                                 : (ke > kp || (ke == kp && sidx[e] < sidx[p]));
          if (swap) {
            skey[e] = kp; skey[p] = ke;
            const int ti = sidx[e]; sidx[e] = sidx[p]; sidx[p] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < cols; e += JAC_THREADS) {
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
    res->nsweeps = sweep + (converged ? 1 : 0);
    res->converged = converged;
    res->norm_kept = sqrt(kept);
    res->tail = sqrt(tail);
  }
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

}  // namespace

void launch_jacobi_svd(c128* X, int rows, int cols, double* sigma, int* perm, SvdResult* res,
                       int chi_max, double eps, int rank_cap, hipStream_t s, JacobiWork& w) {
  QCA_CHECK(rows >= 1 && rows <= 64 * JAC_MAXR, "SVD rows out of range (1..512)");
  QCA_CHECK(cols >= 1 && cols <= JAC_MAXCOLS, "SVD cols out of range (1..512)");
  // block size: a pair of blocks (2 bs columns of `rows` complex doubles) fits
  // in 128 KiB of LDS, and is no wider than the matrix needs
  int bs = 1;
  while (bs * 2 <= 64 && 2L * (bs * 2) * rows * (long)sizeof(c128) <= JAC_LDS_BYTES &&
         2 * bs < cols)
    bs *= 2;
  int nb = ceil_div(cols, bs);
  if (nb < 2) nb = 2;
  if (nb & 1) ++nb;
  const int nwg = nb / 2;
  const size_t shm = std::max<size_t>(2UL * bs * rows * sizeof(c128),
                                      JAC_MAXCOLS * (sizeof(double) + sizeof(int)));
  static size_t attr_bytes = 0;  // dynamic LDS opt-in (static + dynamic <= 160 KiB)
  if (shm > attr_bytes) {
    QCA_HIP(hipFuncSetAttribute((const void*)block_jacobi_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    attr_bytes = shm;
  }
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned int*>(w.sync), 16);
  int max_sweeps = 60;
  double tol = 1.1102230246251565e-16 * sqrt((double)rows);
  unsigned int* bar = reinterpret_cast<unsigned int*>(w.sync);
  int* rotc = reinterpret_cast<int*>(w.sync) + 4;
  double* fro = reinterpret_cast<double*>(reinterpret_cast<char*>(w.sync) + 32);
  void* args[] = {&X, &rows, &cols, &bs, &nb, &max_sweeps, &tol, &bar, &rotc, &fro,
                  &sigma, &perm, &res, &chi_max, &eps, &rank_cap};
  if (nwg == 1) {
    // one workgroup: the grid barrier reduces to a workgroup barrier
    hipLaunchKernelGGL(block_jacobi_kernel, dim3(1), dim3(JAC_THREADS), shm, s, X, rows, cols, bs, nb,
                       max_sweeps, tol, bar, rotc, fro, sigma, perm, res, chi_max, eps, rank_cap);
    QCA_HIP(hipGetLastError());
  } else {
    // several workgroups wait on each other: cooperative launch guarantees
    // they are co-resident (or fails instead of hanging)
    QCA_HIP(hipLaunchCooperativeKernel((const void*)block_jacobi_kernel, dim3(nwg),
                                       dim3(JAC_THREADS), args, (unsigned int)shm, s));
  }
}

void launch_householder_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols,
                           hipStream_t s) {
  QCA_CHECK(rows >= 1 && cols >= 1, "empty QR");
  hipLaunchKernelGGL(householder_qr_kernel, dim3(1), dim3(QR_THREADS), 0, s, M, rows, cols, tau, Q,
                     qcols);
  QCA_HIP(hipGetLastError());
}

}  // namespace qca
