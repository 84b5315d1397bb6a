// FP64 factorisations of the TDVP split (gfx950).
//
// jacobi_svd: one-sided (Hestenes) Jacobi on the columns of a column-major
// matrix, one workgroup of 16 wave64s.  Column pairs follow the round-robin
// tournament so that every round updates n/2 disjoint pairs in parallel, one
// pair per wave at a time, with a wave-shuffle reduction for the 2x2 Gram
// entries.  Jacobi keeps small singular values accurate to working precision
// relative to the largest one, which the 1e-10 truncation tail test needs.
// The kernel also sorts the singular values and evaluates the truncation rank
// (first i with ||s[i:]|| < eps, capped; reference algorithms/tdvp.py:289-292).
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
constexpr int JAC_MAXR = 8;  // rows per lane (rows <= 512)

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(JAC_THREADS) void jacobi_svd_kernel(c128* __restrict__ X, int rows,
                                                                 int cols, double* __restrict__ sigma,
                                                                 int* __restrict__ perm,
                                                                 SvdResult* __restrict__ res,
                                                                 int chi_max, double eps,
                                                                 int rank_cap) {
  __shared__ double nrm[JAC_MAXCOLS];
  __shared__ double skey[JAC_MAXCOLS];
  __shared__ int sidx[JAC_MAXCOLS];
  __shared__ int rot_count;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nr = (rows + 63) / 64;  // rows handled per lane
  const double tol = 1.1102230246251565e-16 * sqrt((double)rows);

  // exact squared column norms
  for (int c = wave; c < cols; c += JAC_WAVES) {
    const c128* x = X + (long)c * rows;
    double a = 0;
    for (int r = lane; r < rows; r += 64) a += x[r].re * x[r].re + x[r].im * x[r].im;
    a = wsum(a);
    if (lane == 0) nrm[c] = a;
  }
  __syncthreads();

  const int n2 = cols + (cols & 1);
  int sweep = 0, converged = 0;
  for (sweep = 0; sweep < 40; ++sweep) {
    if (tid == 0) rot_count = 0;
    __syncthreads();
    for (int round = 0; round < n2 - 1; ++round) {
      for (int k = wave; k < n2 / 2; k += JAC_WAVES) {
        const int t1 = k, t2 = n2 - 1 - k;
        const int i = (t1 == 0) ? 0 : ((t1 - 1 + round) % (n2 - 1)) + 1;
        const int j = ((t2 - 1 + round) % (n2 - 1)) + 1;
        if (i >= cols || j >= cols) continue;
        c128* xi = X + (long)i * rows;
        c128* xj = X + (long)j * rows;
        double ir[JAC_MAXR], ii[JAC_MAXR], jr[JAC_MAXR], ji[JAC_MAXR];
        double gr = 0, gi = 0;
#pragma unroll
        for (int q = 0; q < JAC_MAXR; ++q) {
          const int r = lane + 64 * q;
          if (q < nr && r < rows) {
            ir[q] = xi[r].re; ii[q] = xi[r].im;
            jr[q] = xj[r].re; ji[q] = xj[r].im;
          } else {
            ir[q] = ii[q] = jr[q] = ji[q] = 0.0;
          }
          // g = x_i^H x_j
          gr += ir[q] * jr[q] + ii[q] * ji[q];
          gi += ir[q] * ji[q] - ii[q] * jr[q];
        }
        gr = wsum(gr);
        gi = wsum(gi);
        const double a = nrm[i], b = nrm[j];
        const double ag = sqrt(gr * gr + gi * gi);
        if (ag <= tol * sqrt(a) * sqrt(b) || ag < 1e-280) continue;
        const double zeta = (b - a) / (2.0 * ag);
        double t;
        if (fabs(zeta) > 1e100) t = 0.5 / zeta;
        else t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        const double er = gr / ag, ei = gi / ag;
        // x_i' = c x_i - s conj(e) x_j ;  x_j' = s e x_i + c x_j
        double na = 0, nb = 0;
#pragma unroll
        for (int q = 0; q < JAC_MAXR; ++q) {
          const double ce_r = er * jr[q] + ei * ji[q];   // conj(e) * x_j
          const double ce_i = er * ji[q] - ei * jr[q];
          const double e_r = er * ir[q] - ei * ii[q];    // e * x_i
          const double e_i = er * ii[q] + ei * ir[q];
          const double nir = c * ir[q] - s * ce_r, nii = c * ii[q] - s * ce_i;
          const double njr = s * e_r + c * jr[q], nji = s * e_i + c * ji[q];
          ir[q] = nir; ii[q] = nii; jr[q] = njr; ji[q] = nji;
          na += nir * nir + nii * nii;
          nb += njr * njr + nji * nji;
        }
#pragma unroll
        for (int q = 0; q < JAC_MAXR; ++q) {
          const int r = lane + 64 * q;
          if (q < nr && r < rows) {
            xi[r] = c128{ir[q], ii[q]};
            xj[r] = c128{jr[q], ji[q]};
          }
        }
        na = wsum(na);
        nb = wsum(nb);
        if (lane == 0) {
          nrm[i] = na;
          nrm[j] = nb;
          atomicAdd(&rot_count, 1);
        }
      }
      __syncthreads();
    }
    const int rc = rot_count;
    __syncthreads();
    if (rc == 0) {
      converged = 1;
      break;
    }
  }

  // sort singular values descending (bitonic over the next power of two)
  int np2 = 1;
  while (np2 < cols) np2 <<= 1;
  for (int e = tid; e < np2; e += JAC_THREADS) {
    skey[e] = (e < cols) ? sqrt(nrm[e]) : -1.0;
    sidx[e] = e;
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
    // tail norms over the min(rows, cols) values a reduced SVD returns
    const int ns = rank_cap;
    const int cap = min(chi_max, ns);
    int k = cap;
    double suffix = 0;
    // suffix[i] = sum_{j>=i} s_j^2, scanned from the end; the first i (< cap)
    // whose tail norm is below eps is the rank.
    double tails_cap = 0;
    for (int e = ns - 1; e >= 0; --e) {
      suffix += skey[e] * skey[e];
      if (e < cap && sqrt(suffix) < eps) k = e;  // keeps the smallest such e
      if (e == cap) tails_cap = suffix;
    }
    double kept = 0;
    for (int e = 0; e < k; ++e) kept += skey[e] * skey[e];
    double tail = 0;
    for (int e = k; e < ns; ++e) tail += skey[e] * skey[e];
    (void)tails_cap;
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
                       int chi_max, double eps, int rank_cap, hipStream_t s) {
  QCA_CHECK(rows >= 1 && rows <= 64 * JAC_MAXR, "SVD rows out of range (1..512)");
  QCA_CHECK(cols >= 1 && cols <= JAC_MAXCOLS, "SVD cols out of range (1..512)");
  hipLaunchKernelGGL(jacobi_svd_kernel, dim3(1), dim3(JAC_THREADS), 0, s, X, rows, cols, sigma, perm,
                     res, chi_max, eps, rank_cap);
  QCA_HIP(hipGetLastError());
}

void launch_householder_qr(c128* M, int rows, int cols, c128* tau, c128* Q, int qcols,
                           hipStream_t s) {
  QCA_CHECK(rows >= 1 && cols >= 1, "empty QR");
  hipLaunchKernelGGL(householder_qr_kernel, dim3(1), dim3(QR_THREADS), 0, s, M, rows, cols, tau, Q,
                     qcols);
  QCA_HIP(hipGetLastError());
}

}  // namespace qca
