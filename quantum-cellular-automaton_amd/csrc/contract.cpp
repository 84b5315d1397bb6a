// Contraction launch sequences of the TDVP path; see contract.hpp.
#include "contract.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace qca {

GemmDesc gemm_desc(int M, int N, int K) {
  GemmDesc d;
  std::memset(&d, 0, sizeof(d));
  d.M = M;
  d.N = N;
  d.K = K;
  d.K2 = K > 0 ? K : 1;
  d.NB1 = d.NB2 = 1;
  d.alpha_re = 1.0;
  return d;
}

template <>
void gemm_t<double>(const GemmDesc& d, const c128* A, const c128* B, c128* C, hipStream_t s,
                    const ContractWs<double>& ws) {
  gemm_c128(d, A, B, C, s, ws.splitk, ws.splitk_elems);
}
template <>
void gemm_t<float>(const GemmDesc& d, const c64* A, const c64* B, c64* C, hipStream_t s,
                   const ContractWs<float>& ws) {
  gemm_c64(d, A, B, C, s, ws.splitk, ws.splitk_elems);
}

template <typename T>
ContractPlan<T> heff_plan(int kind, const cplx<T>* x, cplx<T>* y, const cplx<T>* L, const cplx<T>* R,
                          const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                          const ContractWs<T>& ws, const int* skip) {
  ContractPlan<T> p{};
  const long lr = (long)chil * chir;
  if (kind == 2) {
    const int D = Din;
    // T[(w,x)][c] = sum_a L[a][(w,x)] C[a][c]
    GemmDesc g = gemm_desc(D * chil, chir, chil);
    g.a_m = 1; g.a_k2 = (long)D * chil;
    g.b_k2 = chir; g.b_n = 1;
    g.c_m = chir; g.c_n = 1;
    g.skip = skip;
    p.g1 = GemmJob<T>{g, L, x, ws.T1, ws.splitk, ws.splitk_elems};
    // y[x][y] = sum_{(w,c)} T[w][x][c] R[c][w][y]
    GemmDesc h = gemm_desc(chil, chir, D * chir);
    h.K2 = chir;
    h.a_m = chir; h.a_k1 = lr; h.a_k2 = 1;
    h.b_k1 = chir; h.b_k2 = (long)D * chir; h.b_n = 1;
    h.c_m = chir; h.c_n = 1;
    h.skip = skip;
    p.g3 = GemmJob<T>{h, ws.T1, R, y, ws.splitk, ws.splitk_elems};
    p.mix = false;
    return p;
  }
  const int dq = kind == 0 ? 4 : 2;
  QCA_CHECK(W->group_rows == Dout && W->nout == dq * Dout && W->nin == dq * Din,
            "sparse MPO does not match the block");
  // GEMM1, one batch per (q, w): T[q][w][c][x] = sum_a x[q][a][c] L[a][w][x],
  // so that the second GEMM reads T's slices with x contiguous (its M index)
  GemmDesc g = gemm_desc(chir, chil, chil);
  g.NB1 = dq; g.NB2 = Din;
  g.a_m = 1; g.a_k2 = chir; g.a_b1 = lr;
  g.b_k2 = (long)Din * chil; g.b_n = 1; g.b_b2 = chil;
  g.c_m = chil; g.c_n = 1; g.c_b1 = (long)Din * lr; g.c_b2 = lr;
  g.skip = skip;
  p.g1 = GemmJob<T>{g, x, L, ws.T1, ws.splitk, ws.splitk_elems};
  // GEMM3 with the MPO mixing gathered into its K loop:
  //   y[q'][x][y] = sum_{(q,w,v) in W[.][q'][.][.]} W[q][q'][w][v] sum_c T[q][w][c][x] R[c][v][y]
  // one K segment per nonzero of the column block q' (seg_col = q*Din + w,
  // seg_outv = v), each a sum over c.
  p.mix = false;
  GemmDesc h = gemm_desc(chil, chir, W->group_max * chir);
  h.NB2 = dq;
  h.K2 = chir;
  h.seg_rows = Dout;
  h.seg_max = W->group_max;
  h.seg_rowptr = W->rowptr;
  h.seg_col = W->col;
  h.seg_outv = W->outv;
  h.seg_val = W->val;
  if (W->ngroups == dq) {
    h.seg_starts = 1;
    for (int g = 0; g <= dq; ++g) h.seg_start[g] = W->group_start[g];
  }
  h.a_m = 1; h.a_k1 = lr; h.a_k2 = chil;
  h.b_k1 = chir; h.b_k2 = (long)Dout * chir; h.b_n = 1;
  h.c_m = chir; h.c_n = 1; h.c_b2 = lr;
  h.skip = skip;
  p.g3 = GemmJob<T>{h, ws.T1, R, y, ws.splitk, ws.splitk_elems};
  return p;
}

template <typename T>
ContractPlan<T> env_plan(bool left, const cplx<T>* A, const cplx<T>* Eprev, cplx<T>* Enew,
                         const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                         const ContractWs<T>& ws) {
  ContractPlan<T> p{};
  const long lr = (long)chil * chir;
  p.mix = true;
  if (left) {
    // T[p][(w,x)][b] = sum_a L[a][(w,x)] A[p][a][b]
    GemmDesc g = gemm_desc(Din * chil, chir, chil);
    g.NB2 = 2;
    g.a_m = 1; g.a_k2 = (long)Din * chil;
    g.b_k2 = chir; g.b_n = 1; g.b_b2 = lr;
    g.c_m = chir; g.c_n = 1; g.c_b2 = (long)Din * lr;
    p.g1 = GemmJob<T>{g, Eprev, A, ws.T1, ws.splitk, ws.splitk_elems};
    p.wm = WmixArgs<T>{ws.T1, lr, lr, 2 * Din, ws.U, 2, Dout, W->rowptr, W->col, W->val, nullptr};
    // Lnew[(b,w')][b*] = sum_{(p',x)} U[(p',x)][(b,w')] conj(A[(p',x)][b*])
    GemmDesc h = gemm_desc(chir * Dout, chir, 2 * chil);
    h.a_m = 1; h.a_k2 = (long)chir * Dout;
    h.b_k2 = chir; h.b_n = 1; h.conj_b = 1;
    h.c_m = chir; h.c_n = 1;
    p.g3 = GemmJob<T>{h, ws.U, A, Enew, ws.splitk, ws.splitk_elems};
  } else {
    // T[p][(w',y)][a] = sum_b R[b][(w',y)] A[p][a][b]
    GemmDesc g = gemm_desc(Din * chir, chil, chir);
    g.NB2 = 2;
    g.a_m = 1; g.a_k2 = (long)Din * chir;
    g.b_k2 = 1; g.b_n = chir; g.b_b2 = lr;
    g.c_m = chil; g.c_n = 1; g.c_b2 = (long)Din * lr;
    p.g1 = GemmJob<T>{g, Eprev, A, ws.T1, ws.splitk, ws.splitk_elems};
    p.wm = WmixArgs<T>{ws.T1, lr, lr, 2 * Din, ws.U, 2, Dout, W->rowptr, W->col, W->val, nullptr};
    // Rnew[(a,w)][a*] = sum_{(p,y)} U[(p,y)][(a,w)] conj(A[p][a*][y])
    GemmDesc h = gemm_desc(chil * Dout, chil, 2 * chir);
    h.K2 = chir;
    h.a_m = 1; h.a_k2 = (long)chil * Dout; h.a_k1 = (long)chir * chil * Dout;
    h.b_k1 = lr; h.b_k2 = 1; h.b_n = chir; h.conj_b = 1;
    h.c_m = chil; h.c_n = 1;
    p.g3 = GemmJob<T>{h, ws.U, A, Enew, ws.splitk, ws.splitk_elems};
  }
  return p;
}

template <typename T>
void run_plan(const ContractPlan<T>& p, hipStream_t s, const ContractWs<T>& ws) {
  gemm_t<T>(p.g1.d, p.g1.A, p.g1.B, p.g1.C, s, ws);
  if (p.mix) {
    WSparseDev<T> w;
    w.rowptr = const_cast<int*>(p.wm.rowptr);
    w.col = const_cast<int*>(p.wm.col);
    w.val = const_cast<cplx<T>*>(p.wm.val);
    launch_wmix<T>(p.wm.Tin, p.wm.in_stride, p.wm.nxc, p.wm.nin, p.wm.U, p.wm.nq, p.wm.Dout, w, s, p.wm.skip);
  }
  gemm_t<T>(p.g3.d, p.g3.A, p.g3.B, p.g3.C, s, ws);
}

template <typename T>
void heff_apply(int kind, const cplx<T>* x, cplx<T>* y, const cplx<T>* L, const cplx<T>* R,
                const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                const ContractWs<T>& ws, const int* skip, hipStream_t s) {
  run_plan<T>(heff_plan<T>(kind, x, y, L, R, W, Din, Dout, chil, chir, ws, skip), s, ws);
}

template <typename T>
void env_update(bool left, const cplx<T>* A, const cplx<T>* Eprev, cplx<T>* Enew,
                const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                const ContractWs<T>& ws, hipStream_t s) {
  run_plan<T>(env_plan<T>(left, A, Eprev, Enew, W, Din, Dout, chil, chir, ws), s, ws);
}

template ContractPlan<double> heff_plan<double>(int, const c128*, c128*, const c128*, const c128*,
                                                const WSparseDev<double>*, int, int, int, int,
                                                const ContractWs<double>&, const int*);
template ContractPlan<float> heff_plan<float>(int, const c64*, c64*, const c64*, const c64*,
                                              const WSparseDev<float>*, int, int, int, int,
                                              const ContractWs<float>&, const int*);
template ContractPlan<double> env_plan<double>(bool, const c128*, const c128*, c128*,
                                               const WSparseDev<double>*, int, int, int, int,
                                               const ContractWs<double>&);
template ContractPlan<float> env_plan<float>(bool, const c64*, const c64*, c64*, const WSparseDev<float>*,
                                             int, int, int, int, const ContractWs<float>&);
template void heff_apply<double>(int, const c128*, c128*, const c128*, const c128*,
                                 const WSparseDev<double>*, int, int, int, int,
                                 const ContractWs<double>&, const int*, hipStream_t);
template void heff_apply<float>(int, const c64*, c64*, const c64*, const c64*,
                                const WSparseDev<float>*, int, int, int, int,
                                const ContractWs<float>&, const int*, hipStream_t);
template void env_update<double>(bool, const c128*, const c128*, c128*, const WSparseDev<double>*,
                                 int, int, int, int, const ContractWs<double>&, hipStream_t);
template void env_update<float>(bool, const c64*, const c64*, c64*, const WSparseDev<float>*, int,
                                int, int, int, const ContractWs<float>&, hipStream_t);

double heff_flops(int kind, long chil, long chir, long D) {
  // algorithmic complex multiply-adds x 8; the MPO mixing is counted dense
  if (kind == 0) return 8.0 * (4.0 * D * chil * chir * (chil + chir) + 16.0 * D * D * chil * chir);
  if (kind == 1) return 8.0 * (2.0 * D * chil * chir * (chil + chir) + 4.0 * D * D * chil * chir);
  return 8.0 * (double)D * chil * chir * (chil + chir);
}

HostSparse build_sparse(const c128* W, int dq, int Dl, int Dr, bool backward) {
  HostSparse h;
  const int nout = backward ? dq * Dl : dq * Dr;
  h.rowptr.assign(nout + 1, 0);
  std::vector<std::vector<std::pair<int, c128>>> rows(nout);
  for (int p = 0; p < dq; ++p)
    for (int q = 0; q < dq; ++q)
      for (int w = 0; w < Dl; ++w)
        for (int v = 0; v < Dr; ++v) {
          const c128 x = W[((long)(p * dq + q) * Dl + w) * Dr + v];
          if (x.re == 0.0 && x.im == 0.0) continue;
          if (!backward) rows[q * Dr + v].push_back({p * Dl + w, x});
          else rows[p * Dl + w].push_back({q * Dr + v, x});
        }
  h.group_rows = backward ? Dl : Dr;
  for (int o = 0; o < nout; ++o) {
    for (auto& e : rows[o]) {
      h.col.push_back(e.first);
      h.val.push_back(e.second);
      h.outv.push_back(o % h.group_rows);
    }
    h.rowptr[o + 1] = (int)h.col.size();
  }
  for (int g = 0; g < dq; ++g)
    h.group_max = std::max(h.group_max, h.rowptr[(g + 1) * h.group_rows] - h.rowptr[g * h.group_rows]);
  return h;
}

// W2[(p1 p2), (p1' p2'), wl, wr] = sum_m Wl[p1,p1',wl,m] Wr[p2,p2',m,wr]
// (reference tdvp.py:282-284)
std::vector<c128> merge_mpo(const c128* Wl, const c128* Wr, int Dl, int Dm, int Dr) {
  std::vector<c128> out((size_t)16 * Dl * Dr, c128{0, 0});
  for (int p1 = 0; p1 < 2; ++p1)
    for (int q1 = 0; q1 < 2; ++q1)
      for (int p2 = 0; p2 < 2; ++p2)
        for (int q2 = 0; q2 < 2; ++q2)
          for (int a = 0; a < Dl; ++a)
            for (int b = 0; b < Dr; ++b) {
              double re = 0, im = 0;
              for (int m = 0; m < Dm; ++m) {
                const c128 x = Wl[((p1 * 2 + q1) * Dl + a) * Dm + m];
                const c128 y = Wr[((p2 * 2 + q2) * Dm + m) * Dr + b];
                re += x.re * y.re - x.im * y.im;
                im += x.re * y.im + x.im * y.re;
              }
              const int P = p1 * 2 + p2, Q = q1 * 2 + q2;
              out[((size_t)(P * 4 + Q) * Dl + a) * Dr + b] = c128{re, im};
            }
  return out;
}

namespace {
size_t sparse_alloc_size(size_t bytes) { return bytes < 4 ? 4 : bytes; }
}  // namespace

size_t sparse_device_bytes(const HostSparse& h) {
  const size_t nout = h.rowptr.size() - 1, nnz = h.col.size();
  return sparse_alloc_size(sizeof(int) * (nout + 1)) + 2 * sparse_alloc_size(sizeof(int) * nnz) +
         sparse_alloc_size(sizeof(c128) * nnz) + sparse_alloc_size(sizeof(c64) * nnz);
}

void WSparsePair::upload(const HostSparse& h, int nin) {
  const int nout = (int)h.rowptr.size() - 1, nnz = (int)h.col.size();
  auto alloc = [&](size_t bytes) {
    void* p = nullptr;
    QCA_HIP(hipMalloc(&p, sparse_alloc_size(bytes)));
    owned.push_back(p);
    this->bytes += sparse_alloc_size(bytes);
    return p;
  };
  d64.rowptr = (int*)alloc(sizeof(int) * (nout + 1));
  d64.col = (int*)alloc(sizeof(int) * nnz);
  d64.outv = (int*)alloc(sizeof(int) * nnz);
  d64.val = (c128*)alloc(sizeof(c128) * nnz);
  d32.val = (c64*)alloc(sizeof(c64) * nnz);
  QCA_HIP(hipMemcpy(d64.rowptr, h.rowptr.data(), sizeof(int) * (nout + 1), hipMemcpyHostToDevice));
  if (nnz) {
    QCA_HIP(hipMemcpy(d64.col, h.col.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
    QCA_HIP(hipMemcpy(d64.outv, h.outv.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
    QCA_HIP(hipMemcpy(d64.val, h.val.data(), sizeof(c128) * nnz, hipMemcpyHostToDevice));
    std::vector<c64> v32(nnz);
    for (int z = 0; z < nnz; ++z) v32[z] = c64{(float)h.val[z].re, (float)h.val[z].im};
    QCA_HIP(hipMemcpy(d32.val, v32.data(), sizeof(c64) * nnz, hipMemcpyHostToDevice));
  }
  d32.rowptr = d64.rowptr;
  d32.col = d64.col;
  d32.outv = d64.outv;
  d64.group_rows = d32.group_rows = h.group_rows;
  d64.group_max = d32.group_max = h.group_max;
  set_group_starts(d64, h.rowptr, h.group_rows);
  set_group_starts(d32, h.rowptr, h.group_rows);
  d64.nout = d32.nout = nout;
  d64.nin = d32.nin = nin;
  d64.nnz = d32.nnz = nnz;
}

void WSparsePair::release() {
  for (void* p : owned) (void)hipFree(p);
  owned.clear();
  bytes = 0;
}

}  // namespace qca
