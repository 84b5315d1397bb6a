// Device-resident TDVP engine (host orchestration of the gfx950 kernels).
//
// Follows reference algorithms/tdvp.py in gauge-consistent form:
//   * 2TDVP (tdvp.py:107-145): per bond merge -> Krylov exp(+dt/2) -> SVD split
//     and truncation -> left/right environment update -> 1-site exp(-dt/2).
//   * 1TDVP (tdvp.py:65-105): 1-site exp(+dt/2) -> complete QR truncated/padded
//     to the target bond -> environment update -> bond exp(-dt/2) -> absorb.
// The reference's per-step make_site_canonical(0) (tdvp.py:54) only changes the
// gauge of tensors that are already right-canonical and leaves the cached right
// environments stale; here the state stays in the gauge the environments were
// built in, so no re-canonicalisation is done between steps, and measurement
// does not move the orthogonality centre.
//
// Every tensor lives in HBM for the lifetime of the engine.  The host only
// reads back the truncation rank after each SVD and the convergence flag of
// each Krylov exponential.
#include "engine.hpp"

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace qca {

namespace {
constexpr double kHalfPi = 1.5707963267948966;

int sat_pow2_min(int e_left, int e_right, int cap) {
  // min(2^e_left, 2^e_right, cap) without overflow (reference tdvp.py:31-33)
  long v = cap;
  if (e_left < 31) v = std::min<long>(v, 1L << e_left);
  if (e_right < 31) v = std::min<long>(v, 1L << e_right);
  return (int)v;
}

}  // namespace

template <typename T>
T* Engine::dmalloc(long n) {
  void* p = nullptr;
  if (n <= 0) n = 1;
  device_bytes_ += sizeof(T) * (size_t)n;
  if (count_only_) return nullptr;
  QCA_HIP(hipMalloc(&p, sizeof(T) * (size_t)n));
  allocs_.push_back(p);
  return reinterpret_cast<T*>(p);
}

// Uploads a sparse MPO form (or, when only counting, adds its size).
void Engine::upload_sparse(WSparsePair& dst, const HostSparse& h, int nin) {
  if (count_only_) {
    device_bytes_ += sparse_device_bytes(h);
    return;
  }
  dst.upload(h, nin);
  device_bytes_ += dst.bytes;
}

size_t Engine::footprint(const EngineParams& p, const std::vector<const c128*>& W_host,
                         const std::vector<int>& Wdl, const std::vector<int>& Wdr,
                         const std::vector<int>& chis) {
  Engine e(p, W_host, Wdl, Wdr, chis);
  return e.device_bytes_;
}

Engine::Engine(const EngineParams& p, const std::vector<const c128*>& W_host,
               const std::vector<int>& Wdl, const std::vector<int>& Wdr, const std::vector<int>& chis)
    : p_(p), N_(p.ncells), device_(-1), count_only_(true) {
  const std::vector<const c128*> no_state(N_, nullptr);
  init(W_host, Wdl, Wdr, no_state, chis);
}

Engine::Engine(int device, const EngineParams& p, const std::vector<const c128*>& W_host,
               const std::vector<int>& Wdl, const std::vector<int>& Wdr,
               const std::vector<const c128*>& A_host, const std::vector<int>& chis)
    : p_(p), N_(p.ncells), device_(device) {
  // A constructor that throws never reaches the destructor: release whatever
  // was allocated before the failure, then rethrow.
  try {
    init(W_host, Wdl, Wdr, A_host, chis);
  } catch (...) {
    release_all();
    throw;
  }
}

void Engine::init(const std::vector<const c128*>& W_host, const std::vector<int>& Wdl,
                  const std::vector<int>& Wdr, const std::vector<const c128*>& A_host,
                  const std::vector<int>& chis) {
  const EngineParams& p = p_;
  const int device = device_;
  QCA_CHECK(N_ >= 2, "at least 2 cells are required");
  QCA_CHECK((int)W_host.size() == N_ && (int)Wdl.size() == N_ && (int)Wdr.size() == N_,
            "one MPO tensor per cell");
  QCA_CHECK((int)A_host.size() == N_ && (int)chis.size() == N_ + 1, "one MPS tensor per cell");
  QCA_CHECK(chis[0] == 1 && chis[N_] == 1, "open boundary bonds must be 1");
  QCA_CHECK(p.algorithm >= 1 && p.algorithm <= 3,
            "algorithm must be 1 (1TDVP), 2 (2TDVP) or 3 (adaptive 1TDVP)");
  QCA_CHECK(p.precision == PREC_FP64 || p.precision == PREC_MIXED, "unknown precision");
  QCA_CHECK(p.chi_max >= 1 && p.chi_max <= 256, "max bond dimension must be in 1..256");
  QCA_CHECK(Wdl[0] == 1 && Wdr[N_ - 1] == 1, "MPO boundary bonds must be 1");
  for (int i = 0; i + 1 < N_; ++i) QCA_CHECK(Wdr[i] == Wdl[i + 1], "MPO bond mismatch");
  for (int i = 0; i < N_; ++i) QCA_CHECK(chis[i] >= 1 && chis[i] <= 256, "MPS bond out of range");
  if (!count_only_) {
    QCA_HIP(hipSetDevice(device));
    QCA_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    own_stream_ = true;
  }

  chi_ = chis;
  cap_.resize(N_ + 1);
  target_.resize(N_ + 1);
  maxdim_.resize(N_ + 1);
  for (int i = 0; i <= N_; ++i) {
    maxdim_[i] = target_[i] = sat_pow2_min(i, N_ - i, p.chi_max);
    cap_[i] = std::max(target_[i], chi_[i]);
  }
  Dl_ = Wdl;
  Dr_ = Wdr;
  Dmax_ = 1;
  for (int i = 0; i < N_; ++i) Dmax_ = std::max({Dmax_, Dl_[i], Dr_[i]});
  QCA_CHECK(Dmax_ <= 16, "MPO bond dimension above 16");

  // ---- MPO in sparse form, both precisions
  Wf_.resize(N_);
  Wb_.resize(N_);
  for (int i = 0; i < N_; ++i) {
    upload_sparse(Wf_[i], build_sparse(W_host[i], 2, Dl_[i], Dr_[i], false), 2 * Dl_[i]);
    upload_sparse(Wb_[i], build_sparse(W_host[i], 2, Dl_[i], Dr_[i], true), 2 * Dr_[i]);
  }
  if (p.algorithm != 1) {  // two-site blocks: 2TDVP and the adaptive targets
    W2f_.resize(N_ - 1);
    for (int s = 0; s + 1 < N_; ++s) {
      auto w2 = merge_mpo(W_host[s], W_host[s + 1], Dl_[s], Dr_[s], Dr_[s + 1]);
      upload_sparse(W2f_[s], build_sparse(w2.data(), 4, Dl_[s], Dr_[s + 1], false), 4 * Dl_[s]);
    }
  }

  // ---- state
  const bool mixed = p.precision == PREC_MIXED;
  const size_t esz = mixed ? sizeof(c64) : sizeof(c128);
  A_.resize(N_);
  for (int i = 0; i < N_; ++i) {
    A_[i] = dmalloc<c128>(2L * cap_[i] * cap_[i + 1]);
    if (!count_only_)
      QCA_HIP(hipMemcpy(A_[i], A_host[i], sizeof(c128) * 2L * chi_[i] * chi_[i + 1],
                        hipMemcpyHostToDevice));
  }
  L_.assign(N_, nullptr);
  R_.assign(N_, nullptr);
  for (int i = 0; i < N_; ++i) {
    L_[i] = dmalloc<unsigned char>((long)esz * cap_[i + 1] * Dr_[i] * cap_[i + 1]);
    R_[i] = dmalloc<unsigned char>((long)esz * cap_[i] * Dl_[i] * cap_[i]);
  }
  one64_ = dmalloc<c128>(1);
  one32_ = dmalloc<c64>(1);
  if (!count_only_) {
    c128 o{1.0, 0.0};
    c64 o32{1.0f, 0.0f};
    QCA_HIP(hipMemcpy(one64_, &o, sizeof(o), hipMemcpyHostToDevice));
    QCA_HIP(hipMemcpy(one32_, &o32, sizeof(o32), hipMemcpyHostToDevice));
  }

  // ---- workspaces
  long cmax = 1;
  for (int i = 0; i <= N_; ++i) cmax = std::max<long>(cmax, cap_[i]);
  long nmax = 1;  // longest Krylov vector
  for (int s = 0; s < N_; ++s) nmax = std::max<long>(nmax, 2L * cap_[s] * cap_[s + 1]);
  for (int s = 0; s + 2 <= N_; ++s) nmax = std::max<long>(nmax, 4L * cap_[s] * cap_[s + 2]);
  for (int b = 0; b <= N_; ++b) nmax = std::max<long>(nmax, (long)cap_[b] * cap_[b]);
  const long tmax = 4L * Dmax_ * cmax * cmax;
  theta_ = dmalloc<c128>(4 * cmax * cmax);
  theta2_ = dmalloc<c128>(4 * cmax * cmax);
  svdX_ = dmalloc<c128>(4 * cmax * cmax);
  qrQ_ = dmalloc<c128>(4 * cmax * cmax);
  qrTau_ = dmalloc<c128>(2 * cmax);
  cbond_ = dmalloc<c128>(cmax * cmax * 2);
  qr_ws_elems_ = qr_workspace_elems((int)std::min<long>(2 * cmax, 512), (int)cmax,
                                    (int)std::min<long>(2 * cmax, 512));
  qrWs_ = dmalloc<c128>((long)qr_ws_elems_);
  T_ = dmalloc<c128>(tmax);
  U_ = dmalloc<c128>(tmax);
  ws_elems_ = std::max<long>(16 * nmax, 4 * Dmax_ * cmax * cmax);
  ws_ = dmalloc<c128>(ws_elems_);
  ws64_.T1 = reinterpret_cast<c128*>(T_);
  ws64_.U = reinterpret_cast<c128*>(U_);
  ws64_.splitk = reinterpret_cast<c128*>(ws_);
  ws64_.splitk_elems = ws_elems_;
  ws32_.T1 = reinterpret_cast<c64*>(T_);
  ws32_.U = reinterpret_cast<c64*>(U_);
  ws32_.splitk = reinterpret_cast<c64*>(ws_);
  ws32_.splitk_elems = 2 * ws_elems_;
  if (mixed) {
    x32_ = dmalloc<c64>(std::max(nmax, 2 * cmax * cmax));
    y32_ = dmalloc<c64>(nmax);
    a32_ = dmalloc<c64>(2 * cmax * cmax);
  }
  y64_ = dmalloc<c128>(nmax);
  sigma_ = dmalloc<double>(2 * cmax);
  perm_ = dmalloc<int>(2 * cmax);
  svdres_ = dmalloc<SvdResult>(1);
  jw_.ws_bytes = jacobi_workspace_bytes(2 * cmax, 2 * cmax);
  jw_.ws = dmalloc<unsigned char>((long)jw_.ws_bytes);
  if (2 * cmax > 32) {
    jw_.pre_elems = jacobi_precond_elems((int)std::min<long>(2 * cmax, 512),
                                         (int)std::min<long>(2 * cmax, 512));
    jw_.pre = dmalloc<c128>((long)jw_.pre_elems);
  }
  if (!count_only_) QCA_HIP(hipHostMalloc(&jw_.host, 64));
  if (mixed) {
    jw_.tol = MIXED_SVD_TOL;
    if (const char* env = std::getenv("QCA_MIXED_SVD_TOL")) {
      char* end = nullptr;
      const double v = std::strtod(env, &end);
      QCA_CHECK(end != env && *end == '\0' && v >= 0 && v < 1, "QCA_MIXED_SVD_TOL must be in [0, 1)");
      jw_.tol = v;
    }
  }
  if (!count_only_) {
    QCA_HIP(hipHostMalloc((void**)&svdres_host_, sizeof(SvdResult)));
    QCA_HIP(hipHostMalloc((void**)&ctrl_host_, sizeof(LanczosCtrl)));
    QCA_HIP(hipEventCreate(&ev0_));
    QCA_HIP(hipEventCreate(&ev1_));
  }
  lw_.max_n = nmax;
  lw_.max_blocks = lanczos_nblk(nmax);
  lw_.ctrl = dmalloc<LanczosCtrl>(1);
  lw_.V = dmalloc<c128>((long)LZ_MAXK * nmax);
  lw_.w64 = dmalloc<c128>(nmax);
  lw_.part = dmalloc<c128>((long)lw_.max_blocks * LZ_MAXK);
  lw_.part2 = dmalloc<c128>((long)lw_.max_blocks * LZ_MAXK);
  lw_.part_norm = dmalloc<double>(lw_.max_blocks);
  lw_.coef1 = dmalloc<c128>(LZ_MAXK);
  lw_.coef2 = dmalloc<c128>(LZ_MAXK);
  tally_ = dmalloc<KrylovTally>(1);
  if (!count_only_) QCA_HIP(hipMemset(tally_, 0, sizeof(KrylovTally)));
  fused_max_n_ = FUSED_KRYLOV_DEFAULT_N;
  if (const char* env = std::getenv("QCA_FUSED_KRYLOV_MAX_N")) {
    char* end = nullptr;
    const long v = std::strtol(env, &end, 10);
    QCA_CHECK(end != env && *end == '\0' && v >= 0, "QCA_FUSED_KRYLOV_MAX_N must be a count >= 0");
    fused_max_n_ = v;
  }
  fused_max_n_ = std::min(fused_max_n_, FUSED_KRYLOV_MAX_N);
  measE_ = dmalloc<c128>(cmax * cmax);
  measT_ = dmalloc<c128>(2 * cmax * cmax);
  rho_ = dmalloc<c128>(4L * N_);
  obs_ = dmalloc<double>(3L * N_);
  if (count_only_) return;

  // ---- TDVP construction (reference tdvp.py:20-39): canonicalise, then the
  // right environments of every site.
  canonicalize_initial();
  for (int site = N_ - 1; site >= 1; --site) update_env(false, site);
  sync();
}

Engine::~Engine() { release_all(); }

// Frees everything the engine holds; safe on a partly constructed engine and
// when called twice.
void Engine::release_all() {
  if (count_only_) return;
  (void)hipSetDevice(device_);
  if (s_) (void)hipStreamSynchronize(s_);
  for (auto& w : Wf_) w.release();
  for (auto& w : Wb_) w.release();
  for (auto& w : W2f_) w.release();
  for (void* p : allocs_) (void)hipFree(p);
  allocs_.clear();
  if (svdres_host_) (void)hipHostFree(svdres_host_);
  if (ctrl_host_) (void)hipHostFree(ctrl_host_);
  if (jw_.host) (void)hipHostFree(jw_.host);
  svdres_host_ = nullptr;
  ctrl_host_ = nullptr;
  jw_.host = nullptr;
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  ev0_ = ev1_ = nullptr;
  if (own_stream_ && s_) (void)hipStreamDestroy(s_);
  s_ = nullptr;
  own_stream_ = false;
}

void Engine::sync() { QCA_HIP(hipStreamSynchronize(s_)); }

StepStats Engine::stats() {
  KrylovTally t;
  QCA_HIP(hipMemcpyAsync(&t, tally_, sizeof(t), hipMemcpyDeviceToHost, s_));
  sync();
  StepStats r = stats_;
  r.matvec_2site += t.matvecs[0];
  r.matvec_1site += t.matvecs[1];
  r.matvec_bond += t.matvecs[2];
  r.matvecs += t.matvecs[0] + t.matvecs[1] + t.matvecs[2];
  r.lanczos_limit_hits += t.limit_hits;
  r.flops += t.flops;
  int khz = 0;
  QCA_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device_));
  if (khz > 0) r.krylov_ms += (double)t.ticks / (double)khz;
  return r;
}

void Engine::reset_stats() {
  stats_ = StepStats();
  QCA_HIP(hipMemsetAsync(tally_, 0, sizeof(KrylovTally), s_));
}

double Engine::elapsed_ms() {
  float ms = 0.f;
  QCA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  return ms;
}

template <>
void Engine::gemm<double>(const GemmDesc& d, const c128* A, const c128* B, c128* C) {
  gemm_t<double>(d, A, B, C, s_, ws64_);
}
template <>
void Engine::gemm<float>(const GemmDesc& d, const c64* A, const c64* B, c64* C) {
  gemm_t<float>(d, A, B, C, s_, ws32_);
}

template <>
cplx<double>* Engine::Lenv<cplx<double>>(int site) const {
  return site < 0 ? one64_ : reinterpret_cast<c128*>(L_[site]);
}
template <>
cplx<float>* Engine::Lenv<cplx<float>>(int site) const {
  return site < 0 ? one32_ : reinterpret_cast<c64*>(L_[site]);
}
template <>
cplx<double>* Engine::Renv<cplx<double>>(int site) const {
  return site >= N_ ? one64_ : reinterpret_cast<c128*>(R_[site]);
}
template <>
cplx<float>* Engine::Renv<cplx<float>>(int site) const {
  return site >= N_ ? one32_ : reinterpret_cast<c64*>(R_[site]);
}

// ---------------------------------------------------------------- H_eff
// kind 0: 2-site at bond (site, site+1), x = Theta[q][a][c] (4, chil, chir)
// kind 1: 1-site at site, x = A[p][a][b] (2, chil, chir)
// kind 2: bond between site-1 and site, x = C[a][b] (chil, chir)
template <typename T>
void Engine::heff(int kind, int site, const cplx<T>* x, cplx<T>* y, int chil, int chir) {
  const int* skip = lanczos_done_flag(lw_);
  const bool dbl = std::is_same<T, double>::value;
  if (kind == 2) {
    const int D = Dl_[site];
    qca::heff_apply<T>(2, x, y, Lenv<cplx<T>>(site - 1), Renv<cplx<T>>(site), nullptr, D, D, chil,
                       chir, ws_for<T>(), skip, s_);
    return;
  }
  const WSparsePair& wp = kind == 0 ? W2f_[site] : Wf_[site];
  const WSparseDev<T>* W;
  if constexpr (std::is_same<T, double>::value) W = &wp.d64; else W = &wp.d32;
  (void)dbl;
  const int Din = Dl_[site];
  const int Dout = kind == 0 ? Dr_[site + 1] : Dr_[site];
  qca::heff_apply<T>(kind, x, y, Lenv<cplx<T>>(site - 1),
                     Renv<cplx<T>>(kind == 0 ? site + 2 : site + 1), W, Din, Dout, chil, chir,
                     ws_for<T>(), skip, s_);
}

// One-launch exponential: no host round trip, totals kept on the device.
template <typename T>
void Engine::fused_run(int kind, int site, c128* x, long n, int chil, int chir, double delta) {
  FusedKrylovArgs<T> a;
  a.kind = kind;
  a.chil = chil;
  a.chir = chir;
  a.n = n;
  if (kind == 2) {
    a.Din = a.Dout = Dl_[site];
    a.L = Lenv<cplx<T>>(site - 1);
    a.R = Renv<cplx<T>>(site);
  } else {
    const WSparsePair& wp = kind == 0 ? W2f_[site] : Wf_[site];
    a.Din = Dl_[site];
    a.Dout = kind == 0 ? Dr_[site + 1] : Dr_[site];
    a.L = Lenv<cplx<T>>(site - 1);
    a.R = Renv<cplx<T>>(kind == 0 ? site + 2 : site + 1);
    a.rowptr = wp.d64.rowptr;
    a.col = wp.d64.col;
    if constexpr (std::is_same<T, double>::value) a.val = wp.d64.val; else a.val = wp.d32.val;
    a.nnz = wp.d64.nnz;
  }
  a.x = x;
  a.V = lw_.V;
  a.T1 = reinterpret_cast<cplx<T>*>(T_);
  a.U = reinterpret_cast<cplx<T>*>(U_);
  a.tally = tally_;
  a.tau = kHalfPi * delta;
  a.tol = p_.krylov_tol;
  a.max_m = lw_.max_m;
  const int D = kind == 2 ? Dl_[site] : std::max(a.Din, a.Dout);
  a.flops_per_matvec = heff_flops(kind, chil, chir, D);
  launch_fused_krylov<T>(a, s_);
}

template <typename T>
void Engine::lanczos_run(int kind, int site, c128* x, long n, int chil, int chir, double delta) {
  if (n <= fused_max_n_) {
    fused_run<T>(kind, site, x, n, chil, chir, delta);
    return;
  }
  const double tau = kHalfPi * delta;
  QCA_HIP(hipEventRecord(ev0_, s_));
  constexpr bool mixed = !std::is_same<T, double>::value;
  // the mixed path's matvec operand (complex64 copy of V_j) comes with each new
  // basis vector from the Lanczos kernels
  lanczos_begin(x, n, lw_, s_, mixed ? x32_ : nullptr);
  int j = 0;
  int chunk = std::max(2, kpred_[kind]);
  while (true) {
    for (int c = 0; c < chunk && j < lw_.max_m - 1; ++c, ++j) {
      if constexpr (std::is_same<T, double>::value) {
        heff<double>(kind, site, lw_.V + (long)j * n, y64_, chil, chir);
        lanczos_iteration_tail<double>(y64_, n, j, tau, p_.krylov_tol, lw_, s_);
      } else {
        heff<float>(kind, site, x32_, y32_, chil, chir);
        lanczos_iteration_tail<float>(y32_, n, j, tau, p_.krylov_tol, lw_, s_, x32_);
      }
    }
    QCA_HIP(hipMemcpyAsync(ctrl_host_, lw_.ctrl, sizeof(LanczosCtrl), hipMemcpyDeviceToHost, s_));
    QCA_HIP(hipEventRecord(ev1_, s_));
    sync();
    if (ctrl_host_->done) break;
    // the device closes the exponential at the dimension limit (m = max_m - 1)
    // at the latest, so this only fires on non-finite input
    QCA_CHECK(j < lw_.max_m - 1, "Krylov exponential did not converge (non-finite input?)");
    chunk = 2;
  }
  stats_.krylov_ms += elapsed_ms();
  const int m = ctrl_host_->m;
  kpred_[kind] = std::max(2, m);
  if (ctrl_host_->hit_limit) stats_.lanczos_limit_hits++;
  lanczos_finish(n, x, lw_, s_);
  stats_.matvecs += m;
  if (kind == 0) stats_.matvec_2site += m;
  else if (kind == 1) stats_.matvec_1site += m;
  else stats_.matvec_bond += m;
  const int D = kind == 2 ? Dl_[site] : std::max(Dl_[site], kind == 0 ? Dr_[site + 1] : Dr_[site]);
  stats_.flops += m * heff_flops(kind, chil, chir, D);
}

void Engine::evolve(int kind, int site, c128* x, long n, int chil, int chir, double delta) {
  if (p_.precision == PREC_MIXED) lanczos_run<float>(kind, site, x, n, chil, chir, delta);
  else lanczos_run<double>(kind, site, x, n, chil, chir, delta);
}

// ---------------------------------------------------------------- environments
// L[site][b][w'][b*] = sum A[p][a][b] W[p][p'][w][w'] conj(A[p'][a*][b*]) L[site-1][a][w][a*]
// R[site][a][w][a*] = sum A[p][a][b] W[p][p'][w][w'] conj(A[p'][a*][b*]) R[site+1][b][w'][b*]
template <typename T>
void Engine::env_update_t(bool left, int site) {
  const int chil = chi_[site], chir = chi_[site + 1];
  const long lr = (long)chil * chir;
  const cplx<T>* A;
  if constexpr (std::is_same<T, double>::value) {
    A = A_[site];
  } else {
    launch_convert_c128_c64(A_[site], a32_, 2 * lr, s_);
    A = a32_;
  }
  const WSparsePair& wp = left ? Wf_[site] : Wb_[site];
  const WSparseDev<T>* W;
  if constexpr (std::is_same<T, double>::value) W = &wp.d64; else W = &wp.d32;
  if (left)
    env_update<T>(true, A, Lenv<cplx<T>>(site - 1), Lenv<cplx<T>>(site), W, Dl_[site], Dr_[site],
                  chil, chir, ws_for<T>(), s_);
  else
    env_update<T>(false, A, Renv<cplx<T>>(site + 1), Renv<cplx<T>>(site), W, Dr_[site], Dl_[site],
                  chil, chir, ws_for<T>(), s_);
}

void Engine::update_env(bool left, int site) {
  if (p_.precision == PREC_MIXED) env_update_t<float>(left, site);
  else env_update_t<double>(left, site);
  const int chil = chi_[site], chir = chi_[site + 1];
  const int D = std::max(Dl_[site], Dr_[site]);
  stats_.flops += heff_flops(1, chil, chir, D);
}

// ---------------------------------------------------------------- 2-site split
// Theta[p1][p2][a][c] = sum_b A_s[p1][a][b] A_{s+1}[p2][b][c]  (reference tdvp.py:278-280)
GemmDesc merge_theta_desc(int chil, int chim, int chir) {
  GemmDesc g = gemm_desc(chil, chir, chim);
  g.NB1 = 2; g.NB2 = 2;
  g.a_m = chim; g.a_k2 = 1; g.a_b1 = (long)chil * chim;
  g.b_k2 = chir; g.b_n = 1; g.b_b2 = (long)chim * chir;
  g.c_m = chir; g.c_n = 1; g.c_b1 = 2L * chil * chir; g.c_b2 = (long)chil * chir;
  return g;
}

// The factor that takes the (normalised) singular values after the split
// (reference tdvp.py:293-296), as a GEMM of the gathered isometry with Theta.
// right_sweep: A[s+1][p2][k][c] = sum_{(p1,a)} conj(A[s][(p1,a)][k]) Theta[p1][p2][a][c] / |s_k|
//              (A = A[s], B = Theta, C = A[s+1]);
// else:        A[s][p1][a][k] = sum_{(p2,c)} Theta[p1][p2][a][c] conj(A[s+1][p2][k][c]) / |s_k|
//              (A = Theta, B = A[s+1], C = A[s]).
GemmDesc split_theta_desc(bool right_sweep, int chil, int chir, int k, double inv_norm) {
  if (right_sweep) {
    GemmDesc g = gemm_desc(k, chir, 2 * chil);
    g.K2 = chil;
    g.NB2 = 2;
    g.a_m = 1; g.a_k1 = (long)chil * k; g.a_k2 = k; g.conj_a = 1;
    g.b_k1 = 2L * chil * chir; g.b_k2 = chir; g.b_n = 1; g.b_b2 = (long)chil * chir;
    g.c_m = chir; g.c_n = 1; g.c_b2 = (long)k * chir;
    g.alpha_re = inv_norm;
    return g;
  }
  GemmDesc g = gemm_desc(chil, k, 2 * chir);
  g.K2 = chir;
  g.NB2 = 2;
  g.a_m = chir; g.a_k1 = (long)chil * chir; g.a_k2 = 1; g.a_b2 = 2L * chil * chir;
  g.b_k1 = (long)k * chir; g.b_k2 = 1; g.b_n = chir; g.conj_b = 1;
  g.c_m = k; g.c_n = 1; g.c_b2 = (long)chil * k;
  g.alpha_re = inv_norm;
  return g;
}

void Engine::merge_theta(int s, int chil, int chim, int chir) {
  gemm<double>(merge_theta_desc(chil, chim, chir), A_[s], A_[s + 1], theta_);
}

int Engine::split_theta(int s, bool right_sweep, int chil, int chir) {
  const int rows = right_sweep ? 2 * chil : 2 * chir;
  const int cols = right_sweep ? 2 * chir : 2 * chil;
  const int rank_cap = std::min(2 * chil, 2 * chir);
  QCA_HIP(hipEventRecord(ev0_, s_));
  launch_theta_to_svd(theta_, svdX_, chil, chir, right_sweep ? 0 : 1, s_);
  launch_jacobi_svd(svdX_, rows, cols, sigma_, perm_, svdres_, p_.chi_max, p_.eps, rank_cap, s_, jw_);
  QCA_HIP(hipMemcpyAsync(svdres_host_, svdres_, sizeof(SvdResult), hipMemcpyDeviceToHost, s_));
  QCA_HIP(hipEventRecord(ev1_, s_));
  sync();
  stats_.svd_ms += elapsed_ms();
  stats_.svd_calls++;
  const int k = svdres_host_->k;
  stats_.svd_sweeps += svdres_host_->nsweeps;
  if (!svdres_host_->converged)
    throw Error(QCA_ERR_NOT_CONVERGED_CODE, "Jacobi SVD of the two-site split did not converge");
  QCA_CHECK(k >= 1, "SVD truncation produced an empty bond");
  QCA_CHECK(k <= cap_[s + 1], "truncation rank above bond capacity");
  const double inv = 1.0 / svdres_host_->norm_kept;
  if (right_sweep) {
    // A[s][p][a][k] = U[(p,a)][k]
    launch_gather_isometry(svdX_, perm_, sigma_, A_[s], rows, chil, k, 0, s_);
    gemm<double>(split_theta_desc(true, chil, chir, k, inv), A_[s], theta_, A_[s + 1]);
  } else {
    // A[s+1][p2][k][c] = conj(V[(p2,c)][k])
    launch_gather_isometry(svdX_, perm_, sigma_, A_[s + 1], rows, chir, k, 1, s_);
    gemm<double>(split_theta_desc(false, chil, chir, k, inv), theta_, A_[s + 1], A_[s]);
  }
  chi_[s + 1] = k;
  return k;
}

void Engine::sweep_right_2() {
  const double h = p_.dt / 2;
  for (int s = 0; s + 1 < N_; ++s) {
    const int chil = chi_[s], chim = chi_[s + 1], chir = chi_[s + 2];
    merge_theta(s, chil, chim, chir);
    evolve(0, s, theta_, 4L * chil * chir, chil, chir, h);
    const int k = split_theta(s, true, chil, chir);
    stats_.site_updates++;
    if (s < N_ - 2) {
      update_env(true, s);
      evolve(1, s + 1, A_[s + 1], 2L * k * chir, k, chir, -h);
    }
  }
}

void Engine::sweep_left_2() {
  const double h = p_.dt / 2;
  for (int s = N_ - 1; s >= 1; --s) {
    const int chil = chi_[s - 1], chim = chi_[s], chir = chi_[s + 1];
    merge_theta(s - 1, chil, chim, chir);
    evolve(0, s - 1, theta_, 4L * chil * chir, chil, chir, h);
    const int k = split_theta(s - 1, false, chil, chir);
    stats_.site_updates++;
    if (s > 1) {
      update_env(false, s);
      evolve(1, s - 1, A_[s - 1], 2L * chil * k, chil, k, -h);
    }
  }
}

// ---------------------------------------------------------------- QR shifts
// Left QR of A[site] as (2 chil x chir): A[site] <- Q[:, :target] (zero padded),
// C <- R[:target, :chir].  (reference mps.py:83-88, tdvp.py:74-78)
void Engine::qr_left_site(int site, int target, c128* C, int& crow, int& ccol, bool) {
  const int chil = chi_[site], chir = chi_[site + 1];
  const int rows = 2 * chil, cols = chir;
  QCA_CHECK(target <= cap_[site + 1], "target bond above capacity");
  launch_to_colmajor(A_[site], svdX_, rows, cols, s_);
  const int qcols = std::min(target, rows);
  launch_qr(svdX_, rows, cols, qrTau_, qrQ_, qcols, qrWs_, qr_ws_elems_, s_);
  launch_qr_gather_left(qrQ_, svdX_, A_[site], C, rows, cols, qcols, target, s_);
  crow = target;
  ccol = chir;
}

// Right QR of A[site]^T as (2 chir x chil): A[site] <- Q^T[:target] (2, target, chir),
// C <- R^T[:, :target] (chil x target).  (reference mps.py:90-98, tdvp.py:95-99)
void Engine::qr_right_site(int site, int target, c128* C, int& crow, int& ccol, bool) {
  const int chil = chi_[site], chir = chi_[site + 1];
  const int rows = 2 * chir, cols = chil;
  QCA_CHECK(target <= cap_[site], "target bond above capacity");
  launch_site_to_rightqr(A_[site], svdX_, chil, chir, s_);
  const int qcols = std::min(target, rows);
  launch_qr(svdX_, rows, cols, qrTau_, qrQ_, qcols, qrWs_, qr_ws_elems_, s_);
  launch_qr_gather_right(qrQ_, svdX_, A_[site], C, chir, cols, qcols, target, s_);
  crow = chil;
  ccol = target;
}

// A[site] (2, ccol, chi_r) <- C (crow x ccol) * A[site]  => (2, crow, chi_r)
void Engine::absorb_left(int site, const c128* C, int crow, int ccol) {
  const int chir = chi_[site + 1];
  GemmDesc g = gemm_desc(crow, chir, ccol);
  g.NB2 = 2;
  g.a_m = ccol; g.a_k2 = 1;
  g.b_k2 = chir; g.b_n = 1; g.b_b2 = (long)ccol * chir;
  g.c_m = chir; g.c_n = 1; g.c_b2 = (long)crow * chir;
  gemm<double>(g, C, A_[site], theta2_);
  QCA_HIP(hipMemcpyAsync(A_[site], theta2_, sizeof(c128) * 2L * crow * chir,
                         hipMemcpyDeviceToDevice, s_));
}

// A[site] (2, chi_l, crow) <- A[site] * C (crow x ccol)  => (2, chi_l, ccol)
void Engine::absorb_right(int site, const c128* C, int crow, int ccol) {
  const int chil = chi_[site];
  GemmDesc g = gemm_desc(2 * chil, ccol, crow);
  g.a_m = crow; g.a_k2 = 1;
  g.b_k2 = ccol; g.b_n = 1;
  g.c_m = ccol; g.c_n = 1;
  gemm<double>(g, A_[site], C, theta2_);
  QCA_HIP(hipMemcpyAsync(A_[site], theta2_, sizeof(c128) * 2L * chil * ccol,
                         hipMemcpyDeviceToDevice, s_));
}

// make_site_canonical(N-1) then make_site_canonical(0) with reduced QRs whose
// factors are cut or zero-padded to keep every bond dimension
// (reference tdvp.py:24-26, mps.py:146-192).
void Engine::canonicalize_initial() {
  int cr, cc;
  for (int j = 0; j + 1 < N_; ++j) {
    qr_left_site(j, chi_[j + 1], cbond_, cr, cc, false);
    absorb_left(j + 1, cbond_, cr, cc);
  }
  for (int j = N_ - 1; j >= 1; --j) {
    qr_right_site(j, chi_[j], cbond_, cr, cc, false);
    absorb_right(j - 1, cbond_, cr, cc);
  }
}

void Engine::sweep_right_1() {
  const double h = p_.dt / 2;
  for (int site = 0; site < N_; ++site) {
    const int chil = chi_[site], chir = chi_[site + 1];
    evolve(1, site, A_[site], 2L * chil * chir, chil, chir, h);
    stats_.site_updates++;
    if (site == N_ - 1) continue;
    const int tgt = target_[site + 1];
    int cr, cc;
    qr_left_site(site, tgt, cbond_, cr, cc, true);
    chi_[site + 1] = tgt;  // A[site] is (2, chil, tgt); A[site+1] still has left dim chir
    update_env(true, site);
    evolve(2, site + 1, cbond_, (long)cr * cc, cr, cc, -h);
    // absorb: A[site+1] (2, chir, chi2) -> (2, tgt, chi2)
    const int chi2 = chi_[site + 2];
    GemmDesc g = gemm_desc(cr, chi2, cc);
    g.NB2 = 2;
    g.a_m = cc; g.a_k2 = 1;
    g.b_k2 = chi2; g.b_n = 1; g.b_b2 = (long)cc * chi2;
    g.c_m = chi2; g.c_n = 1; g.c_b2 = (long)cr * chi2;
    gemm<double>(g, cbond_, A_[site + 1], theta2_);
    QCA_HIP(hipMemcpyAsync(A_[site + 1], theta2_, sizeof(c128) * 2L * cr * chi2,
                           hipMemcpyDeviceToDevice, s_));
  }
}

void Engine::sweep_left_1() {
  const double h = p_.dt / 2;
  for (int site = N_ - 1; site >= 0; --site) {
    const int chil = chi_[site], chir = chi_[site + 1];
    evolve(1, site, A_[site], 2L * chil * chir, chil, chir, h);
    stats_.site_updates++;
    if (site == 0) continue;
    const int tgt = target_[site];
    int cr, cc;
    qr_right_site(site, tgt, cbond_, cr, cc, true);
    chi_[site] = tgt;  // A[site] is (2, tgt, chir); A[site-1] still has right dim chil
    update_env(false, site);
    evolve(2, site, cbond_, (long)cr * cc, cr, cc, -h);
    const int chi0 = chi_[site - 1];
    GemmDesc g = gemm_desc(2 * chi0, cc, cr);
    g.a_m = cr; g.a_k2 = 1;
    g.b_k2 = cc; g.b_n = 1;
    g.c_m = cc; g.c_n = 1;
    gemm<double>(g, A_[site - 1], cbond_, theta2_);
    QCA_HIP(hipMemcpyAsync(A_[site - 1], theta2_, sizeof(c128) * 2L * chi0 * cc,
                           hipMemcpyDeviceToDevice, s_));
  }
}

void Engine::step() {
  if (p_.algorithm == 2) {
    sweep_right_2();
    sweep_left_2();
  } else {
    if (p_.algorithm == 3) compute_targets();
    sweep_right_1();
    sweep_left_1();
  }
  sync();
}

// Adaptive 1TDVP bond targets (reference tdvp.py:212-266 with the
// exact_exponential propagator): walk the centre to the right end with reduced
// QRs, refreshing the left environments; then from the right end, evolve a
// copy of the two-site block of every bond by dt/2, take its truncation rank
// k (first i with ||s[i:]|| < eps, capped by min(max bond, len(s))) as the
// target max(1, k), and move the centre left with a reduced QR, refreshing the
// right environment the next block uses.  The state itself is only regauged.
void Engine::compute_targets() {
  int cr, cc;
  for (int site = 0; site + 1 < N_; ++site) {
    qr_left_site(site, chi_[site + 1], cbond_, cr, cc, false);
    absorb_left(site + 1, cbond_, cr, cc);
    update_env(true, site);
  }
  for (int bond = N_ - 1; bond >= 1; --bond) {
    const int s = bond - 1;
    const int chil = chi_[s], chim = chi_[s + 1], chir = chi_[s + 2];
    merge_theta(s, chil, chim, chir);
    evolve(0, s, theta_, 4L * chil * chir, chil, chir, p_.dt / 2);
    const int rows = 2 * chil, cols = 2 * chir;
    launch_theta_to_svd(theta_, svdX_, chil, chir, 0, s_);
    launch_jacobi_svd(svdX_, rows, cols, sigma_, perm_, svdres_, maxdim_[bond], p_.eps,
                      std::min(rows, cols), s_, jw_);
    QCA_HIP(hipMemcpyAsync(svdres_host_, svdres_, sizeof(SvdResult), hipMemcpyDeviceToHost, s_));
    sync();
    stats_.svd_sweeps += svdres_host_->nsweeps;
    if (!svdres_host_->converged)
      throw Error(QCA_ERR_NOT_CONVERGED_CODE, "Jacobi SVD of the adaptive bond target did not converge");
    target_[bond] = std::max(1, svdres_host_->k);
    qr_right_site(bond, chi_[bond], cbond_, cr, cc, false);
    absorb_right(bond - 1, cbond_, cr, cc);
    update_env(false, bond);
  }
}

// ---------------------------------------------------------------- measurement
// rho_i[p][q] = sum E[x][y] A[p][x][b] conj(A[q][y][b]) with the left transfer
// E_i = sum_p A_i^T E_{i-1} conj(A_i); equals the reference's centre-site rho
// (mps.py:100-140) because sites 1..N-1 are right-isometric.
void Engine::measure(c128* rho_host) {
  measure_device();
  QCA_HIP(hipMemcpyAsync(rho_host, rho_, sizeof(c128) * 4L * N_, hipMemcpyDeviceToHost, s_));
  sync();
}

void Engine::observables(double* population, double* d_population, double* entropy) {
  measure_device();
  launch_observables(rho_, N_, obs_, s_);
  std::vector<double> h(3L * N_);
  QCA_HIP(hipMemcpyAsync(h.data(), obs_, sizeof(double) * 3L * N_, hipMemcpyDeviceToHost, s_));
  sync();
  std::copy(h.begin(), h.begin() + N_, population);
  std::copy(h.begin() + N_, h.begin() + 2L * N_, d_population);
  std::copy(h.begin() + 2L * N_, h.end(), entropy);
}

void Engine::measure_device() {
  c128 one{1.0, 0.0};
  QCA_HIP(hipMemcpyAsync(measE_, &one, sizeof(one), hipMemcpyHostToDevice, s_));
  for (int site = 0; site < N_; ++site) {
    const int chil = chi_[site], chir = chi_[site + 1];
    const long lr = (long)chil * chir;
    // T[p][x][b] = sum_y E[x][y] conj(A[p][y][b])
    GemmDesc g = gemm_desc(chil, chir, chil);
    g.NB2 = 2;
    g.a_m = chil; g.a_k2 = 1;
    g.b_k2 = chir; g.b_n = 1; g.b_b2 = lr; g.conj_b = 1;
    g.c_m = chir; g.c_n = 1; g.c_b2 = lr;
    gemm<double>(g, measE_, A_[site], measT_);
    launch_rho(A_[site], measT_, lr, rho_ + 4L * site, s_);
    if (site + 1 < N_) {
      // E[b][c] = sum_{(p,x)} A[(p,x)][b] T[(p,x)][c]
      GemmDesc h = gemm_desc(chir, chir, 2 * chil);
      h.a_m = 1; h.a_k2 = chir;
      h.b_k2 = chir; h.b_n = 1;
      h.c_m = chir; h.c_n = 1;
      gemm<double>(h, A_[site], measT_, measE_);
    }
  }
}

void Engine::get_site(int i, c128* host) const {
  QCA_CHECK(i >= 0 && i < N_, "site index out of range");
  QCA_HIP(hipMemcpyAsync(host, A_[i], sizeof(c128) * 2L * chi_[i] * chi_[i + 1],
                         hipMemcpyDeviceToHost, s_));
  QCA_HIP(hipStreamSynchronize(s_));
}

}  // namespace qca
