// Device-resident TDVP engine for one MPS instance.
#pragma once

#include <type_traits>
#include <vector>

#include "contract.hpp"

namespace qca {

enum Precision { PREC_FP64 = 0, PREC_MIXED = 1 };

// Default size limit of the one-launch Krylov exponential (elements).
constexpr long FUSED_KRYLOV_DEFAULT_N = 1024;

// Jacobi rotation threshold of the mixed path's SVD, relative to the column
// norms (|g_ij| <= tol * sqrt(g_ii g_jj) leaves a pair alone).  Its complex64
// contractions carry ~1e-7 relative error; the plain path keeps eps*sqrt(rows).
constexpr double MIXED_SVD_TOL = 1e-10;

struct EngineParams {
  int ncells = 0;
  int algorithm = 2;       // 1 = 1TDVP, 2 = 2TDVP, 3 = adaptive 1TDVP
  int precision = PREC_FP64;
  int chi_max = 32;
  double dt = 0.1;         // time step (propagators use exp(-i pi/2 dt H))
  double eps = 1e-10;      // SVD truncation cutoff on ||s[k:]||
  double krylov_tol = 1e-14;
};

struct StepStats {
  long matvecs = 0;        // effective-Hamiltonian applications (all kinds)
  long matvec_2site = 0;
  long matvec_1site = 0;
  long matvec_bond = 0;
  long site_updates = 0;
  long svd_sweeps = 0;
  long lanczos_limit_hits = 0;
  long svd_calls = 0;
  double flops = 0;        // algorithmic contraction flops
  double svd_ms = 0;       // device time of the SVDs (layout change to truncation rank)
  double krylov_ms = 0;    // device time of the Krylov exponentials
};

// GEMM of the two-site merge and of the split's second factor (shared by the
// engine and the grouped stepping of many engines).
GemmDesc merge_theta_desc(int chil, int chim, int chir);
GemmDesc split_theta_desc(bool right_sweep, int chil, int chir, int k, double inv_norm);

class Group;

class Engine {
  friend class Group;

 public:
  Engine(int device, const EngineParams& p, const std::vector<const c128*>& W_host,
         const std::vector<int>& Wdl, const std::vector<int>& Wdr,
         const std::vector<const c128*>& A_host, const std::vector<int>& chis);
  ~Engine();
  Engine(const Engine&) = delete;
  // Device bytes an engine of these parameters allocates (MPS, environments,
  // sparse MPO forms, every workspace), computed by running the constructor's
  // sizing without touching a device.
  static size_t footprint(const EngineParams& p, const std::vector<const c128*>& W_host,
                          const std::vector<int>& Wdl, const std::vector<int>& Wdr,
                          const std::vector<int>& chis);
  // Device bytes this engine holds (the same sum, from its allocations).
  size_t device_bytes() const { return device_bytes_; }
  Engine& operator=(const Engine&) = delete;

  void step();                                        // one time step (both half sweeps)
  void measure(c128* rho_host);                       // 4 complex per site
  // population, rounded population, single-site entropy (N each, host)
  void observables(double* population, double* d_population, double* entropy);
  int chi(int bond) const { return chi_[bond]; }
  int ncells() const { return N_; }
  int device() const { return device_; }
  void get_site(int i, c128* host) const;             // copies (2, chi_i, chi_{i+1})
  void set_stream(hipStream_t s) { s_ = s; }
  hipStream_t stream() const { return s_; }
  // Counters since the last reset; waits for the engine's stream so that the
  // device-side totals of the one-launch exponentials are included.
  StepStats stats();
  void reset_stats();
  // Krylov blocks of at most this many elements use the one-launch
  // exponential (0: never).  Default FUSED_KRYLOV_DEFAULT_N, or the
  // QCA_FUSED_KRYLOV_MAX_N environment variable read at creation.
  long fused_max_n() const { return fused_max_n_; }
  void set_fused_max_n(long n) { fused_max_n_ = n; }

 private:
  // counting constructor behind footprint(): sizes only, no HIP call
  Engine(const EngineParams& p, const std::vector<const c128*>& W_host, const std::vector<int>& Wdl,
         const std::vector<int>& Wdr, const std::vector<int>& chis);
  void upload_sparse(WSparsePair& dst, const HostSparse& h, int nin);
  bool count_only_ = false;
  size_t device_bytes_ = 0;
  void init(const std::vector<const c128*>& W_host, const std::vector<int>& Wdl,
            const std::vector<int>& Wdr, const std::vector<const c128*>& A_host,
            const std::vector<int>& chis);
  void release_all();
  // ---------------- helpers
  template <typename T> T* Lenv(int site) const;       // L[site], site >= -1
  template <typename T> T* Renv(int site) const;       // R[site], site <= N
  template <typename T> void heff(int kind, int site, const cplx<T>* x, cplx<T>* y, int chil,
                                  int chir);
  template <typename T> void env_update_t(bool left, int site);
  template <typename T> const ContractWs<T>& ws_for() const {
    if constexpr (std::is_same<T, double>::value) return ws64_; else return ws32_;
  }
  template <typename T> void lanczos_run(int kind, int site, c128* x, long n, int chil, int chir,
                                         double delta);
  void evolve(int kind, int site, c128* x, long n, int chil, int chir, double delta);
  void update_env(bool left, int site);
  void merge_theta(int s, int chil, int chim, int chir);
  int split_theta(int s, bool right_sweep, int chil, int chir);
  void sweep_right_2();
  void sweep_left_2();
  void sweep_right_1();
  void sweep_left_1();
  void canonicalize_initial();
  void compute_targets();
  void measure_device();                               // rho_ <- per-site rho
  void qr_left_site(int site, int target, c128* Cout, int& crow, int& ccol, bool complete);
  void qr_right_site(int site, int target, c128* Cout, int& crow, int& ccol, bool complete);
  void absorb_left(int site, const c128* C, int crow, int ccol);   // A[site] = C * A[site]
  void absorb_right(int site, const c128* C, int crow, int ccol);  // A[site] = A[site] * C
  template <typename T> void gemm(const GemmDesc& d, const cplx<T>* A, const cplx<T>* B, cplx<T>* C);
  void sync();

  EngineParams p_;
  int N_;
  int device_;
  hipStream_t s_ = nullptr;
  bool own_stream_ = false;
  std::vector<int> chi_, cap_, target_, maxdim_;
  std::vector<int> Dl_, Dr_;
  int Dmax_ = 1;
  // MPO (forward/backward sparse forms per site, merged pairs for 2TDVP)
  std::vector<WSparsePair> Wf_, Wb_, W2f_;
  std::vector<void*> allocs_;
  // state
  std::vector<c128*> A_;
  std::vector<void*> L_, R_;   // c128 or c64 by precision
  c64* one32_ = nullptr;       // boundary environment ones((1,1,1)), both precisions
  c128* one64_ = nullptr;
  // workspaces
  c128 *theta_ = nullptr, *theta2_ = nullptr, *svdX_ = nullptr, *qrQ_ = nullptr, *qrTau_ = nullptr;
  c128 *cbond_ = nullptr;
  c128* qrWs_ = nullptr;        // blocked-QR workspace
  size_t qr_ws_elems_ = 0;
  void *T_ = nullptr, *U_ = nullptr, *ws_ = nullptr;
  long ws_elems_ = 0;
  ContractWs<double> ws64_;
  ContractWs<float> ws32_;
  c64 *x32_ = nullptr, *y32_ = nullptr, *a32_ = nullptr;
  c128* y64_ = nullptr;
  double* sigma_ = nullptr;
  int* perm_ = nullptr;
  SvdResult* svdres_ = nullptr;
  JacobiWork jw_;
  SvdResult* svdres_host_ = nullptr;
  LanczosCtrl* ctrl_host_ = nullptr;
  LanczosWork lw_;
  c128 *measE_ = nullptr, *measT_ = nullptr, *rho_ = nullptr;
  double* obs_ = nullptr;
  int kpred_[3] = {8, 8, 8};
  StepStats stats_;
  KrylovTally* tally_ = nullptr;               // device totals of the one-launch exponentials
  long fused_max_n_ = 0;
  template <typename T> void fused_run(int kind, int site, c128* x, long n, int chil, int chir,
                                       double delta);
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;  // phase timers
  double elapsed_ms();                         // ev0_ -> ev1_, both complete

  template <typename T> T* dmalloc(long n);
};

}  // namespace qca
