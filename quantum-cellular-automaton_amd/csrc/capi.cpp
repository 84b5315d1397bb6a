// extern "C" boundary of libqca.so (declared in include/qca.h).
#include "qca.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "group.hpp"

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return QCA_OK;
  } catch (const qca::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return QCA_ERR_INTERNAL;
  } catch (...) {
    g_last_error = "unknown error";
    return QCA_ERR_INTERNAL;
  }
}

// Growable per-device scratch for the primitive entry points.
struct Scratch {
  std::map<std::string, std::pair<void*, size_t>> bufs;
  void* get(const std::string& name, size_t bytes) {
    auto it = bufs.find(name);
    if (it != bufs.end() && it->second.second >= bytes) return it->second.first;
    if (it != bufs.end()) {
      QCA_HIP(hipDeviceSynchronize());
      QCA_HIP(hipFree(it->second.first));
      bufs.erase(it);
    }
    void* p = nullptr;
    QCA_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
    bufs[name] = {p, bytes};
    return p;
  }
  // Device copies of the sparse MPO forms passed to the primitives, keyed by
  // their defining bytes, so that repeated calls upload nothing and stay
  // stream-asynchronous.
  std::map<std::string, qca::WSparsePair> mpos;
  qca::WSparsePair& mpo(const double* W_host, int dq, int Dl, int Dr, bool backward, int nin) {
    const size_t bytes = sizeof(qca::c128) * (size_t)dq * dq * Dl * Dr;
    std::string key(reinterpret_cast<const char*>(W_host), bytes);
    key += std::to_string(dq) + "/" + std::to_string(Dl) + "/" + std::to_string(Dr) + "/" +
           std::to_string((int)backward) + "/" + std::to_string(nin);
    auto it = mpos.find(key);
    if (it != mpos.end()) return it->second;
    if (mpos.size() >= 64) release_mpos();
    qca::WSparsePair wp;
    try {
      wp.upload(qca::build_sparse(reinterpret_cast<const qca::c128*>(W_host), dq, Dl, Dr, backward),
                nin);
    } catch (...) {
      wp.release();
      throw;
    }
    return mpos.emplace(std::move(key), std::move(wp)).first->second;
  }
  void release_mpos() {
    QCA_HIP(hipDeviceSynchronize());
    for (auto& kv : mpos) kv.second.release();
    mpos.clear();
  }
  void release() {
    for (auto& kv : bufs) (void)hipFree(kv.second.first);
    bufs.clear();
    for (auto& kv : mpos) kv.second.release();
    mpos.clear();
  }
};
std::map<int, Scratch> g_scratch;

Scratch& prim_scratch() {
  int dev = 0;
  QCA_HIP(hipGetDevice(&dev));
  return g_scratch[dev];
}

template <typename T>
T* sbuf(const char* name, long n) {
  return reinterpret_cast<T*>(prim_scratch().get(name, sizeof(T) * (size_t)std::max<long>(n, 1)));
}

template <typename T>
qca::ContractWs<T> contract_ws(long tsize, long outsize) {
  qca::ContractWs<T> w;
  w.T1 = sbuf<qca::cplx<T>>("T1", tsize);
  w.U = sbuf<qca::cplx<T>>("U", tsize);
  w.splitk_elems = 16 * std::max(outsize, tsize / 4);
  w.splitk = sbuf<qca::cplx<T>>("splitk", w.splitk_elems);
  return w;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

qca::LanczosWork lanczos_ws(long n, qca::LanczosCtrl** ctrl_host, int max_m = qca::LZ_MAXK) {
  qca::LanczosWork lw;
  lw.max_n = n;
  lw.max_m = max_m;
  lw.max_blocks = qca::lanczos_nblk(n);
  lw.ctrl = sbuf<qca::LanczosCtrl>("lz_ctrl", 1);
  lw.V = sbuf<qca::c128>("lz_V", (long)max_m * n);
  lw.w64 = sbuf<qca::c128>("lz_w", n);
  lw.part = sbuf<qca::c128>("lz_part", (long)lw.max_blocks * qca::LZ_MAXK);
  lw.part2 = sbuf<qca::c128>("lz_part2", (long)lw.max_blocks * qca::LZ_MAXK);
  lw.part_norm = sbuf<double>("lz_pnorm", lw.max_blocks);
  lw.coef1 = sbuf<qca::c128>("lz_c1", qca::LZ_MAXK);
  lw.coef2 = sbuf<qca::c128>("lz_c2", qca::LZ_MAXK);
  // ordinary host memory: the control block is read once per exponential
  // after a stream synchronisation, and no HIP resource outlives the calls
  static thread_local qca::LanczosCtrl host;
  *ctrl_host = &host;
  return lw;
}

struct KrylovOutcome {
  int m = 0;           // Krylov dimension used
  bool hit_limit = false;
};

// Runs the Krylov loop: mv(j, V_j) must enqueue w = H V_j and the iteration
// tail.  Iterations j = 0 .. max_m-2 may run; the device closes the
// exponential at the latest at dimension max_m-1 and then flags a limit hit
// when its error estimate is still above the tolerance.  x is overwritten with
// the result unless the limit was hit and keep_on_limit is false.
template <typename MV>
KrylovOutcome krylov_loop(qca::c128* x, long n, qca::LanczosWork& lw, qca::LanczosCtrl* host,
                          MV&& mv, hipStream_t s, bool keep_on_limit = true) {
  qca::lanczos_begin(x, n, lw, s);
  int j = 0, chunk = 8;
  while (true) {
    for (int c = 0; c < chunk && j < lw.max_m - 1; ++c, ++j) mv(j, lw.V + (long)j * n);
    QCA_HIP(hipMemcpyAsync(host, lw.ctrl, sizeof(qca::LanczosCtrl), hipMemcpyDeviceToHost, s));
    QCA_HIP(hipStreamSynchronize(s));
    if (host->done) break;
    if (j >= lw.max_m - 1)
      throw qca::Error(qca::QCA_ERR_NOT_CONVERGED_CODE,
                       "Krylov exponential did not converge (non-finite input?)");
    chunk = 4;
  }
  KrylovOutcome r;
  r.m = host->m;
  r.hit_limit = host->hit_limit != 0;
  if (r.hit_limit && !keep_on_limit) return r;
  qca::lanczos_finish(n, x, lw, s);
  QCA_HIP(hipStreamSynchronize(s));
  return r;
}

}  // namespace

struct qca_engine {
  qca::Engine* eng = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double gpu_ms = 0;
  // Set when a time step fails part-way: the tensors and environments are then
  // mutually inconsistent, so every later call on the engine is refused.
  std::string failed;
};

namespace {

void require_usable(const qca_engine* e) {
  QCA_CHECK(e && e->eng, "null engine");
  if (!e->failed.empty())
    throw qca::Error(QCA_ERR_ARG, "engine unusable after an earlier failed step: " + e->failed);
}

// One time step of `e` on its stream, with its GPU time added to e->gpu_ms.
// A failure marks the engine as failed.
void engine_one_step(qca_engine* e) {
  require_usable(e);
  try {
    QCA_HIP(hipEventRecord(e->ev0, e->eng->stream()));
    e->eng->step();
    QCA_HIP(hipEventRecord(e->ev1, e->eng->stream()));
    QCA_HIP(hipEventSynchronize(e->ev1));
  } catch (const std::exception& ex) {
    e->failed = ex.what();
    throw;
  }
  float ms = 0;
  QCA_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
  e->gpu_ms += ms;
}

}  // namespace

extern "C" {

int qca_version(void) { return 10000; }

const char* qca_last_error(void) { return g_last_error.c_str(); }

int qca_device_count(int* count) {
  return guarded([&] {
    QCA_CHECK(count != nullptr, "null count");
    QCA_HIP(hipGetDeviceCount(count));
  });
}

namespace {
qca::EngineParams to_params(const qca_engine_params* params) {
  const int n = params->ncells;
  QCA_CHECK(n >= 2 && n <= 4096, "ncells out of range");
  qca::EngineParams p;
  p.ncells = n;
  p.algorithm = params->algorithm;
  p.precision = params->precision;
  p.chi_max = params->chi_max;
  p.dt = params->dt;
  p.eps = params->eps;
  p.krylov_tol = params->krylov_tol > 0 ? params->krylov_tol : 1e-14;
  return p;
}
}  // namespace

int qca_engine_footprint(const qca_engine_params* params, const double* const* W, const int* Wdl,
                         const int* Wdr, const int* chis, size_t* bytes) {
  return guarded([&] {
    QCA_CHECK(params && W && Wdl && Wdr && chis && bytes, "null argument");
    const qca::EngineParams p = to_params(params);
    const int n = p.ncells;
    std::vector<const qca::c128*> Wv(n);
    for (int i = 0; i < n; ++i) {
      QCA_CHECK(W[i], "null tensor pointer");
      Wv[i] = reinterpret_cast<const qca::c128*>(W[i]);
    }
    std::vector<int> dl(Wdl, Wdl + n), dr(Wdr, Wdr + n), ch(chis, chis + n + 1);
    *bytes = qca::Engine::footprint(p, Wv, dl, dr, ch);
  });
}

int qca_engine_device_bytes(qca_engine* e, size_t* bytes) {
  return guarded([&] {
    QCA_CHECK(e && e->eng && bytes, "null argument");
    *bytes = e->eng->device_bytes();
  });
}

int qca_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    QCA_CHECK(free_bytes && total_bytes, "null argument");
    int cur = 0;
    QCA_HIP(hipGetDevice(&cur));
    QCA_HIP(hipSetDevice(device));
    const hipError_t rc = hipMemGetInfo(free_bytes, total_bytes);
    (void)hipSetDevice(cur);
    QCA_HIP(rc);
  });
}

int qca_engine_create(const qca_engine_params* params, const double* const* W, const int* Wdl,
                      const int* Wdr, const double* const* A, const int* chis, qca_engine** out) {
  return guarded([&] {
    QCA_CHECK(params && W && Wdl && Wdr && A && chis && out, "null argument");
    const qca::EngineParams p = to_params(params);
    const int n = p.ncells;
    std::vector<const qca::c128*> Wv(n), Av(n);
    for (int i = 0; i < n; ++i) {
      QCA_CHECK(W[i] && A[i], "null tensor pointer");
      Wv[i] = reinterpret_cast<const qca::c128*>(W[i]);
      Av[i] = reinterpret_cast<const qca::c128*>(A[i]);
    }
    std::vector<int> dl(Wdl, Wdl + n), dr(Wdr, Wdr + n), ch(chis, chis + n + 1);
    auto* h = new qca_engine();
    try {
      h->eng = new qca::Engine(params->device, p, Wv, dl, dr, Av, ch);
      QCA_HIP(hipEventCreate(&h->ev0));
      QCA_HIP(hipEventCreate(&h->ev1));
    } catch (...) {
      delete h->eng;
      delete h;
      throw;
    }
    *out = h;
  });
}

int qca_engine_destroy(qca_engine* e) {
  return guarded([&] {
    if (!e) return;
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    delete e->eng;
    delete e;
  });
}

int qca_engine_step(qca_engine* e, int nsteps) {
  return guarded([&] {
    QCA_CHECK(e && e->eng, "null engine");
    QCA_CHECK(nsteps >= 0, "negative step count");
    QCA_HIP(hipSetDevice(e->eng->device()));
    for (int i = 0; i < nsteps; ++i) engine_one_step(e);
  });
}

int qca_batch_step(qca_engine* const* engines, int n, int nsteps, int nthreads, int* failed) {
  std::string first_error;
  int first_code = QCA_OK;
  const int rc = guarded([&] {
    QCA_CHECK(engines != nullptr || n == 0, "null engine list");
    QCA_CHECK(n >= 0 && nsteps >= 0, "negative count");
    QCA_CHECK(nthreads >= 1 && nthreads <= 256, "nthreads out of range (1..256)");
    for (int i = 0; i < n; ++i) QCA_CHECK(engines[i] && engines[i]->eng, "null engine in batch");
    if (failed) *failed = -1;
    if (n == 0 || nsteps == 0) return;
    // round-robin queue of (engine, steps left): a worker takes the front
    // engine, advances it one time step on its own stream and requeues it
    std::mutex mu;
    std::deque<std::pair<int, int>> queue;
    for (int i = 0; i < n; ++i) queue.emplace_back(i, nsteps);
    auto worker = [&] {
      while (true) {
        std::pair<int, int> job;
        {
          std::lock_guard<std::mutex> lock(mu);
          if (queue.empty() || first_code != QCA_OK) return;
          job = queue.front();
          queue.pop_front();
        }
        qca_engine* e = engines[job.first];
        const int code = guarded([&] {
          QCA_HIP(hipSetDevice(e->eng->device()));
          engine_one_step(e);
        });
        std::lock_guard<std::mutex> lock(mu);
        if (code != QCA_OK) {
          if (first_code == QCA_OK) {
            first_code = code;
            first_error = "engine " + std::to_string(job.first) + ": " + g_last_error;
            if (failed) *failed = job.first;
          }
          return;
        }
        if (job.second > 1) queue.emplace_back(job.first, job.second - 1);
      }
    };
    const int nt = std::min(nthreads, n);
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
  });
  if (rc != QCA_OK) return rc;
  if (first_code != QCA_OK) {
    g_last_error = first_error;
    return first_code;
  }
  return QCA_OK;
}

struct qca_group {
  qca::Group* grp = nullptr;
  std::vector<qca_engine*> engines;
};

int qca_group_create(qca_engine* const* engines, int n, qca_group** out) {
  return guarded([&] {
    QCA_CHECK(engines && out && n >= 1, "null argument or empty group");
    QCA_CHECK(n <= 4096, "at most 4096 engines per group");
    std::vector<qca::Engine*> v;
    for (int i = 0; i < n; ++i) {
      require_usable(engines[i]);
      v.push_back(engines[i]->eng);
    }
    auto* g = new qca_group();
    try {
      g->grp = new qca::Group(v);
    } catch (...) {
      delete g;
      throw;
    }
    g->engines.assign(engines, engines + n);
    *out = g;
  });
}

int qca_group_step(qca_group* g, int nsteps) {
  return guarded([&] {
    QCA_CHECK(g && g->grp, "null group");
    QCA_CHECK(nsteps >= 0, "negative step count");
    for (qca_engine* e : g->engines) require_usable(e);
    try {
      g->grp->step(nsteps);
    } catch (const std::exception& ex) {
      // the instances stopped part-way through a sweep
      for (qca_engine* e : g->engines) e->failed = ex.what();
      throw;
    }
  });
}

int qca_group_get_stats(qca_group* g, qca_group_stats* out) {
  return guarded([&] {
    QCA_CHECK(g && g->grp && out, "null argument");
    const qca::GroupStats& s = g->grp->stats();
    out->steps = s.steps;
    out->krylov_iterations = s.launches_krylov;
    out->host_syncs = s.host_syncs;
    out->krylov_ms = s.krylov_ms;
    out->svd_ms = s.svd_ms;
    out->total_ms = s.total_ms;
  });
}

int qca_group_reset_stats(qca_group* g) {
  return guarded([&] {
    QCA_CHECK(g && g->grp, "null group");
    g->grp->reset_stats();
  });
}

int qca_group_destroy(qca_group* g) {
  return guarded([&] {
    if (!g) return;
    delete g->grp;
    delete g;
  });
}

int qca_engine_measure(qca_engine* e, double* rho_out) {
  return guarded([&] {
    QCA_CHECK(e && e->eng && rho_out, "null argument");
    require_usable(e);
    e->eng->measure(reinterpret_cast<qca::c128*>(rho_out));
  });
}

int qca_engine_observables(qca_engine* e, double* population, double* d_population,
                           double* single_site_entropy, double* bond_dims) {
  return guarded([&] {
    QCA_CHECK(e && e->eng && population && d_population && single_site_entropy && bond_dims,
              "null argument");
    require_usable(e);
    QCA_HIP(hipSetDevice(e->eng->device()));
    const int n = e->eng->ncells();
    for (int i = 0; i <= n; ++i) bond_dims[i] = e->eng->chi(i);
    e->eng->observables(population, d_population, single_site_entropy);
  });
}

int qca_engine_bond_dims(qca_engine* e, int* out) {
  return guarded([&] {
    QCA_CHECK(e && e->eng && out, "null argument");
    for (int i = 0; i <= e->eng->ncells(); ++i) out[i] = e->eng->chi(i);
  });
}

int qca_engine_get_site(qca_engine* e, int site, double* out) {
  return guarded([&] {
    QCA_CHECK(e && e->eng && out, "null argument");
    require_usable(e);
    e->eng->get_site(site, reinterpret_cast<qca::c128*>(out));
  });
}

int qca_engine_stats(qca_engine* e, qca_stats* out) {
  return guarded([&] {
    QCA_CHECK(e && e->eng && out, "null argument");
    const qca::StepStats s = e->eng->stats();
    out->matvecs = s.matvecs;
    out->matvec_2site = s.matvec_2site;
    out->matvec_1site = s.matvec_1site;
    out->matvec_bond = s.matvec_bond;
    out->site_updates = s.site_updates;
    out->svd_sweeps = s.svd_sweeps;
    out->lanczos_limit_hits = s.lanczos_limit_hits;
    out->svd_calls = s.svd_calls;
    out->flops = s.flops;
    out->svd_ms = s.svd_ms;
    out->krylov_ms = s.krylov_ms;
  });
}

int qca_engine_reset_stats(qca_engine* e) {
  return guarded([&] {
    QCA_CHECK(e && e->eng, "null engine");
    e->eng->reset_stats();
    e->gpu_ms = 0;
  });
}

int qca_engine_gpu_ms(qca_engine* e, double* ms) {
  return guarded([&] {
    QCA_CHECK(e && ms, "null argument");
    *ms = e->gpu_ms;
  });
}

int qca_engine_set_fused_krylov_limit(qca_engine* e, long max_n) {
  return guarded([&] {
    QCA_CHECK(e, "null engine");
    QCA_CHECK(max_n >= 0 && max_n <= qca::FUSED_KRYLOV_MAX_N, "limit must be in 0..4096");
    e->eng->set_fused_max_n(max_n);
  });
}

// ------------------------------------------------------------------ primitives

int qca_heff_apply(int precision, int kind, const void* x, void* y, const void* L, const void* R,
                   const double* W_host, int Din, int Dout, int chil, int chir, void* stream) {
  return guarded([&] {
    QCA_CHECK(precision == 0 || precision == 1, "precision must be 0 or 1");
    QCA_CHECK(kind >= 0 && kind <= 2, "kind must be 0, 1 or 2");
    QCA_CHECK(x && y && L && R, "null tensor");
    QCA_CHECK(chil >= 1 && chir >= 1 && chil <= 512 && chir <= 512, "bond out of range");
    QCA_CHECK(Din >= 1 && Dout >= 1 && Din <= 16 && Dout <= 16, "MPO bond out of range");
    QCA_CHECK(kind == 2 ? Din == Dout : W_host != nullptr, "bad MPO arguments");
    hipStream_t s = as_stream(stream);
    const int dq = kind == 0 ? 4 : (kind == 1 ? 2 : 1);
    const long tsize = (long)std::max(dq, 1) * std::max(Din, Dout) * chil * chir;
    qca::WSparsePair* wp = kind == 2 ? nullptr : &prim_scratch().mpo(W_host, dq, Din, Dout, false, dq * Din);
    if (precision == 0) {
      auto ws = contract_ws<double>(tsize, (long)dq * chil * chir);
      qca::heff_apply<double>(kind, (const qca::c128*)x, (qca::c128*)y, (const qca::c128*)L,
                              (const qca::c128*)R, wp ? &wp->d64 : nullptr, Din, Dout, chil, chir,
                              ws, nullptr, s);
    } else {
      auto ws = contract_ws<float>(tsize, (long)dq * chil * chir);
      qca::heff_apply<float>(kind, (const qca::c64*)x, (qca::c64*)y, (const qca::c64*)L,
                             (const qca::c64*)R, wp ? &wp->d32 : nullptr, Din, Dout, chil, chir, ws,
                             nullptr, s);
    }
  });
}

int qca_env_update(int precision, int side, const void* A, const void* Eprev, void* Enew,
                   const double* W_host, int Dl, int Dr, int chil, int chir, void* stream) {
  return guarded([&] {
    QCA_CHECK(precision == 0 || precision == 1, "precision must be 0 or 1");
    QCA_CHECK(side == 0 || side == 1, "side must be 0 (left) or 1 (right)");
    QCA_CHECK(A && Eprev && Enew && W_host, "null tensor");
    QCA_CHECK(chil >= 1 && chir >= 1 && chil <= 512 && chir <= 512, "bond out of range");
    QCA_CHECK(Dl >= 1 && Dr >= 1 && Dl <= 16 && Dr <= 16, "MPO bond out of range");
    hipStream_t s = as_stream(stream);
    const bool left = side == 0;
    qca::WSparsePair& wp = prim_scratch().mpo(W_host, 2, Dl, Dr, !left, left ? 2 * Dl : 2 * Dr);
    const int Din = left ? Dl : Dr, Dout = left ? Dr : Dl;
    const long tsize = 2L * std::max(Dl, Dr) * chil * chir;
    const long osize = (long)std::max(chil, chir) * std::max(Dl, Dr) * std::max(chil, chir);
    if (precision == 0) {
      auto ws = contract_ws<double>(tsize, osize);
      qca::env_update<double>(left, (const qca::c128*)A, (const qca::c128*)Eprev, (qca::c128*)Enew,
                              &wp.d64, Din, Dout, chil, chir, ws, s);
    } else {
      auto ws = contract_ws<float>(tsize, osize);
      qca::env_update<float>(left, (const qca::c64*)A, (const qca::c64*)Eprev, (qca::c64*)Enew,
                             &wp.d32, Din, Dout, chil, chir, ws, s);
    }
  });
}

int qca_krylov_expm(int precision, int kind, void* x, const void* L, const void* R,
                    const double* W_host, int Din, int Dout, int chil, int chir, double delta,
                    double tol, int* iters, void* stream) {
  return guarded([&] {
    QCA_CHECK(precision == 0 || precision == 1, "precision must be 0 or 1");
    QCA_CHECK(kind >= 0 && kind <= 2, "kind must be 0, 1 or 2");
    QCA_CHECK(x && L && R, "null tensor");
    QCA_CHECK(chil >= 1 && chir >= 1 && chil <= 512 && chir <= 512, "bond out of range");
    QCA_CHECK(Din >= 1 && Dout >= 1 && Din <= 16 && Dout <= 16, "MPO bond out of range");
    QCA_CHECK(kind == 2 ? Din == Dout : W_host != nullptr, "bad MPO arguments");
    hipStream_t s = as_stream(stream);
    const int dq = kind == 0 ? 4 : (kind == 1 ? 2 : 1);
    const long n = (long)dq * chil * chir;
    const long tsize = (long)dq * std::max(Din, Dout) * chil * chir;
    qca::WSparsePair* wp = kind == 2 ? nullptr : &prim_scratch().mpo(W_host, dq, Din, Dout, false, dq * Din);
    qca::LanczosCtrl* host = nullptr;
    qca::LanczosWork lw = lanczos_ws(n, &host);
    const double tau = 1.5707963267948966 * delta;
    const double tl = tol > 0 ? tol : 1e-14;
    {
      int m = 0;
      if (precision == 0) {
        auto ws = contract_ws<double>(tsize, n);
        qca::c128* y = sbuf<qca::c128>("kr_y", n);
        m = krylov_loop((qca::c128*)x, n, lw, host, [&](int j, const qca::c128* vj) {
          qca::heff_apply<double>(kind, vj, y, (const qca::c128*)L, (const qca::c128*)R,
                                  wp ? &wp->d64 : nullptr, Din, Dout, chil, chir, ws,
                                  qca::lanczos_done_flag(lw), s);
          qca::lanczos_iteration_tail<double>(y, n, j, tau, tl, lw, s);
        }, s).m;
      } else {
        auto ws = contract_ws<float>(tsize, n);
        qca::c64* x32 = sbuf<qca::c64>("kr_x32", n);
        qca::c64* y = sbuf<qca::c64>("kr_y32", n);
        m = krylov_loop((qca::c128*)x, n, lw, host, [&](int j, const qca::c128* vj) {
          qca::launch_convert_c128_c64(vj, x32, n, s, qca::lanczos_done_flag(lw));
          qca::heff_apply<float>(kind, x32, y, (const qca::c64*)L, (const qca::c64*)R,
                                 wp ? &wp->d32 : nullptr, Din, Dout, chil, chir, ws,
                                 qca::lanczos_done_flag(lw), s);
          qca::lanczos_iteration_tail<float>(y, n, j, tau, tl, lw, s);
        }, s).m;
      }
      if (iters) *iters = m;
    }
  });
}

int qca_svd_split(const void* theta, void* Al, void* Ar, int chil, int chir, int chi_max,
                  double eps, int right_sweep, int* k, double* sigma_host, void* stream) {
  return guarded([&] {
    QCA_CHECK(theta && Al && Ar && k, "null argument");
    QCA_CHECK(chil >= 1 && chir >= 1 && chil <= 256 && chir <= 256, "bond out of range");
    QCA_CHECK(chi_max >= 1, "chi_max must be positive");
    hipStream_t s = as_stream(stream);
    const int rows = right_sweep ? 2 * chil : 2 * chir;
    const int cols = right_sweep ? 2 * chir : 2 * chil;
    const int rank_cap = std::min(2 * chil, 2 * chir);
    const qca::c128* th = (const qca::c128*)theta;
    qca::c128* X = sbuf<qca::c128>("svd_X", 4L * chil * chir);
    double* sigma = sbuf<double>("svd_sigma", cols);
    int* perm = sbuf<int>("svd_perm", cols);
    qca::SvdResult* res = sbuf<qca::SvdResult>("svd_res", 1);
    qca::launch_theta_to_svd(th, X, chil, chir, right_sweep ? 0 : 1, s);
    // keeps the sweep count of the previous call; the convergence words are
    // read into ordinary host memory (no HIP resource outlives the call)
    static thread_local qca::JacobiWork jw;
    static thread_local long long jw_words[8];
    jw.ws_bytes = qca::jacobi_workspace_bytes(rows, cols);
    jw.ws = sbuf<unsigned char>("svd_ws", (long)jw.ws_bytes);
    jw.pre_elems = qca::jacobi_precond_elems(rows, cols);
    jw.pre = sbuf<qca::c128>("svd_pre", (long)jw.pre_elems);
    jw.host = jw_words;
    qca::launch_jacobi_svd(X, rows, cols, sigma, perm, res, chi_max, eps, rank_cap, s, jw);
    qca::SvdResult hres;
    QCA_HIP(hipMemcpyAsync(&hres, res, sizeof(hres), hipMemcpyDeviceToHost, s));
    QCA_HIP(hipStreamSynchronize(s));
    const int kk = hres.k;
    if (!hres.converged)
      throw qca::Error(qca::QCA_ERR_NOT_CONVERGED_CODE, "Jacobi SVD did not converge");
    QCA_CHECK(kk >= 1, "empty truncation");
    qca::ContractWs<double> ws = contract_ws<double>(4L * chil * chir, 4L * chil * chir);
    const double inv = 1.0 / hres.norm_kept;
    if (right_sweep) {
      qca::launch_gather_isometry(X, perm, sigma, (qca::c128*)Al, rows, chil, kk, 0, s);
      qca::GemmDesc g = qca::gemm_desc(kk, chir, 2 * chil);
      g.K2 = chil;
      g.NB2 = 2;
      g.a_m = 1; g.a_k1 = (long)chil * kk; g.a_k2 = kk; g.conj_a = 1;
      g.b_k1 = 2L * chil * chir; g.b_k2 = chir; g.b_n = 1; g.b_b2 = (long)chil * chir;
      g.c_m = chir; g.c_n = 1; g.c_b2 = (long)kk * chir;
      g.alpha_re = inv;
      qca::gemm_t<double>(g, (const qca::c128*)Al, th, (qca::c128*)Ar, s, ws);
    } else {
      qca::launch_gather_isometry(X, perm, sigma, (qca::c128*)Ar, rows, chir, kk, 1, s);
      qca::GemmDesc g = qca::gemm_desc(chil, kk, 2 * chir);
      g.K2 = chir;
      g.NB2 = 2;
      g.a_m = chir; g.a_k1 = (long)chil * chir; g.a_k2 = 1; g.a_b2 = 2L * chil * chir;
      g.b_k1 = (long)kk * chir; g.b_k2 = 1; g.b_n = chir; g.conj_b = 1;
      g.c_m = kk; g.c_n = 1; g.c_b2 = (long)chil * kk;
      g.alpha_re = inv;
      qca::gemm_t<double>(g, th, (const qca::c128*)Ar, (qca::c128*)Al, s, ws);
    }
    if (sigma_host)
      QCA_HIP(hipMemcpyAsync(sigma_host, sigma, sizeof(double) * rank_cap, hipMemcpyDeviceToHost, s));
    QCA_HIP(hipStreamSynchronize(s));
    *k = kk;
  });
}

int qca_qr_site(int side, const void* A, int chil, int chir, int target, void* Anew, void* C,
                void* stream) {
  return guarded([&] {
    QCA_CHECK(side == 0 || side == 1, "side must be 0 or 1");
    QCA_CHECK(A && Anew && C, "null tensor");
    QCA_CHECK(chil >= 1 && chir >= 1 && chil <= 512 && chir <= 512 && target >= 1, "bad shape");
    hipStream_t s = as_stream(stream);
    const int rows = side == 0 ? 2 * chil : 2 * chir;
    const int cols = side == 0 ? chir : chil;
    qca::c128* X = sbuf<qca::c128>("qr_X", (long)rows * cols);
    qca::c128* Q = sbuf<qca::c128>("qr_Q", (long)rows * std::min(target, rows));
    qca::c128* tau = sbuf<qca::c128>("qr_tau", std::min(rows, cols));
    const int qcols = std::min(target, rows);
    if (side == 0) qca::launch_to_colmajor((const qca::c128*)A, X, rows, cols, s);
    else qca::launch_site_to_rightqr((const qca::c128*)A, X, chil, chir, s);
    const size_t qws = qca::qr_workspace_elems(rows, cols, std::max(qcols, 1));
    qca::launch_qr(X, rows, cols, tau, Q, qcols, rows <= 512 ? sbuf<qca::c128>("qr_ws", (long)qws) : nullptr,
                   qws, s);
    if (side == 0) qca::launch_qr_gather_left(Q, X, (qca::c128*)Anew, (qca::c128*)C, rows, cols, qcols, target, s);
    else qca::launch_qr_gather_right(Q, X, (qca::c128*)Anew, (qca::c128*)C, chir, cols, qcols, target, s);
    QCA_HIP(hipStreamSynchronize(s));
  });
}

int qca_exact_apply_h(int precision, const void* psi, void* out, int ncells, int distance, int lo,
                      int hi, void* stream) {
  return guarded([&] {
    QCA_CHECK(precision == 0 || precision == 1, "precision must be 0 or 1");
    QCA_CHECK(psi && out, "null vector");
    QCA_CHECK(ncells >= 1 && ncells <= 34, "ncells out of range for the state vector");
    QCA_CHECK(distance >= 1, "distance must be positive");
    hipStream_t s = as_stream(stream);
    if (precision == 0)
      qca::launch_exact_apply<double>((const qca::c128*)psi, (qca::c128*)out, ncells, distance, lo, hi, s);
    else
      qca::launch_exact_apply<float>((const qca::c64*)psi, (qca::c64*)out, ncells, distance, lo, hi, s);
    QCA_HIP(hipStreamSynchronize(s));
  });
}

int qca_exact_expm(void* psi, int ncells, int distance, int lo, int hi, double dt, double tol,
                   int* iters, void* stream) {
  return guarded([&] {
    QCA_CHECK(psi, "null vector");
    QCA_CHECK(ncells >= 2 && ncells <= 28, "ncells out of range for the Krylov exponential (2..28)");
    QCA_CHECK(distance >= 1, "distance must be positive");
    QCA_CHECK(std::isfinite(dt), "time step must be finite");
    hipStream_t s = as_stream(stream);
    const long n = 1L << ncells;
    // Krylov basis capped at 16 GiB (at least 24 vectors)
    const int max_m = (int)std::max<long>(24, std::min<long>(qca::LZ_MAXK, (16L << 30) / (16L * n)));
    qca::LanczosCtrl* host = nullptr;
    qca::LanczosWork lw = lanczos_ws(n, &host, max_m);
    qca::c128* y = sbuf<qca::c128>("ex_y", n);
    const double tl = tol > 0 ? tol : 1e-14;
    // The time step is split into 2^level equal substeps.  A short basis
    // starts from about one substep per 8 units of |tau| * N (each row of H
    // has at most N unit entries, so ||H|| <= N); a substep whose exponential
    // reaches the basis limit leaves psi untouched and is retried at half the
    // length.  Progress is counted in units of dt / 2^level, so every substep
    // is an exact binary fraction of dt.
    const double tau_full = 1.5707963267948966 * dt;
    int level = 0;
    if (max_m < 48) {
      const double est = std::fabs(tau_full) * ncells / 8.0;
      while (level < 20 && (double)(1L << level) < est) ++level;
    }
    long done_units = 0;
    int total = 0;
    while (done_units < (1L << level)) {
      const double tau = tau_full / (double)(1L << level);
      const KrylovOutcome r = krylov_loop((qca::c128*)psi, n, lw, host, [&](int j, const qca::c128* vj) {
        qca::launch_exact_apply<double>(vj, y, ncells, distance, lo, hi, s);
        qca::lanczos_iteration_tail<double>(y, n, j, tau, tl, lw, s);
      }, s, /*keep_on_limit=*/false);
      total += r.m;
      if (r.hit_limit) {
        QCA_CHECK(level < 30, "Krylov exponential of the state vector did not converge");
        ++level;
        done_units *= 2;
        continue;
      }
      ++done_units;
    }
    if (iters) *iters = total;
  });
}

int qca_exact_rho(const void* psi, int ncells, double* rho_host, void* stream) {
  return guarded([&] {
    QCA_CHECK(psi && rho_host, "null argument");
    QCA_CHECK(ncells >= 1 && ncells <= 34, "ncells out of range for the state vector");
    hipStream_t s = as_stream(stream);
    double* part = sbuf<double>("rho_part", 4L * ncells * qca::exact_rho_blocks(ncells));
    qca::c128* rho = sbuf<qca::c128>("rho_out", 4L * ncells);
    qca::launch_exact_rho((const qca::c128*)psi, ncells, part, rho, s);
    QCA_HIP(hipMemcpyAsync(rho_host, rho, sizeof(qca::c128) * 4L * ncells, hipMemcpyDeviceToHost, s));
    QCA_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

// ------------------------------------------------------------------ no-allocation variant
// Static device scratch of qca_heff_apply_noalloc: the sparse MPO form and the
// contraction workspace (both intermediates and the split-K partials) of one
// application, in the layout and sizes qca_heff_apply allocates, so both
// entry points run the same GEMM plans.  One copy per device, held by the
// loaded code object until the process ends.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_noalloc_mu;
std::map<int, unsigned char*> g_noalloc_ptr;   // device -> address of `scratch` there

// Device address of `scratch` on the current device (looked up once per
// device with that device current; the array's name is only a host stand-in).
unsigned char* noalloc_scratch() {
  int dev = 0;
  QCA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  auto it = g_noalloc_ptr.find(dev);
  if (it != g_noalloc_ptr.end()) return it->second;
  void* p = nullptr;
  QCA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  QCA_CHECK(p != nullptr, "no device address for the static scratch");
  g_noalloc_ptr[dev] = reinterpret_cast<unsigned char*>(p);
  return reinterpret_cast<unsigned char*>(p);
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

template <typename T>
void heff_noalloc(int kind, const void* x, void* y, const void* L, const void* R, const double* W_host,
                  int Din, int Dout, int chil, int chir, hipStream_t s) {
  const int dq = kind == 0 ? 4 : (kind == 1 ? 2 : 1);
  const long tsize = (long)std::max(dq, 1) * std::max(Din, Dout) * chil * chir;
  const long outsize = (long)dq * chil * chir;
  const long splitk_elems = 16 * std::max(outsize, tsize / 4);   // as contract_ws
  qca::HostSparse hs;
  size_t mpo_bytes = 0;
  if (kind != 2) {
    hs = qca::build_sparse(reinterpret_cast<const qca::c128*>(W_host), dq, Din, Dout, false);
    mpo_bytes = align256(sizeof(int) * hs.rowptr.size()) + 2 * align256(sizeof(int) * std::max<size_t>(1, hs.col.size())) +
                align256(sizeof(qca::cplx<T>) * std::max<size_t>(1, hs.val.size()));
  }
  // the effective-Hamiltonian chain needs T1 and the split-K partials only
  const size_t need = mpo_bytes + align256(sizeof(qca::cplx<T>) * tsize) +
                      align256(sizeof(qca::cplx<T>) * splitk_elems);
  if (need > kNoAllocScratchBytes)
    throw qca::Error(QCA_ERR_ARG, "qca_heff_apply_noalloc: this application needs " + std::to_string(need) +
                                      " bytes of workspace, above the 64 MiB static scratch");
  unsigned char* base = noalloc_scratch();
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base + off;
    off += align256(bytes);
    return p;
  };
  qca::WSparseDev<T> w;
  if (kind != 2) {
    w.rowptr = reinterpret_cast<int*>(take(sizeof(int) * hs.rowptr.size()));
    w.col = reinterpret_cast<int*>(take(sizeof(int) * std::max<size_t>(1, hs.col.size())));
    w.outv = reinterpret_cast<int*>(take(sizeof(int) * std::max<size_t>(1, hs.outv.size())));
    w.val = reinterpret_cast<qca::cplx<T>*>(take(sizeof(qca::cplx<T>) * std::max<size_t>(1, hs.val.size())));
    w.nout = (int)hs.rowptr.size() - 1;
    w.nin = dq * Din;
    w.nnz = (int)hs.col.size();
    w.group_rows = hs.group_rows;
    w.group_max = hs.group_max;
    qca::set_group_starts(w, hs.rowptr, hs.group_rows);
    std::vector<qca::cplx<T>> val(hs.val.size());
    for (size_t z = 0; z < val.size(); ++z) val[z] = qca::cplx<T>{(T)hs.val[z].re, (T)hs.val[z].im};
    // pageable sources: the copies are complete when the calls return
    QCA_HIP(hipMemcpyAsync(w.rowptr, hs.rowptr.data(), sizeof(int) * hs.rowptr.size(), hipMemcpyHostToDevice, s));
    if (!hs.col.empty()) {
      QCA_HIP(hipMemcpyAsync(w.col, hs.col.data(), sizeof(int) * hs.col.size(), hipMemcpyHostToDevice, s));
      QCA_HIP(hipMemcpyAsync(w.outv, hs.outv.data(), sizeof(int) * hs.outv.size(), hipMemcpyHostToDevice, s));
      QCA_HIP(hipMemcpyAsync(w.val, val.data(), sizeof(qca::cplx<T>) * val.size(), hipMemcpyHostToDevice, s));
    }
    QCA_HIP(hipStreamSynchronize(s));
  }
  qca::ContractWs<T> ws;
  ws.T1 = reinterpret_cast<qca::cplx<T>*>(take(sizeof(qca::cplx<T>) * tsize));
  ws.splitk_elems = splitk_elems;
  ws.splitk = reinterpret_cast<qca::cplx<T>*>(take(sizeof(qca::cplx<T>) * splitk_elems));
  qca::heff_apply<T>(kind, reinterpret_cast<const qca::cplx<T>*>(x), reinterpret_cast<qca::cplx<T>*>(y),
                     reinterpret_cast<const qca::cplx<T>*>(L), reinterpret_cast<const qca::cplx<T>*>(R),
                     kind == 2 ? nullptr : &w, Din, Dout, chil, chir, ws, nullptr, s);
}

}  // namespace

extern "C" {

int qca_heff_apply_noalloc(int precision, int kind, const void* x, void* y, const void* L, const void* R,
                           const double* W_host, int Din, int Dout, int chil, int chir, void* stream) {
  return guarded([&] {
    QCA_CHECK(precision == 0 || precision == 1, "precision must be 0 or 1");
    QCA_CHECK(kind >= 0 && kind <= 2, "kind must be 0, 1 or 2");
    QCA_CHECK(x && y && L && R, "null tensor");
    QCA_CHECK(chil >= 1 && chir >= 1 && chil <= 512 && chir <= 512, "bond out of range");
    QCA_CHECK(Din >= 1 && Dout >= 1 && Din <= 16 && Dout <= 16, "MPO bond out of range");
    QCA_CHECK(kind == 2 ? Din == Dout : W_host != nullptr, "bad MPO arguments");
    hipStream_t s = as_stream(stream);
    if (precision == 0) heff_noalloc<double>(kind, x, y, L, R, W_host, Din, Dout, chil, chir, s);
    else heff_noalloc<float>(kind, x, y, L, R, W_host, Din, Dout, chil, chir, s);
  });
}

int qca_release_workspace(void) {
  return guarded([&] {
    QCA_HIP(hipDeviceSynchronize());
    prim_scratch().release();
  });
}

}  // extern "C"
