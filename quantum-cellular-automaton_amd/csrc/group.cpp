// Lockstep 2TDVP over many engines with grouped launches (see group.hpp).
//
// Per sweep position the host makes two passes over the instances:
//   A: merge the two-site block (reference tdvp.py:278-284), evolve it by
//      +dt/2 (tdvp.py:349-365 as a Krylov exponential), split it by SVD
//      (tdvp.py:286-296); then one host round trip reads every rank;
//   B: gather the isometry and form the other factor (normalised), update the
//      environment (tdvp.py:328-347) and evolve the new centre site by -dt/2
//      (tdvp.py:147-162), skipped at the chain end as in tdvp.py:117-125.
// Each step of a pass is one grouped launch over the instances that take part.
// Krylov exponentials of at most fused_max_n elements run as one workgroup per
// instance (one launch for all of them); longer ones iterate in lockstep, with
// one host check of every instance's convergence flag per chunk of iterations.
#include "group.hpp"

#include <algorithm>
#include <cstring>
#include <string>

namespace qca {

namespace {
constexpr double kHalfPi = 1.5707963267948966;
constexpr int kGatherWords = 4;   // 32 bytes per instance: SvdResult, head of LanczosCtrl
}  // namespace

Group::Group(const std::vector<Engine*>& engines) : e_(engines) {
  QCA_CHECK(!e_.empty(), "empty engine group");
  device_ = e_[0]->device_;
  precision_ = e_[0]->p_.precision;
  for (Engine* e : e_) {
    QCA_CHECK(e != nullptr, "null engine in group");
    QCA_CHECK(e->p_.algorithm == 2, "grouped stepping supports 2TDVP engines only");
    QCA_CHECK(e->device_ == device_, "engines of a group must share one device");
    QCA_CHECK(e->p_.precision == precision_, "engines of a group must share one precision");
  }
  for (size_t i = 0; i < e_.size(); ++i)
    for (size_t j = i + 1; j < e_.size(); ++j)
      QCA_CHECK(e_[i] != e_[j], "an engine appears twice in a group");
  const int n = (int)e_.size();
  try {
    QCA_HIP(hipSetDevice(device_));
    QCA_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    arena_.reserve((size_t)(1 << 20) + (size_t)n * (160 << 10));
    QCA_HIP(hipMalloc((void**)&flags_.dev, sizeof(int) * n));
    QCA_HIP(hipHostMalloc((void**)&flags_.host, sizeof(int) * n));
    flags_.cap = n;
    QCA_HIP(hipMalloc((void**)&gather_dev_, sizeof(unsigned long long) * kGatherWords * n));
    QCA_HIP(hipHostMalloc((void**)&gather_host_, sizeof(unsigned long long) * kGatherWords * n));
    QCA_HIP(hipEventCreate(&ev0_));
    QCA_HIP(hipEventCreate(&ev1_));
    QCA_HIP(hipEventCreate(&ev2_));
    QCA_HIP(hipEventCreate(&ev3_));
  } catch (...) {
    release();
    throw;
  }
}

Group::~Group() { release(); }

void Group::release() {
  (void)hipSetDevice(device_);
  if (s_) (void)hipStreamSynchronize(s_);
  arena_.release();
  if (flags_.dev) (void)hipFree(flags_.dev);
  if (flags_.host) (void)hipHostFree(flags_.host);
  if (gather_dev_) (void)hipFree(gather_dev_);
  if (gather_host_) (void)hipHostFree(gather_host_);
  flags_ = GroupFlags();
  gather_dev_ = nullptr;
  gather_host_ = nullptr;
  for (hipEvent_t* ev : {&ev0_, &ev1_, &ev2_, &ev3_})
    if (*ev) {
      (void)hipEventDestroy(*ev);
      *ev = nullptr;
    }
  if (s_) (void)hipStreamDestroy(s_);
  s_ = nullptr;
}

void Group::sync() {
  QCA_HIP(hipStreamSynchronize(s_));
  stats_.host_syncs++;
}

// Copies the first kGatherWords 8-byte words at every src[i] into
// gather_host_[i * kGatherWords ...] (one launch, one transfer, one sync).
void Group::gather_words(const std::vector<const void*>& src, int words) {
  QCA_CHECK(words <= kGatherWords && (int)src.size() <= flags_.cap, "gather too large");
  const void* const* dsrc = arena_.push(src);
  arena_.upload(s_);
  launch_gather_words(dsrc, words, (int)src.size(), gather_dev_, s_);
  QCA_HIP(hipMemcpyAsync(gather_host_, gather_dev_, sizeof(unsigned long long) * words * src.size(),
                         hipMemcpyDeviceToHost, s_));
  sync();
}

void Group::step(int nsteps) {
  QCA_HIP(hipSetDevice(device_));
  for (int it = 0; it < nsteps; ++it) {
    // work queued on the engines' own streams (creation, measurements) first
    for (Engine* e : e_) QCA_HIP(hipStreamSynchronize(e->s_));
    QCA_HIP(hipEventRecord(ev0_, s_));
    sweep(true);
    sweep(false);
    QCA_HIP(hipEventRecord(ev1_, s_));
    sync();
    float ms = 0;
    QCA_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    stats_.total_ms += ms;
    stats_.steps++;
  }
}

// Right sweep: position t is bond (t, t+1) of every chain longer than t+1
// (reference tdvp.py:107-125).  Left sweep: position t is bond (N-2-t, N-1-t)
// (tdvp.py:127-145).
void Group::sweep(bool right) {
  int maxpos = 0;
  for (Engine* e : e_) maxpos = std::max(maxpos, e->N_ - 1);
  std::vector<Bond> bonds;
  for (int t = 0; t < maxpos; ++t) {
    bonds.clear();
    for (int i = 0; i < (int)e_.size(); ++i) {
      const Engine* e = e_[i];
      const int N = e->N_;
      if (t >= N - 1) continue;
      Bond b;
      b.inst = i;
      b.right = right;
      b.s = right ? t : N - 2 - t;
      b.chil = e->chi_[b.s];
      b.chir = e->chi_[b.s + 2];
      b.back = right ? (b.s < N - 2) : (b.s > 0);
      bonds.push_back(b);
    }
    position(bonds);
  }
}

void Group::fail(int inst, const char* what) {
  throw Error(QCA_ERR_NOT_CONVERGED_CODE, std::string("instance ") + std::to_string(inst) + ": " + what);
}

void Group::position(std::vector<Bond>& bonds) {
  const bool mixed = precision_ == PREC_MIXED;
  // ---------------------------------------------------------------- pass A
  arena_.reset();
  {
    std::vector<GemmJob<double>> merge;
    for (const Bond& b : bonds) {
      Engine* e = e_[b.inst];
      const int chim = e->chi_[b.s + 1];
      GemmDesc g = merge_theta_desc(b.chil, chim, b.chir);
      merge.push_back(GemmJob<double>{g, e->A_[b.s], e->A_[b.s + 1], e->theta_, e->ws64_.splitk,
                                      e->ws64_.splitk_elems});
    }
    GemmGroup<double> gm;
    gm.prepare(merge, arena_);
    arena_.upload(s_);
    gm.launch(s_);
  }
  {
    std::vector<int> inst, kind, site, chil, chir;
    std::vector<c128*> x;
    std::vector<long> n;
    std::vector<double> delta;
    for (const Bond& b : bonds) {
      Engine* e = e_[b.inst];
      inst.push_back(b.inst);
      kind.push_back(0);
      site.push_back(b.s);
      x.push_back(e->theta_);
      n.push_back(4L * b.chil * b.chir);
      chil.push_back(b.chil);
      chir.push_back(b.chir);
      delta.push_back(e->p_.dt / 2);
    }
    if (mixed) evolve<float>(inst, kind, site, x, n, chil, chir, delta);
    else evolve<double>(inst, kind, site, x, n, chil, chir, delta);
  }
  QCA_HIP(hipEventRecord(ev2_, s_));
  {
    std::vector<GRec<ThetaSvdArgs>> th;
    std::vector<SvdJob> jobs;
    for (const Bond& b : bonds) {
      Engine* e = e_[b.inst];
      const int rows = b.right ? 2 * b.chil : 2 * b.chir;
      const int cols = b.right ? 2 * b.chir : 2 * b.chil;
      th.push_back(theta_svd_rec(e->theta_, e->svdX_, b.chil, b.chir, b.right ? 0 : 1));
      jobs.push_back(SvdJob{e->svdX_, rows, cols, e->sigma_, e->perm_, e->svdres_, e->p_.chi_max, e->p_.eps,
                            std::min(2 * b.chil, 2 * b.chir), &e->jw_});
    }
    KGroup<ThetaSvdArgs> gth;
    gth.prepare(th, arena_);
    arena_.upload(s_);
    launch_theta_svd_grouped(gth, s_);
    grouped_jacobi_svd(jobs, arena_, flags_, s_);
    QCA_HIP(hipEventRecord(ev3_, s_));
    std::vector<const void*> src;
    for (const Bond& b : bonds) src.push_back(e_[b.inst]->svdres_);
    gather_words(src, kGatherWords);
  }
  {
    float ms = 0;
    QCA_HIP(hipEventElapsedTime(&ms, ev2_, ev3_));
    stats_.svd_ms += ms;
  }
  static_assert(sizeof(SvdResult) == kGatherWords * 8, "SvdResult is four words");
  std::vector<double> inv(bonds.size());
  for (size_t i = 0; i < bonds.size(); ++i) {
    Bond& b = bonds[i];
    Engine* e = e_[b.inst];
    SvdResult r;
    std::memcpy(&r, gather_host_ + kGatherWords * i, sizeof(r));
    e->stats_.svd_calls++;
    e->stats_.svd_sweeps += r.nsweeps;
    const int rows = b.right ? 2 * b.chil : 2 * b.chir;
    const int cols = b.right ? 2 * b.chir : 2 * b.chil;
    if (cols > 32) {
      const int kq = std::min(rows, cols);
      const bool pre = e->jw_.pre != nullptr && kq > 32 &&
                       e->jw_.pre_elems >= jacobi_precond_elems(rows, cols);
      (pre ? e->jw_.pred_sweeps_pre : e->jw_.pred_sweeps) = r.nsweeps;
    }
    if (!r.converged) fail(b.inst, "Jacobi SVD of the two-site split did not converge");
    QCA_CHECK(r.k >= 1, "SVD truncation produced an empty bond");
    QCA_CHECK(r.k <= e->cap_[b.s + 1], "truncation rank above bond capacity");
    b.k = r.k;
    inv[i] = 1.0 / r.norm_kept;
  }
  // ---------------------------------------------------------------- pass B
  sync();
  arena_.reset();
  {
    std::vector<GRec<GatherIsoArgs>> gi;
    std::vector<GemmJob<double>> split;
    for (size_t i = 0; i < bonds.size(); ++i) {
      const Bond& b = bonds[i];
      Engine* e = e_[b.inst];
      const int rows = b.right ? 2 * b.chil : 2 * b.chir;
      if (b.right) {
        gi.push_back(gather_iso_rec(e->svdX_, e->perm_, e->sigma_, e->A_[b.s], rows, b.chil, b.k, 0));
        split.push_back(GemmJob<double>{split_theta_desc(true, b.chil, b.chir, b.k, inv[i]), e->A_[b.s],
                                        e->theta_, e->A_[b.s + 1], e->ws64_.splitk, e->ws64_.splitk_elems});
      } else {
        gi.push_back(gather_iso_rec(e->svdX_, e->perm_, e->sigma_, e->A_[b.s + 1], rows, b.chir, b.k, 1));
        split.push_back(GemmJob<double>{split_theta_desc(false, b.chil, b.chir, b.k, inv[i]), e->theta_,
                                        e->A_[b.s + 1], e->A_[b.s], e->ws64_.splitk, e->ws64_.splitk_elems});
      }
      e->chi_[b.s + 1] = b.k;
      e->stats_.site_updates++;
    }
    KGroup<GatherIsoArgs> ggi;
    ggi.prepare(gi, arena_);
    GemmGroup<double> gsp;
    gsp.prepare(split, arena_);
    arena_.upload(s_);
    launch_gather_iso_grouped(ggi, s_);
    gsp.launch(s_);
  }
  // environment of the new left (right sweep) or right (left sweep) site
  {
    std::vector<GRec<ConvArgs>> conv;
    std::vector<GemmJob<double>> g1d, g3d;
    std::vector<GemmJob<float>> g1f, g3f;
    std::vector<GRec<WmixArgs<double>>> wmd;
    std::vector<GRec<WmixArgs<float>>> wmf;
    size_t lds = 0;
    for (const Bond& b : bonds) {
      if (!b.back) continue;
      Engine* e = e_[b.inst];
      const bool left = b.right;
      const int site = b.right ? b.s : b.s + 1;
      const int chil = e->chi_[site], chir = e->chi_[site + 1];
      const long lr = (long)chil * chir;
      const WSparsePair& wp = left ? e->Wf_[site] : e->Wb_[site];
      const int Din = left ? e->Dl_[site] : e->Dr_[site];
      const int Dout = left ? e->Dr_[site] : e->Dl_[site];
      if (mixed) {
        conv.push_back(conv_rec(e->A_[site], e->a32_, 2 * lr, nullptr));
        const c64* Eprev = left ? e->Lenv<cplx<float>>(site - 1) : e->Renv<cplx<float>>(site + 1);
        c64* Enew = left ? e->Lenv<cplx<float>>(site) : e->Renv<cplx<float>>(site);
        const ContractPlan<float> p = env_plan<float>(left, e->a32_, Eprev, Enew, &wp.d32, Din, Dout, chil, chir,
                                                      e->ws32_);
        g1f.push_back(p.g1);
        wmf.push_back(wmix_rec<float>(p.wm.Tin, p.wm.in_stride, p.wm.nxc, p.wm.nin, p.wm.U, p.wm.nq, p.wm.Dout,
                                      wp.d32, nullptr));
        lds = std::max(lds, wmix_lds_bytes<float>(p.wm.nin, p.wm.nq, p.wm.Dout));
        g3f.push_back(p.g3);
      } else {
        const c128* Eprev = left ? e->Lenv<cplx<double>>(site - 1) : e->Renv<cplx<double>>(site + 1);
        c128* Enew = left ? e->Lenv<cplx<double>>(site) : e->Renv<cplx<double>>(site);
        const ContractPlan<double> p = env_plan<double>(left, e->A_[site], Eprev, Enew, &wp.d64, Din, Dout,
                                                        chil, chir, e->ws64_);
        g1d.push_back(p.g1);
        wmd.push_back(wmix_rec<double>(p.wm.Tin, p.wm.in_stride, p.wm.nxc, p.wm.nin, p.wm.U, p.wm.nq,
                                       p.wm.Dout, wp.d64, nullptr));
        lds = std::max(lds, wmix_lds_bytes<double>(p.wm.nin, p.wm.nq, p.wm.Dout));
        g3d.push_back(p.g3);
      }
      e->stats_.flops += heff_flops(1, chil, chir, std::max(e->Dl_[site], e->Dr_[site]));
    }
    if (mixed) {
      KGroup<ConvArgs> gc;
      gc.prepare(conv, arena_);
      GemmGroup<float> a, c;
      KGroup<WmixArgs<float>> w;
      a.prepare(g1f, arena_);
      w.prepare(wmf, arena_);
      c.prepare(g3f, arena_);
      arena_.upload(s_);
      launch_conv_grouped(gc, s_);
      a.launch(s_);
      launch_wmix_grouped<float>(w, lds, s_);
      c.launch(s_);
    } else {
      GemmGroup<double> a, c;
      KGroup<WmixArgs<double>> w;
      a.prepare(g1d, arena_);
      w.prepare(wmd, arena_);
      c.prepare(g3d, arena_);
      arena_.upload(s_);
      a.launch(s_);
      launch_wmix_grouped<double>(w, lds, s_);
      c.launch(s_);
    }
  }
  // back-evolution of the new centre site by -dt/2
  {
    std::vector<int> inst, kind, site, chil, chir;
    std::vector<c128*> x;
    std::vector<long> n;
    std::vector<double> delta;
    for (const Bond& b : bonds) {
      if (!b.back) continue;
      Engine* e = e_[b.inst];
      const int st = b.right ? b.s + 1 : b.s;
      const int cl = e->chi_[st], cr = e->chi_[st + 1];
      inst.push_back(b.inst);
      kind.push_back(1);
      site.push_back(st);
      x.push_back(e->A_[st]);
      n.push_back(2L * cl * cr);
      chil.push_back(cl);
      chir.push_back(cr);
      delta.push_back(-e->p_.dt / 2);
    }
    if (!inst.empty()) {
      if (mixed) evolve<float>(inst, kind, site, x, n, chil, chir, delta);
      else evolve<double>(inst, kind, site, x, n, chil, chir, delta);
    }
  }
  sync();
}

// Krylov exponentials x_i <- exp(-i (pi/2) delta_i H_eff) x_i, one per listed
// instance (Engine::evolve for each, grouped).
template <typename T>
void Group::evolve(std::vector<int> inst, std::vector<int> kind, std::vector<int> site, std::vector<c128*> x,
                   std::vector<long> n, std::vector<int> chil, std::vector<int> chir,
                   std::vector<double> delta) {
  constexpr bool mixed = !std::is_same<T, double>::value;
  std::vector<GRec<FusedKrylovArgs<T>>> fused;
  size_t fused_lds = 0;
  std::vector<LzArgs> lz;
  std::vector<int> lz_item;
  std::vector<GemmJob<T>> g1, g3;
  std::vector<GRec<WmixArgs<T>>> wm;
  size_t wm_lds = 0;
  int chunk = 2, max_m = LZ_MAXK;
  for (size_t i = 0; i < inst.size(); ++i) {
    Engine* e = e_[inst[i]];
    const int k = kind[i], st = site[i];
    const WSparsePair* wp = k == 0 ? &e->W2f_[st] : &e->Wf_[st];
    const int Din = e->Dl_[st];
    const int Dout = k == 0 ? e->Dr_[st + 1] : e->Dr_[st];
    const cplx<T>* L = e->Lenv<cplx<T>>(st - 1);
    const cplx<T>* R = e->Renv<cplx<T>>(k == 0 ? st + 2 : st + 1);
    const WSparseDev<T>* W;
    if constexpr (mixed) W = &wp->d32; else W = &wp->d64;
    const double tau = kHalfPi * delta[i];
    if (n[i] <= e->fused_max_n_) {
      FusedKrylovArgs<T> a;
      a.kind = k;
      a.chil = chil[i];
      a.chir = chir[i];
      a.n = n[i];
      a.Din = Din;
      a.Dout = Dout;
      a.L = L;
      a.R = R;
      a.rowptr = wp->d64.rowptr;
      a.col = wp->d64.col;
      a.val = W->val;
      a.nnz = wp->d64.nnz;
      a.x = x[i];
      a.V = e->lw_.V;
      a.T1 = reinterpret_cast<cplx<T>*>(e->T_);
      a.U = reinterpret_cast<cplx<T>*>(e->U_);
      a.tally = e->tally_;
      a.tau = tau;
      a.tol = e->p_.krylov_tol;
      a.max_m = e->lw_.max_m;
      a.flops_per_matvec = heff_flops(k, chil[i], chir[i], std::max(a.Din, a.Dout));
      FusedKrylovArgs<T> placed;
      fused_lds = std::max(fused_lds, fused_krylov_place(a, placed));
      fused.push_back(grec(placed, 1));
      continue;
    }
    LzArgs z{};
    z.ctrl = e->lw_.ctrl;
    z.x = x[i];
    z.y = x[i];
    z.V = e->lw_.V;
    z.w64 = e->lw_.w64;
    z.part = e->lw_.part;
    z.part2 = e->lw_.part2;
    z.part_norm = e->lw_.part_norm;
    z.coef1 = e->lw_.coef1;
    z.coef2 = e->lw_.coef2;
    z.n = n[i];
    z.nblk = lanczos_nblk(n[i]);
    QCA_CHECK(z.nblk <= e->lw_.max_blocks, "Lanczos vector longer than the workspace");
    z.max_m = e->lw_.max_m;
    z.tau = tau;
    z.tol = e->p_.krylov_tol;
    const int* skip = &e->lw_.ctrl->done;
    ContractPlan<T> p;
    if constexpr (mixed) {
      z.w = e->y32_;
      z.v32 = e->x32_;
      p = heff_plan<T>(k, e->x32_, e->y32_, L, R, W, Din, Dout, chil[i], chir[i], e->ws32_, skip);
    } else {
      z.w = e->y64_;
      z.v32 = nullptr;
      p = heff_plan<T>(k, e->lw_.V, e->y64_, L, R, W, Din, Dout, chil[i], chir[i], e->ws64_, skip);
      p.g1.a_jstride = n[i];   // the matvec reads basis vector V_j (the first GEMM's A)
    }
    g1.push_back(p.g1);
    if (p.mix) {
      wm.push_back(wmix_rec<T>(p.wm.Tin, p.wm.in_stride, p.wm.nxc, p.wm.nin, p.wm.U, p.wm.nq, p.wm.Dout, *W, skip));
      wm_lds = std::max(wm_lds, wmix_lds_bytes<T>(p.wm.nin, p.wm.nq, p.wm.Dout));
    }
    g3.push_back(p.g3);
    lz.push_back(z);
    lz_item.push_back((int)i);
    chunk = std::max(chunk, e->kpred_[k]);
    max_m = std::min(max_m, e->lw_.max_m);
  }
  KGroup<FusedKrylovArgs<T>> gf;
  gf.prepare(fused, arena_);
  LzGroup glz;
  GemmGroup<T> G1, G3;
  KGroup<WmixArgs<T>> GW;
  if (!lz.empty()) {
    glz.prepare(lz, arena_);
    G1.prepare(g1, arena_);
    GW.prepare(wm, arena_);
    G3.prepare(g3, arena_);
  }
  arena_.upload(s_);
  launch_fused_krylov_grouped<T>(gf, fused_lds, s_);
  if (lz.empty()) return;
  QCA_HIP(hipEventRecord(ev2_, s_));
  glz.begin(s_);
  int j = 0;
  std::vector<const void*> src;
  for (const LzArgs& z : lz) src.push_back(z.ctrl);
  while (true) {
    for (int c = 0; c < chunk && j < max_m - 1; ++c, ++j) {
      G1.launch(s_, j);
      launch_wmix_grouped<T>(GW, wm_lds, s_);
      G3.launch(s_, j);
      glz.tail<T>(j, s_);
      stats_.launches_krylov++;
    }
    gather_words(src, kGatherWords);
    bool all = true;
    for (size_t i = 0; i < lz.size(); ++i) {
      LanczosCtrl head;
      std::memcpy(&head, gather_host_ + kGatherWords * i, 32);
      all = all && head.done != 0;
    }
    if (all) break;
    if (j >= max_m - 1)
      throw Error(QCA_ERR_NOT_CONVERGED_CODE, "Krylov exponential did not converge (non-finite input?)");
    chunk = 2;
  }
  glz.finish(s_);
  QCA_HIP(hipEventRecord(ev3_, s_));
  sync();
  float ms = 0;
  QCA_HIP(hipEventElapsedTime(&ms, ev2_, ev3_));
  stats_.krylov_ms += ms;
  for (size_t q = 0; q < lz.size(); ++q) {
    const int i = lz_item[q];
    Engine* e = e_[inst[i]];
    LanczosCtrl head;
    std::memcpy(&head, gather_host_ + kGatherWords * q, 32);
    const int m = head.m, k = kind[i], st = site[i];
    e->kpred_[k] = std::max(2, m);
    if (head.hit_limit) e->stats_.lanczos_limit_hits++;
    e->stats_.matvecs += m;
    if (k == 0) e->stats_.matvec_2site += m;
    else if (k == 1) e->stats_.matvec_1site += m;
    else e->stats_.matvec_bond += m;
    const int D = std::max(e->Dl_[st], k == 0 ? e->Dr_[st + 1] : e->Dr_[st]);
    e->stats_.flops += m * heff_flops(k, chil[i], chir[i], D);
  }
}

template void Group::evolve<double>(std::vector<int>, std::vector<int>, std::vector<int>, std::vector<c128*>,
                                    std::vector<long>, std::vector<int>, std::vector<int>, std::vector<double>);
template void Group::evolve<float>(std::vector<int>, std::vector<int>, std::vector<int>, std::vector<c128*>,
                                   std::vector<long>, std::vector<int>, std::vector<int>, std::vector<double>);

}  // namespace qca
