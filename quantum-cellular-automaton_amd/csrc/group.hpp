// Lockstep stepping of many device-resident 2TDVP engines with grouped
// launches: every kernel of a sweep position runs once for all instances.
#pragma once

#include <vector>

#include "engine.hpp"

namespace qca {

struct GroupStats {
  long steps = 0;           // group time steps
  long launches_krylov = 0; // grouped Lanczos iterations launched
  long svd_sweeps = 0;      // grouped Jacobi sweeps launched
  long host_syncs = 0;      // batch-wide host round trips
  double krylov_ms = 0;     // device time of the Krylov phases (events)
  double svd_ms = 0;        // device time of the split phases (layout change to ranks)
  double total_ms = 0;      // device time of whole group steps
};

// A group of engines (same device, 2TDVP, same precision) advanced together.
// Each engine keeps its own tensors, environments and workspaces; the group
// holds only launch records and convergence words.  Instance i at sweep
// position t works on its own bond (right sweep: bond t; left sweep: bond
// N_i - 1 - t) and sits out positions beyond its chain.  Every instance
// computes exactly what Engine::step computes for it alone: the same GEMM
// plans (split-K chosen from its own shapes), the same Lanczos and Jacobi
// kernels and reductions, the same stopping tests.
class Group {
 public:
  explicit Group(const std::vector<Engine*>& engines);
  ~Group();
  Group(const Group&) = delete;
  Group& operator=(const Group&) = delete;
  void step(int nsteps);
  const GroupStats& stats() const { return stats_; }
  void reset_stats() { stats_ = GroupStats(); }
  int size() const { return (int)e_.size(); }

 private:
  struct Bond {        // one instance's work at the current sweep position
    int inst;
    int s;             // left site of the two-site block
    bool right;        // right sweep
    int chil, chir;    // outer bonds of the block
    int k = 0;         // truncation rank
    bool back = false; // back-evolution follows (not at the chain end)
  };
  void sweep(bool right);
  void position(std::vector<Bond>& bonds);
  template <typename T>
  void evolve(std::vector<int> inst, std::vector<int> kind, std::vector<int> site, std::vector<c128*> x,
              std::vector<long> n, std::vector<int> chil, std::vector<int> chir, std::vector<double> delta);
  void sync();
  void release();
  void fail(int inst, const char* what);

  std::vector<Engine*> e_;
  int device_ = 0;
  int precision_ = 0;
  hipStream_t s_ = nullptr;
  RecordArena arena_;
  GroupFlags flags_;
  // gather buffers: one row of words per instance
  unsigned long long* gather_dev_ = nullptr;
  unsigned long long* gather_host_ = nullptr;
  const void** gather_src_host_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr, ev2_ = nullptr, ev3_ = nullptr;
  GroupStats stats_;
  void gather_words(const std::vector<const void*>& src, int words);
};

}  // namespace qca
