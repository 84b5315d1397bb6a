// TDVP contractions as launch sequences (shared by the engine and the C-ABI).
//
// Layouts (row-major, interleaved complex), as in reference algorithms/tdvp.py:
//   site tensor  A[p][a][b]                (2, chil, chir)
//   two-site     Theta[(p1 p2)][a][c]      (4, chil, chir)
//   environments L[a][w][a*], R[b][w][b*]  (ket, MPO, bra)
// Each effective-Hamiltonian application is GEMM(L) -> sparse W mixing ->
// GEMM(R); environment updates are the same chain with conj(A) in place of R.
#pragma once

#include <vector>

#include "grouped.hpp"

namespace qca {

template <typename T>
struct ContractWs {
  cplx<T>* T1 = nullptr;  // >= dq * Dmax * chil * chir
  cplx<T>* U = nullptr;   // same size
  cplx<T>* splitk = nullptr;
  long splitk_elems = 0;
};

template <typename T>
void gemm_t(const GemmDesc& d, const cplx<T>* A, const cplx<T>* B, cplx<T>* C, hipStream_t s,
            const ContractWs<T>& ws);

GemmDesc gemm_desc(int M, int N, int K);

// A contraction as its launches: GEMM over the first environment, sparse MPO
// mixing (absent for the bond matrix), GEMM over the second operand.  The
// single-instance launchers run it; the grouped engine batches the same plans.
template <typename T>
struct ContractPlan {
  GemmJob<T> g1;
  bool mix;
  WmixArgs<T> wm;
  GemmJob<T> g3;
};
template <typename T>
ContractPlan<T> heff_plan(int kind, const cplx<T>* x, cplx<T>* y, const cplx<T>* L, const cplx<T>* R,
                          const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                          const ContractWs<T>& ws, const int* skip);
template <typename T>
ContractPlan<T> env_plan(bool left, const cplx<T>* A, const cplx<T>* Eprev, cplx<T>* Enew,
                         const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                         const ContractWs<T>& ws);

// kind 0: 2-site (W = merged, dq = 4), kind 1: 1-site (dq = 2), kind 2: bond.
// Din = MPO dimension of L, Dout = MPO dimension of R (kind 2: Din == Dout).
template <typename T>
void heff_apply(int kind, const cplx<T>* x, cplx<T>* y, const cplx<T>* L, const cplx<T>* R,
                const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                const ContractWs<T>& ws, const int* skip, hipStream_t s);

// Left: Lnew (chir, Dout, chir) from Lprev (chil, Din, chil), W forward form.
// Right: Rnew (chil, Dout, chil) from Rprev (chir, Din, chir), W backward form.
template <typename T>
void env_update(bool left, const cplx<T>* A, const cplx<T>* Eprev, cplx<T>* Enew,
                const WSparseDev<T>* W, int Din, int Dout, int chil, int chir,
                const ContractWs<T>& ws, hipStream_t s);

double heff_flops(int kind, long chil, long chir, long D);

// Sparse MPO construction on the host (dense W[p][q][w][v], dq x dq x Dl x Dr).
struct HostSparse {
  std::vector<int> rowptr, col;
  std::vector<c128> val;
  std::vector<int> outv;  // row % group_rows of each nonzero
  int group_rows = 0;     // rows per output physical index
  int group_max = 0;      // most nonzeros in any group
};
HostSparse build_sparse(const c128* W, int dq, int Dl, int Dr, bool backward);
std::vector<c128> merge_mpo(const c128* Wl, const c128* Wr, int Dl, int Dm, int Dr);

// Device copy of a sparse MPO in both precisions (rowptr/col shared).
struct WSparsePair {
  WSparseDev<double> d64;
  WSparseDev<float> d32;
  std::vector<void*> owned;
  size_t bytes = 0;   // device bytes of the uploaded form
  void upload(const HostSparse& h, int nin);
  void release();
};
// Device bytes WSparsePair::upload allocates for h.
size_t sparse_device_bytes(const HostSparse& h);

}  // namespace qca
