/*
 * qca.h -- C-ABI of the MI355X (gfx950) TDVP time-evolution library libqca.so.
 *
 * The reference (BenjaminDecker/quantum-cellular-automaton) is pure Python and
 * has no FFI; its boundary for this path is the Algorithm interface
 * (algorithms/algorithm.py:10-63) implemented by TDVP (algorithms/tdvp.py) and
 * Exact (algorithms/exact.py).  Each entry point below names the reference
 * function(s) it replaces.  The Python classes under
 * quantum-cellular-automaton_amd/algorithms/ bind these through ctypes.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on failure
 *    (-1 internal, -2 invalid argument, -3 HIP runtime error, -4 an iterative
 *    method did not converge: a Jacobi SVD that made rotations in every one
 *    of its 60 sweeps, or a Krylov exponential fed non-finite numbers);
 *    qca_last_error() returns the message of the last failure on the calling
 *    thread.  No C++ exception crosses this boundary.  An engine whose time
 *    step failed part-way refuses every later call except destroy (its
 *    tensors and cached environments no longer agree).
 *  - Complex numbers are interleaved (re, im) doubles (complex128) or floats
 *    (complex64), the layout of numpy/torch complex arrays.
 *  - Tensor layouts are the reference's: MPS site A[p][a][b] (2, chi_l, chi_r),
 *    MPO W[p][p'][w_l][w_r], environments L[a][w][a*] and R[b][w][b*]
 *    (ket, MPO, bra), two-site Theta[(p1 p2)][a][c] (4, chi_l, chi_r).
 *  - "device" pointers are HIP device memory owned by the caller (for example
 *    torch tensor data_ptr()); "host" pointers are ordinary host memory.
 *    Primitive calls are ordered on the given stream (NULL = default stream).
 *    qca_heff_apply and qca_env_update return once their kernels are queued;
 *    calls that report a host-side result (rank, iteration count) return after
 *    the stream has drained.  The sparse MPO form of each distinct W is built
 *    and uploaded on first use and kept until qca_release_workspace.
 *    Primitives on one device share the library's workspace: calls that
 *    overlap on one device race.
 *  - precision: 0 = complex128 contractions (plain FP64 path),
 *               1 = complex64 contractions on FP32 matrix cores (mixed path;
 *                   Krylov vectors, SVD and QR stay FP64).  An engine of the
 *                   mixed path stops its Jacobi SVD at a relative off-diagonal
 *                   of 1e-10 (plain path: eps * sqrt(rows)); the environment
 *                   variable QCA_MIXED_SVD_TOL, read at creation, overrides it.
 */
#ifndef QCA_H
#define QCA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QCA_OK 0
#define QCA_ERR_INTERNAL (-1)
#define QCA_ERR_ARG (-2)
#define QCA_ERR_HIP (-3)
#define QCA_ERR_NOT_CONVERGED (-4)

/* Library version (major*10000 + minor*100 + patch). */
int qca_version(void);
/* Message of the last failed call on this thread ("" if none).  No
 * counterpart in the reference (error reporting of the C-ABI). */
const char* qca_last_error(void);
/* Number of visible HIP devices.  No counterpart in the reference. */
int qca_device_count(int* count);

/* ------------------------------------------------------------------ engine
 * A device-resident TDVP instance: MPS tensors, environments and workspaces
 * stay in HBM between calls.  Replaces reference TDVP.__init__ /
 * TDVP.do_time_step / TDVP.psi (algorithms/tdvp.py:12-63) together with the
 * MPS.measure observables it feeds (tensor_networks/mps.py:100-140).
 */
typedef struct qca_engine qca_engine;

typedef struct qca_engine_params {
  int device;        /* HIP device ordinal */
  int ncells;        /* number of sites N >= 2 */
  int algorithm;     /* 1 = 1TDVP, 2 = 2TDVP, 3 = adaptive 1TDVP (reference
                        args.algorithm 1tdvp / 2tdvp / a1tdvp; a1tdvp targets use the
                        exact_exponential propagator, tdvp.py:212-266) */
  int precision;     /* 0 plain FP64, 1 mixed */
  int chi_max;       /* reference args.max_bond_dim, 1..256 */
  int reserved;
  double dt;         /* reference args.step_size; propagator exp(-i pi/2 dt H) */
  double eps;        /* reference args.svd_epsilon (2TDVP truncation) */
  double krylov_tol; /* Krylov a-posteriori error target (relative to |x|);
                        0: 1e-14 (the Python drop-in passes 1e-10 for mixed) */
} qca_engine_params;

typedef struct qca_stats {
  long matvecs, matvec_2site, matvec_1site, matvec_bond;
  long site_updates, svd_sweeps, lanczos_limit_hits, svd_calls;
  double flops;      /* algorithmic contraction flops executed */
  double svd_ms;     /* device time of the 2-site SVD splits */
  double krylov_ms;  /* device time of the Krylov exponentials (event-timed on the
                        multi-launch path, kernel clock on the one-launch path) */
} qca_stats;

/* W[i]: host, complex128 (2,2,Wdl[i],Wdr[i]); A[i]: host, complex128
 * (2, chis[i], chis[i+1]).  Canonicalises the state (reference
 * tdvp.py:24-26) and builds all right environments (tdvp.py:37-39). */
int qca_engine_create(const qca_engine_params* params, const double* const* W, const int* Wdl,
                      const int* Wdr, const double* const* A, const int* chis, qca_engine** out);
/* Device bytes qca_engine_create with the same arguments (the state itself
 * is not needed) would allocate: MPS and environments at their bond
 * capacities, sparse MPO forms, Krylov basis and every workspace.  Computed on
 * the host by the constructor's own sizing; needs no device.  No counterpart
 * in the reference (admission of served requests by device footprint). */
int qca_engine_footprint(const qca_engine_params* params, const double* const* W, const int* Wdl,
                         const int* Wdr, const int* chis, size_t* bytes);
/* Device bytes held by an engine (the same sum, from its allocations).  No
 * counterpart in the reference. */
int qca_engine_device_bytes(qca_engine* e, size_t* bytes);
/* Free and total device memory of a HIP device (hipMemGetInfo).  No
 * counterpart in the reference. */
int qca_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
/* Frees the engine's device state (the reference's TDVP object going out of
 * scope, algorithms/tdvp.py:10). */
int qca_engine_destroy(qca_engine* e);
/* nsteps calls of reference TDVP.do_time_step (tdvp.py:50-63), gauge-consistent. */
int qca_engine_step(qca_engine* e, int nsteps);
/* Single-site reduced density matrices, host out[8*N] = rho_i[p][q] as complex.
 * <n_i> = Re rho_i[1][1] (reference mps.py:122-132). */
int qca_engine_measure(qca_engine* e, double* rho_out);
/* The four rows reference Algorithm.measure writes (algorithm.py:58-63 ->
 * mps.py:100-140), computed on the device: population[N] = Re rho_i[1][1],
 * d_population[N] = its rounding (half to even), single_site_entropy[N] =
 * -Tr rho_i log2 rho_i (eigenvalues of the 2x2 Hermitian part, 0 log 0 = 0),
 * bond_dims[N+1] = current bond dimensions (recorded before any gauge change,
 * mps.py:106-109).  All host arrays of doubles. */
int qca_engine_observables(qca_engine* e, double* population, double* d_population,
                           double* single_site_entropy, double* bond_dims);
/* Current bond dimensions, host out[N+1] (reference mps.py:106-109). */
int qca_engine_bond_dims(qca_engine* e, int* out);
/* Copy site tensor i to host complex128 (2, chi_i, chi_{i+1}): the tensor
 * psi.A[i] of the reference's evolved MPS (algorithms/tdvp.py:42-44). */
int qca_engine_get_site(qca_engine* e, int site, double* out);
/* Counters since the last reset.  Waits for the engine's stream: the
 * one-launch Krylov exponentials keep their counts (and kernel time, in
 * krylov_ms) on the device.  No counterpart in the reference (instrumentation
 * for the site-update and flop accounting of the benchmark). */
int qca_engine_stats(qca_engine* e, qca_stats* out);
/* Zeroes the counters above.  No counterpart in the reference. */
int qca_engine_reset_stats(qca_engine* e);
/* Accumulated GPU time (ms) of the engine's stream since the last reset,
 * measured with HIP events around each qca_engine_step call.  No
 * counterpart in the reference (benchmark instrumentation). */
int qca_engine_gpu_ms(qca_engine* e, double* ms);
/* Krylov blocks of at most max_n elements (0..4096) use the one-launch
 * exponential (one workgroup, no host round trip); larger ones use the
 * multi-launch path.  Default 1024, or QCA_FUSED_KRYLOV_MAX_N at creation.
 * Raising it trades single-instance latency for throughput when many engines
 * share the GPU.  No counterpart in the reference (a scheduling choice of this
 * implementation; results agree to rounding either way). */
int qca_engine_set_fused_krylov_limit(qca_engine* e, long max_n);

/* Advances each of the n engines by nsteps time steps using nthreads host
 * worker threads (1..256).  Each engine runs on its own stream, so up to
 * nthreads instances are in flight on the GPU at once; the workers take engines
 * from a round-robin queue one time step at a time, so instances of unequal
 * cost stay balanced.  Equivalent to calling qca_engine_step(engines[i], nsteps)
 * for every i.  On failure returns the first error, stops handing out work and
 * sets *failed (if not NULL) to the failing engine's index, else to -1.
 * No counterpart in the reference, which runs one simulation per process
 * (quantum_game.py:75-119 loops over initial states sequentially). */
int qca_batch_step(qca_engine* const* engines, int n, int nsteps, int nthreads, int* failed);

/* ------------------------------------------------------------------ groups
 * Lockstep 2TDVP of many engines with grouped launches: at every sweep
 * position each kernel runs once for all instances (one record per instance
 * holding its operands and shapes), with one host round trip per position
 * for the truncation ranks of all splits and one per chunk of Krylov
 * iterations for all convergence flags.  Every instance computes exactly what
 * qca_engine_step computes for it alone (same GEMM plans, kernels,
 * reductions and stopping tests), so results are bit-identical.  Instances
 * may differ in cell count, bond cap, MPO, time step and cutoff.  No
 * counterpart in the reference, which runs one simulation per process
 * (quantum_game.py:75-119); each instance's step is reference
 * TDVP.do_time_step (tdvp.py:50-63, 107-145).
 */
typedef struct qca_group qca_group;

typedef struct qca_group_stats {
  long steps;              /* group time steps taken */
  long krylov_iterations;  /* grouped Lanczos iterations launched */
  long host_syncs;         /* batch-wide host round trips */
  double krylov_ms;        /* device time of the multi-launch Krylov phases */
  double svd_ms;           /* device time of the splits */
  double total_ms;         /* device time of the group steps */
} qca_group_stats;

/* Groups n engines: all 2TDVP, one device, one precision, each at most once.
 * The engines stay owned by the caller and must outlive the group; they can
 * still be measured, read and stepped alone between group steps (not
 * concurrently with them). */
int qca_group_create(qca_engine* const* engines, int n, qca_group** out);
/* nsteps time steps of every engine of the group.  A failure marks every
 * engine of the group as failed (their sweeps stopped part-way). */
int qca_group_step(qca_group* g, int nsteps);
/* Counters since creation or the last reset. */
int qca_group_get_stats(qca_group* g, qca_group_stats* out);
int qca_group_reset_stats(qca_group* g);
/* Frees the group's records and buffers (not the engines). */
int qca_group_destroy(qca_group* g);

/* ------------------------------------------------------------------ primitives */

/* y = H_eff x (matrix-free).  Replaces the dense assembly + product of
 * reference tdvp.py:298-310 (_assemble_H_eff), 312-326 (_assemble_K_eff) and
 * lautils.py:80-81.  kind 0: 2-site, W (4,4,Din,Dout) merged (tdvp.py:282-284),
 * x,y (4,chil,chir); kind 1: 1-site, W (2,2,Din,Dout), x,y (2,chil,chir);
 * kind 2: bond, W unused, x,y (chil,chir), Din == Dout.  L (chil,Din,chil),
 * R (chir,Dout,chir).  All device tensors in the precision's type. */
int qca_heff_apply(int precision, int kind, const void* x, void* y, const void* L, const void* R,
                   const double* W_host, int Din, int Dout, int chil, int chir, void* stream);

/* qca_heff_apply without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and result as qca_heff_apply (the same GEMM plans, so the
 * output is identical); its sparse MPO form and contraction workspace live in
 * a static device array of 64 MiB per device, so an application whose
 * workspace exceeds it (e.g. the complex128 two-site block at chi = 128,
 * D = 13) is rejected with QCA_ERR_ARG.  The MPO form is copied in on every
 * call, which returns after that copy; the kernels are queued on the stream.
 * All calls on one device share the array: calls that overlap on one device
 * (different streams or threads) race. */
int qca_heff_apply_noalloc(int precision, int kind, const void* x, void* y, const void* L,
                           const void* R, const double* W_host, int Din, int Dout, int chil,
                           int chir, void* stream);

/* New environment layer.  Replaces reference tdvp.py:328-347
 * (_assemble_new_layer_H_eff).  side 0 = left: Eprev (chil,Dl,chil) ->
 * Enew (chir,Dr,chir); side 1 = right: Eprev (chir,Dr,chir) -> Enew (chil,Dl,chil).
 * A (2,chil,chir) and environments in the precision's type; W host (2,2,Dl,Dr). */
int qca_env_update(int precision, int side, const void* A, const void* Eprev, void* Enew,
                   const double* W_host, int Dl, int Dr, int chil, int chir, void* stream);

/* x <- exp(-i (pi/2) delta H_eff) x by Krylov (FP64 recurrence), x complex128
 * on the device.  Replaces reference lautils.py:44-81 (calculate_U/timestep)
 * as used by tdvp.py:147-188, 349-365.  Arguments as qca_heff_apply. */
int qca_krylov_expm(int precision, int kind, void* x, const void* L, const void* R,
                    const double* W_host, int Din, int Dout, int chil, int chir, double delta,
                    double tol, int* iters, void* stream);

/* SVD split and truncation of an evolved two-site tensor (complex128, device).
 * Replaces reference tdvp.py:286-296.  right_sweep != 0: Al (2,chil,k) is the
 * left isometry and Ar (2,k,chir) = diag(s/|s|) V^H; right_sweep == 0: Ar is
 * the right isometry and Al = U diag(s/|s|).  k is returned; sigma_host (may be
 * NULL) receives min(2chil,2chir) singular values, descending. */
int qca_svd_split(const void* theta, void* Al, void* Ar, int chil, int chir, int chi_max,
                  double eps, int right_sweep, int* k, double* sigma_host, void* stream);

/* Complete-QR shift of one site truncated/padded to a target bond, LAPACK
 * zgeqrf/zungqr conventions.  Replaces reference mps.py:83-98 with
 * truncate_and_pad_into_shape (mps.py:226-236) as used by tdvp.py:74-78
 * (side 0: Anew (2,chil,target), C (target,chir)) and tdvp.py:95-99
 * (side 1: Anew (2,target,chir), C (chil,target)).  complex128, device. */
int qca_qr_site(int side, const void* A, int chil, int chir, int target, void* Anew, void* C,
                void* stream);

/* out = H psi on the 2^N basis (site 0 = most significant bit), closed form of
 * the rule Hamiltonian with activation interval [lo, hi) (Python range).
 * Replaces the dense matrix of reference mpo.py:202-228 + exact.py:28. */
int qca_exact_apply_h(int precision, const void* psi, void* out, int ncells, int distance, int lo,
                      int hi, void* stream);

/* psi <- exp(-i (pi/2) dt H) psi on the 2^N basis (complex128, device) by
 * Krylov with the closed-form H.  Replaces reference exact.py:15-28.  The
 * step is split into 2^l equal substeps when the Krylov basis (at most 64
 * vectors, at least 24, within 16 GiB) cannot hold the whole step: a substep
 * that reaches the basis limit is discarded and retried at half the length.
 * *iters receives the matvec count summed over the substeps. */
int qca_exact_expm(void* psi, int ncells, int distance, int lo, int hi, double dt, double tol,
                   int* iters, void* stream);

/* Single-site reduced density matrices of a 2^N state (complex128, device):
 * host rho_out[8*N] = rho_i[p][q].  Replaces MPS.from_vector + MPS.measure as
 * used on the exact state (reference exact.py:19-21, mps.py:54-73, 100-140). */
int qca_exact_rho(const void* psi, int ncells, double* rho_out, void* stream);

/* Release the primitives' workspace on the current device.  No counterpart
 * in the reference (device memory management of the C-ABI). */
int qca_release_workspace(void);

#ifdef __cplusplus
}
#endif

#endif /* QCA_H */
