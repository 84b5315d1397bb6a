"""CPU restatement of the reference's TDVP time-evolution path (test infrastructure).

THIS PACKAGE IS THE ORACLE.  It exists so that tests, the smoke check and the
CPU baseline leg of the benchmark can check the GPU implementation against an
independent complex128 NumPy computation.  It is never imported by the product
code under ``quantum-cellular-automaton_amd/``; the product path fails loudly
when its HIP library is missing instead of falling back to anything here.

Modules (each function cites the reference file:line it restates):

* ``automaton``  -- rule Hamiltonian as an MPO (reference tensor_networks/mpo.py),
                    plus the closed-form bit-flip Hamiltonian used by the exact path.
* ``mps``        -- MPS container helpers: product states, QR shifts, measure,
                    vector conversion (reference tensor_networks/mps.py).
* ``local``      -- the local contractions, dense and Krylov exponentials
                    (reference algorithms/tdvp.py:298-365, lautils/lautils.py:44-81).
* ``tdvp``       -- 1-site / 2-site TDVP sweeps in two gauge modes
                    (reference algorithms/tdvp.py:12-188).
* ``exact``      -- exact state-vector evolution (reference algorithms/exact.py).

Gauge modes of ``tdvp.TDVPOracle``:

* ``"reference"`` replays the reference's operation order exactly, including
  the per-step ``make_site_canonical(0)`` re-QR that leaves right environments
  stale and the gauge side effect of ``measure``.  Only reproducible with the
  same LAPACK (numpy) factorisations, so this mode always uses numpy's QR/SVD.
* ``"consistent"`` keeps the environments consistent with the tensors: no
  redundant re-canonicalisation at the start of a step and a measurement that
  does not move the orthogonality centre.  This is the physically correct
  projector-splitting integrator; at full bond dimension it reproduces exact
  evolution.  The GPU implementation follows this mode.

Parity pins (what checks each function; all in tests/test_oracle.py):

* automaton: golden MPO tensors from the reference for every (d, a, b), and the
  closed-form Hamiltonian vs the MPO contracted to a dense matrix (N <= 9).
* exact: golden Exact trajectories (N = 6..10), the analytic 3-level case,
  dense eigh vs Krylov, norm conservation.
* local: golden single-call vectors of the reference's local operations
  (environment updates, 1-site and 2-site evolution, split/truncate, bond
  evolution), and dense expm vs Krylov expm vs scipy.linalg.expm.
* tdvp reference mode: golden trajectories of the reference (full-chi N=8 runs,
  short truncating runs, and config 1 in full).
* tdvp consistent mode: step-1 rows of every golden trajectory (both modes agree
  on the first step from a product state), and full-rank runs vs exact
  (2, 3 and 7 cells, all three algorithms).
* tdvp adaptive targets (a1tdvp): golden reference runs replayed in reference
  mode, bond targets included.
* planted mistakes (transposed MPO indices, wrong propagator sign, wrong
  neighbour window) fail these pins.
"""
