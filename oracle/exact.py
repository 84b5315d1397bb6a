"""Exact state-vector evolution restated from reference algorithms/exact.py (oracle).

The reference forms the dense 2^N propagator U = exp(-i (pi/2) dt H) by eigh of
the MPO contracted to a matrix (exact.py:15-17, mpo.py:219-228) and applies
psi <- U psi at every plot step with dt scaled by the plot interval
(quantum_game.py:70-73, 116-117).  Two evaluation modes:

* ``dense``   -- eigh of the closed-form H (N <= ~12), the reference's arithmetic;
* ``krylov``  -- matrix-free closed-form H application + Lanczos expm, for any N
                 that fits in host memory.
"""
from __future__ import annotations

import numpy as np

from . import automaton as AU
from . import local as loc
from . import mps as M


class ExactOracle:
    """State-vector evolution of reference algorithms/exact.py:9-28."""

    def __init__(self, As, ncells: int, distance: int, interval: range, dt: float,
                 mode: str = "dense", tol: float = 1e-14):
        self.n, self.d, self.interval = ncells, distance, interval
        self.dt = dt
        self.mode = mode
        self.tol = tol
        self.psi = M.as_vector(As)
        if mode == "dense":
            H = AU.dense_h(ncells, distance, interval)
            w, V = np.linalg.eigh(H)
            self.U = (np.exp(-1j * loc.HALF_PI * dt * w) * V) @ V.conj().T
        elif mode != "krylov":
            raise ValueError(mode)

    def do_time_step(self) -> None:
        """psi <- exp(-i (pi/2) dt H) psi (reference algorithms/exact.py:27-28)."""
        if self.mode == "dense":
            self.psi = self.U @ self.psi
        else:
            self.psi, _ = loc.expm_krylov(
                lambda v: AU.apply_h(v, self.n, self.d, self.interval), self.psi, self.dt,
                tol=self.tol, max_iter=200)

    def populations(self) -> np.ndarray:
        """<n_i> = sum_x |psi_x|^2 x_i (marginals of |psi|^2): the population
        MPS.measure (reference tensor_networks/mps.py:100-132) reads from
        MPS.from_vector(psi) (algorithms/exact.py:20-21)."""
        p = np.abs(self.psi) ** 2
        x = np.arange(1 << self.n)
        return np.array([p[(x >> (self.n - 1 - i)) & 1 == 1].sum() for i in range(self.n)])


def run_exact(As, ncells, distance, interval, dt, steps, plot_interval, mode="dense"):
    """quantum_game.py:70-73, 82-117 for the exact branch: one propagation of
    dt * interval per plot step, measured before each propagation."""
    ex = ExactOracle(As, ncells, distance, interval, dt * plot_interval, mode=mode)
    rows = steps // plot_interval + (1 if steps % plot_interval else 0)
    pop = np.zeros((rows, ncells))
    for r in range(rows):
        pop[r] = ex.populations()
        ex.do_time_step()
    return pop
