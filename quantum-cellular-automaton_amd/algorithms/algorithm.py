"""The Algorithm interface every time-evolution backend implements
(reference algorithms/algorithm.py:10-63).

Same constructor, ``do_time_step``, ``psi`` property and ``measure`` signature
as the reference, so drivers written against it run unchanged.  ``measure``
writes into caller-owned float64 row views: population <n_i>, its rounding,
the single-site entropy and the N+1 bond dimensions.
"""
from abc import ABC, abstractmethod

import numpy as np

from parameters import Rules


def entropy_bits(rho: np.ndarray) -> float:
    """-Tr(rho log2 rho) of a 2x2 density matrix from its eigenvalues
    (0 log 0 = 0); the reference takes logm of rho (tensor_networks/mps.py:135-140)."""
    w = np.linalg.eigvalsh(0.5 * (rho + rho.conj().T))
    w = w[w > 0]
    return float(-(w * np.log2(w)).sum())


def write_observables(rho: np.ndarray, bond_dims_now, population, d_population,
                      single_site_entropy, bond_dims) -> None:
    """Fill the reference's measure() outputs from per-site density matrices
    (reference mps.py:100-140).  <n_i> = Re rho_i[1][1], not normalised."""
    for i in range(rho.shape[0]):
        p = rho[i, 1, 1].real
        population[i] = p
        d_population[i] = np.round(p)
        single_site_entropy[i] = entropy_bits(rho[i])
    for i, b in enumerate(bond_dims_now):
        bond_dims[i] = b


class Algorithm(ABC):
    """Time-evolution backend (reference algorithms/algorithm.py:10-63)."""

    _H = None
    args = None

    @abstractmethod
    def __init__(self, psi_0, H, args) -> None:
        """Initial state, rule MPO and parsed arguments (reference algorithm.py:14-18)."""
        self.psi = psi_0
        self._H = H
        self.args = args

    @abstractmethod
    def do_time_step(self) -> None:
        """Advance the state by one step (reference algorithm.py:20-22)."""

    @property
    @abstractmethod
    def psi(self):
        """The current state as an MPS (reference algorithm.py:24-27)."""

    @psi.setter
    @abstractmethod
    def psi(self, value) -> None:
        """Replace the current state (reference algorithm.py:29-32)."""

    @classmethod
    def classical_evolution(cls, first_column: np.ndarray, rules: Rules, plot_steps: int) -> np.ndarray:
        """Classical automaton overlay (reference algorithm.py:34-56): a cell flips
        when its alive-neighbour count is in the activation interval."""
        heat = np.zeros([plot_steps, len(first_column)])
        heat[0, :] = first_column
        n = len(first_column)
        for step in range(1, plot_steps):
            prev = heat[step - 1]
            for site in range(n):
                lo, hi = max(0, site - rules.distance), min(n, site + rules.distance + 1)
                count = prev[lo:hi].sum() - prev[site]
                heat[step, site] = 1.0 - prev[site] if count in rules.activation_interval else prev[site]
        return heat

    @abstractmethod
    def measure(self, population, d_population, single_site_entropy, bond_dims) -> None:
        """Write <n_i>, its rounding, the single-site entropy and the bond
        dimensions into the given rows (reference algorithm.py:58-63)."""
