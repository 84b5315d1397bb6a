"""Synthetic multi-instance workloads: independent simulations drawn from a
seeded generator, for batched and served runs.

The reference runs one simulation per process (quantum_game.py:75-119 loops
over its initial states one after another); these generators describe many
independent runs of the same TDVP step (reference algorithms/tdvp.py:50-63)
with the reference's own parameters (parameters/parser.py flags: number of
cells, distance, activation interval, max bond dimension, step size, SVD
epsilon).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Instance:
    ncells: int
    chi_max: int
    distance: int
    interval: range              # closed [a, b] as range(a, b + 1)
    plist: list = field(repr=False)
    algorithm: str = "2tdvp"
    step_size: float = 0.1
    svd_epsilon: float = 1e-10

    @property
    def site_updates_per_step(self) -> int:
        """Two-site block updates per 2TDVP step (2(N-1)); 1-site updates per
        1TDVP step (2N)."""
        return 2 * (self.ncells - 1) if self.algorithm == "2tdvp" else 2 * self.ncells


def draw_batch_instance(rng: np.random.Generator) -> Instance:
    """One instance of the batched 2TDVP workload, drawn in this order:
    N ~ U{15..61}; chi_max ~ {16, 32, 64, 128}; d ~ U{1, 2}; a ~ U{1..2d};
    b ~ U{a..2d}; cells alive with probability 0.3."""
    n = int(rng.integers(15, 62))
    chi = int(rng.choice([16, 32, 64, 128]))
    d = int(rng.integers(1, 3))
    a = int(rng.integers(1, 2 * d + 1))
    b = int(rng.integers(a, 2 * d + 1))
    plist = [1.0 if x < 0.3 else 0.0 for x in rng.random(n)]
    return Instance(n, chi, d, range(a, b + 1), plist)


def batch_instances(count: int = 256, seed: int = 4) -> list:
    """The batched workload: `count` independent 2TDVP instances (dt 0.1,
    eps 1e-10) from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return [draw_batch_instance(rng) for _ in range(count)]


def serving_request(rng: np.random.Generator, big_fraction: float = 0.1) -> tuple:
    """One request of the served workload: with probability 1 - big_fraction
    a batched-workload instance, else the large 2TDVP case (N=101, chi 256,
    d=2, interval [2,3]); steps ~ U{50..500}; <n_i> returned every 10 steps.
    Returns (instance, steps)."""
    if rng.random() < big_fraction:
        plist = [1.0 if x < 0.3 else 0.0 for x in rng.random(101)]
        inst = Instance(101, 256, 2, range(2, 4), plist)
    else:
        inst = draw_batch_instance(rng)
    steps = int(rng.integers(50, 501))
    return inst, steps


def mpo_tensors(inst: Instance) -> list:
    from parameters import Rules
    from tensor_networks import MPO
    return MPO.hamiltonian_from_rules(Rules(inst.ncells, inst.interval, inst.distance, False)).W


def initial_tensors(inst: Instance) -> list:
    from tensor_networks import MPS
    return MPS.from_density_distribution(inst.plist).A
