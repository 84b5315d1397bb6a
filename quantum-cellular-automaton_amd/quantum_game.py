"""Driver with the reference's step/measure cadence and output files
(reference quantum_game.py:16-146), minus plotting.

Kept: measure at steps 0, interval, 2*interval, ... before stepping; CSV files
data/{name}-population.csv, -d_population.csv, -single_site_entropy.csv and
-bond_dims.csv rewritten at every plot step with rows [:plot_step] (so the
newest row is written at the next plot step, as in the reference); the npz
checkpoint data/{name}.npz of the current MPS; the exact algorithm propagating
dt * interval per plot step.
Dropped on purpose: plots (presentation), and the reference's in-place
``args.step_size *= interval`` that compounds over several initial states
(quantum_game.py:70-73) -- here the scaled step is local to each state.
"""
from __future__ import annotations

import copy
import logging
import os

import numpy as np

import states
from algorithms import Exact, TDVP
from parameters import Parser
from tensor_networks import MPO, MPS


class QuantumGame(object):
    """Runs every initial state and writes the reference's output files
    (reference quantum_game.py:16-146)."""

    def __init__(self, args: Parser) -> None:
        """Rule MPO and the named or file initial states (reference quantum_game.py:18-26)."""
        logging.basicConfig(level=logging.INFO)
        self.args = args
        self.H = MPO.hamiltonian_from_rules(args.rules)
        n, d = args.rules.ncells, args.rules.distance
        self.initial_states = ([states.NAMED[name](n, d) for name in args.initial_states] +
                               [MPS.from_file(f) for f in args.initial_state_files])

    def file_name(self, state_name: str) -> str:
        """Output file stem (reference quantum_game.py:49-58)."""
        a = self.args
        iv = a.rules.activation_interval
        name = (f"{state_name}{a.rules.ncells}_{a.rules.distance}{iv.start}{iv.stop}_"
                f"{str(a.step_size).replace('.', '')}_{a.algorithm}")
        if a.algorithm != "exact":
            name += f"_{a.max_bond_dim}"
            if a.algorithm == "a1tdvp":
                name += f"_{a.approximative_evolution_method}"
                if a.approximative_evolution_method == "taylor":
                    name += f"_{a.taylor_steps}"
        return name

    def run_state(self, index: int, initial_state, data_dir: str = "data") -> dict:
        """One state: measure every plot step, evolve between them, write the
        CSV and npz outputs (reference quantum_game.py:30-146)."""
        args = self.args
        n = args.rules.ncells
        rows = args.plot_steps
        population = np.zeros([rows, n])
        d_population = np.zeros([rows, n])
        sse = np.zeros([rows, n])
        bond_dims = np.zeros([rows, n + 1])
        state_name = args.initial_states[index] if index < len(args.initial_states) else str(index)
        name = self.file_name(state_name)
        local = copy.copy(args)
        if args.algorithm == "exact":
            local.step_size = args.step_size * args.plot_step_interval
            algorithm = Exact(psi_0=initial_state, H=self.H, args=local)
        elif args.algorithm in ("1tdvp", "2tdvp", "a1tdvp"):
            algorithm = TDVP(psi_0=initial_state, H=self.H, args=local)
        else:
            raise NotImplementedError(f"algorithm {args.algorithm!r} is not provided")
        os.makedirs(data_dir, exist_ok=True)
        logging.info(f"Running {name}")
        for step in range(args.num_steps):
            if step % args.plot_step_interval == 0:
                ps = step // args.plot_step_interval
                algorithm.measure(population=population[ps], d_population=d_population[ps],
                                  single_site_entropy=sse[ps], bond_dims=bond_dims[ps])
                for tag, arr in (("population", population), ("d_population", d_population),
                                 ("single_site_entropy", sse), ("bond_dims", bond_dims)):
                    np.savetxt(os.path.join(data_dir, f"{name}-{tag}.csv"), arr[:ps, :], delimiter=",")
                algorithm.psi.write_to_file(os.path.join(data_dir, f"{name}.npz"))
                if args.algorithm == "exact":
                    algorithm.do_time_step()
            if args.algorithm != "exact":
                algorithm.do_time_step()
        logging.info(f"Finished {name}")
        return dict(population=population, d_population=d_population, single_site_entropy=sse,
                    bond_dims=bond_dims, name=name)

    def start(self, data_dir: str = "data") -> list:
        """Every initial state in order (reference quantum_game.py:28-30)."""
        return [self.run_state(i, s, data_dir) for i, s in enumerate(self.initial_states)]
