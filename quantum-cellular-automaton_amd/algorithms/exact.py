"""Exact state-vector evolution on the GPU (drop-in for reference algorithms/exact.py:9-28).

The reference forms the dense 2^N x 2^N propagator; here the 2^N amplitudes
live in a device buffer and every step applies exp(-i (pi/2) dt H) by a Krylov
exponential whose matrix-vector product is the closed-form rule Hamiltonian
(one bit flip per active cell), so N is bounded by memory for vectors, not
matrices.  ``args.step_size`` is the propagation time per ``do_time_step``;
the reference driver scales it by the plot interval before construction
(quantum_game.py:70-73) and that contract is kept.
"""
from __future__ import annotations

import ctypes

import numpy as np

from qca_native import call, ptr
from tensor_networks import MPS

from .algorithm import Algorithm, write_observables


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the GPU exact evolution needs a HIP device")
    return torch


class Exact(Algorithm):
    """Exact 2^N state-vector evolution (reference algorithms/exact.py:9-28)."""

    def __init__(self, psi_0, H, args) -> None:
        """The rule comes from ``args.rules`` (distance, activation interval), as
        the MPO H was built from it (reference exact.py:15-17); the state is held
        on the device."""
        self._psi_dev = None
        self._n = len(psi_0.A)
        rules = args.rules
        self._d = int(rules.distance)
        self._lo, self._hi = int(rules.activation_interval.start), int(rules.activation_interval.stop)
        super().__init__(psi_0, H, args)
        self.tol = float(getattr(args, "krylov_tol", 0.0) or 0.0) or 1e-14
        self.last_iters = 0

    @property
    def psi(self) -> MPS:
        """The state as an MPS (reference exact.py:19-21)."""
        return MPS.from_vector(self.vector())

    @psi.setter
    def psi(self, value) -> None:
        """Upload value.as_vector() (reference exact.py:23-25)."""
        torch = _torch()
        v = np.ascontiguousarray(value.as_vector(), dtype=np.complex128)
        dev = int(getattr(getattr(self, "args", None), "device", 0) or 0)
        self._psi_dev = torch.from_numpy(v).to(f"cuda:{dev}")

    def vector(self) -> np.ndarray:
        """The state vector on the host (site 0 most significant, as MPS.as_vector)."""
        return self._psi_dev.cpu().numpy()

    def do_time_step(self) -> None:
        """psi <- exp(-i (pi/2) dt H) psi on the device (reference exact.py:27-28
        applies the dense propagator of exact.py:17)."""
        iters = ctypes.c_int(0)
        call("qca_exact_expm", ptr(self._psi_dev), self._n, self._d, self._lo, self._hi,
             float(self.args.step_size), self.tol, ctypes.byref(iters), None)
        self.last_iters = iters.value

    def rho(self) -> np.ndarray:
        """Single-site density matrices (N, 2, 2), the quantities MPS.measure
        reduces (reference tensor_networks/mps.py:118-132)."""
        out = np.zeros((self._n, 2, 2), dtype=np.complex128)
        call("qca_exact_rho", ptr(self._psi_dev), self._n, out.ctypes.data, None)
        return out

    def measure(self, population, d_population, single_site_entropy, bond_dims) -> None:
        """Observables of the exact state (reference algorithm.py:58-63 on
        MPS.from_vector(psi), mps.py:100-140)."""
        # bond dims of the exact state as an MPS (reference exact.py:19-21 -> from_vector)
        n = self._n
        bd = [min(2 ** i, 2 ** (n - i)) for i in range(n + 1)]
        write_observables(self.rho(), bd, population, d_population, single_site_entropy, bond_dims)
