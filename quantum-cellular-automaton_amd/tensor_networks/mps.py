"""MPS container at the host boundary (reference tensor_networks/mps.py).

Holds the site tensors A[p][a][b] (2, chi_l, chi_r) as complex128 numpy arrays
and the file/vector conversions the simulation loop uses.  Time evolution and
measurement of an MPS run on the GPU through ``algorithms.TDVP``; this module
only moves data in and out.
"""
from __future__ import annotations

import numpy as np


class MPS(object):
    """Tensors A[i] of shape (2, D[i], D[i+1]) (reference tensor_networks/mps.py:14-25)."""

    A: list = []

    def __init__(self, Alist: list) -> None:
        self.A = Alist

    @classmethod
    def from_tensors(cls, Alist) -> "MPS":
        """Complex copies of the given tensors (reference mps.py:27-32)."""
        return cls(Alist=[np.array(A, dtype=complex) for A in Alist])

    @classmethod
    def from_density_distribution(cls, plist) -> "MPS":
        """Bond-dimension-1 product state with amplitudes (sqrt(1-p), sqrt(p))
        per cell (reference mps.py:34-52)."""
        out = []
        for p in plist:
            A = np.zeros((2, 1, 1), dtype=complex)
            A[:, 0, 0] = ((1.0 - p) ** 0.5, p ** 0.5)
            out.append(A)
        return cls.from_tensors(out)

    @classmethod
    def from_vector(cls, psi) -> "MPS":
        """Left-canonical MPS of a 2^N vector by successive QRs (reference mps.py:54-73)."""
        out = []
        M = np.asarray(psi, dtype=complex).reshape(2, -1)
        while M.shape[1] > 1:
            Q, R = np.linalg.qr(M)
            out.append(np.transpose(Q.reshape(-1, 2, R.shape[0]), (1, 0, 2)))
            M = R.reshape(R.shape[0] * 2, -1)
        out.append(np.transpose(M.reshape(-1, 2, 1), (1, 0, 2)))
        return cls.from_tensors(out)

    @classmethod
    def from_file(cls, path: str) -> "MPS":
        """npz with arr_0 .. arr_{N-1} (reference mps.py:75-80); pickles are refused."""
        with np.load(path, allow_pickle=False) as f:
            return cls(Alist=[f[f"arr_{i}"] for i in range(len(f.files))])

    def write_to_file(self, path: str) -> None:
        """npz with arr_0 .. arr_{N-1} (reference mps.py:142-144)."""
        with open(path, "wb") as f:
            np.savez(f, *self.A)

    def as_vector(self) -> np.ndarray:
        """2^N vector, site 0 = most significant bit (reference mps.py:194-208)."""
        psi = self.A[0]
        for A in self.A[1:]:
            psi = np.einsum("xab,pbc->xpac", psi, A)
            s = psi.shape
            psi = psi.reshape(s[0] * s[1], s[2], s[3])
        return np.trace(psi, axis1=1, axis2=2)

    def bond_dims(self) -> list:
        """The N+1 bond dimensions as measure() records them (reference mps.py:106-109)."""
        return [A.shape[1] for A in self.A] + [self.A[-1].shape[2]]

    def is_valid_mps(self) -> bool:
        """Open boundaries and matching neighbour bonds (reference mps.py:216-225)."""
        if self.A[0].shape[1] != 1 or self.A[-1].shape[2] != 1:
            return False
        return all(self.A[i].shape[2] == self.A[i + 1].shape[1] for i in range(len(self.A) - 1))
