"""Initial states (reference states.py:11-88) plus seeded generators.

The reference's named states take the cell count from the parsed arguments;
here every generator takes ``ncells`` (and ``distance`` where the reference
uses it) explicitly.  ``rand`` draws from a seeded numpy Generator instead of
the unseeded ``random.random`` (reference states.py:86-88).  ``padded`` builds
the bond-padded product state of the 1TDVP benchmark: a product state whose
site tensors are embedded in (2, chi_l, chi_r) with every other entry a small
complex Gaussian, so that all bonds are full rank before TDVP canonicalises
it (a bare product state leaves the 1TDVP padding subspace to rounding).
"""
from __future__ import annotations

import numpy as np

from tensor_networks import MPS


def _product(plist) -> MPS:
    return MPS.from_density_distribution(plist)


def blinker(ncells: int, width: int = 1) -> MPS:
    """Cells mid-width and mid+width alive (reference states.py:11-15)."""
    p = [0.0] * ncells
    mid = int(ncells / 2)
    p[mid - width] = p[mid + width] = 1.0
    return _product(p)


def triple_blinker(ncells: int) -> MPS:
    """Cells mid-2, mid and mid+2 alive (reference states.py:18-22)."""
    p = [0.0] * ncells
    mid = int(ncells / 2)
    p[mid - 2] = p[mid] = p[mid + 2] = 1.0
    return _product(p)


def three_adjacent(ncells: int) -> MPS:
    """Cells mid-1, mid, mid+1 alive (the 41-cell 2TDVP benchmark state)."""
    p = [0.0] * ncells
    mid = int(ncells / 2)
    p[mid - 1] = p[mid] = p[mid + 1] = 1.0
    return _product(p)


def full_blinker(ncells: int) -> MPS:
    """Every odd cell alive (reference states.py:25-31)."""
    return _product([float(i % 2) for i in range(ncells)])


def single(ncells: int, position: int | None = None) -> MPS:
    """One alive cell, the middle one by default (reference states.py:34-37)."""
    p = [0.0] * ncells
    p[int(ncells / 2) if position is None else position] = 1.0
    return _product(p)


def single_bottom(ncells: int) -> MPS:
    """Cell 0 alive (reference states.py:40-41)."""
    return single(ncells, 0)


def all_ket_0(ncells: int) -> MPS:
    """Every cell dead (reference states.py:44-46)."""
    return _product([0.0] * ncells)


def all_ket_1(ncells: int) -> MPS:
    """Every cell alive (reference states.py:49-51)."""
    return _product([1.0] * ncells)


def only_outer(ncells: int) -> MPS:
    """The two end cells alive (reference states.py:54-57)."""
    p = [0.0] * ncells
    p[0] = p[-1] = 1.0
    return _product(p)


def all_ket_1_but_outer(ncells: int, distance: int) -> MPS:
    """All alive except `distance` dead cells at each end (reference states.py:60-64)."""
    return _product([0.0] * distance + [1.0] * (ncells - 2 * distance) + [0.0] * distance)


def equal_superposition(ncells: int) -> MPS:
    """p = 1/2 on every cell (reference states.py:67-69)."""
    return _product([0.5] * ncells)


def equal_superposition_but_outer(ncells: int, distance: int) -> MPS:
    """p = 1/2 except `distance` dead cells at each end (reference states.py:72-76)."""
    return _product([0.0] * distance + [0.5] * (ncells - 2 * distance) + [0.0] * distance)


def gradient(ncells: int, reverse: bool = False) -> MPS:
    """p_i = sin(pi i / (N-1) / 2), optionally reversed (reference states.py:79-83)."""
    p = [np.sin(np.pi * i / (ncells - 1) / 2) for i in range(ncells)]
    return _product(list(reversed(p)) if reverse else p)


def rand_plist(ncells: int, p: float = 0.5, seed: int = 0) -> list:
    """Each cell alive with probability p, from a seeded generator (the reference's
    states.py:86-88 draws from the unseeded `random` module)."""
    rng = np.random.default_rng(seed)
    return [1.0 if x < p else 0.0 for x in rng.random(ncells)]


def rand(ncells: int, p: float = 0.5, seed: int = 0) -> MPS:
    """Product state of rand_plist (reference states.py:86-88)."""
    return _product(rand_plist(ncells, p, seed))


def padded(plist, chi_max: int, scale: float = 1e-8, seed: int = 0) -> MPS:
    """Product state embedded at bond dims min(2^i, 2^(N-i), chi_max) with
    complex Gaussian filler of size ``scale`` in every other entry."""
    n = len(plist)
    rng = np.random.default_rng(seed)
    dims = [min(2 ** min(i, 62), 2 ** min(n - i, 62), chi_max) for i in range(n + 1)]
    out = []
    for i, p in enumerate(plist):
        shape = (2, dims[i], dims[i + 1])
        A = scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
        A[:, 0, 0] = ((1.0 - p) ** 0.5, p ** 0.5)
        out.append(A)
    return MPS.from_tensors(out)


NAMED = {
    "blinker": lambda n, d: blinker(n), "triple_blinker": lambda n, d: triple_blinker(n),
    "full_blinker": lambda n, d: full_blinker(n), "single": lambda n, d: single(n),
    "single_bottom": lambda n, d: single_bottom(n), "all_ket_0": lambda n, d: all_ket_0(n),
    "all_ket_1": lambda n, d: all_ket_1(n), "only_outer": lambda n, d: only_outer(n),
    "all_ket_1_but_outer": all_ket_1_but_outer, "equal_superposition": lambda n, d: equal_superposition(n),
    "equal_superposition_but_outer": equal_superposition_but_outer, "gradient": lambda n, d: gradient(n),
    "rand": lambda n, d: rand(n),
}
