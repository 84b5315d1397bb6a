"""MPS helpers restated from reference tensor_networks/mps.py (oracle).

An MPS is a Python list of complex128 arrays A[p, a, b] of shape
(2, chi_left, chi_right), chi_0 = chi_N = 1.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import logm

from .automaton import P1


def product_state(plist) -> list[np.ndarray]:
    """Bond-dimension-1 MPS with amplitudes (sqrt(1-p), sqrt(p)) per cell
    (reference mps.py:34-52)."""
    out = []
    for p in plist:
        A = np.zeros((2, 1, 1), dtype=complex)
        A[0, 0, 0] = (1.0 - p) ** 0.5
        A[1, 0, 0] = p ** 0.5
        out.append(A)
    return out


def truncate_pad(T: np.ndarray, shape) -> np.ndarray:
    """Slice each axis to ``shape`` and zero-pad where ``T`` is smaller
    (reference mps.py:226-236)."""
    sl = tuple(slice(0, min(a, b)) for a, b in zip(T.shape, shape))
    if all(a >= b for a, b in zip(T.shape, shape)):
        return T[sl]
    out = np.zeros(shape, dtype=T.dtype)
    out[sl] = T[sl]
    return out


def left_qr(A: np.ndarray, reduced: bool = True):
    """QR of A reshaped to (2*chi_l, chi_r); Q back to (2, chi_l, k)
    (reference mps.py:83-88).  numpy/LAPACK zgeqrf+zungqr conventions."""
    p, l, r = A.shape
    Q, R = np.linalg.qr(A.reshape(p * l, r), mode="reduced" if reduced else "complete")
    return Q.reshape(p, l, -1), R


def right_qr(A: np.ndarray, reduced: bool = True):
    """QR of A^T (over the right bond); returns Q as (2, k, chi_r) and R^T
    (reference mps.py:90-98)."""
    Q, R = left_qr(np.transpose(A, (0, 2, 1)), reduced)
    return np.transpose(Q, (0, 2, 1)), R.T


def orthonormalize_left(As: list, i: int) -> None:
    """Left-isometrise site i, push R into site i+1 (reference mps.py:146-164)."""
    shape = As[i].shape
    Q, R = left_qr(As[i])
    Q = truncate_pad(Q, shape)
    R = truncate_pad(R, (shape[2], As[i + 1].shape[1]))
    As[i + 1] = np.einsum("xb,pbc->pxc", R, As[i + 1])
    As[i] = Q


def orthonormalize_right(As: list, i: int) -> None:
    """Right-isometrise site i, push R into site i-1 (reference mps.py:166-181)."""
    shape = As[i].shape
    Q, R = right_qr(As[i])
    Q = truncate_pad(Q, shape)
    R = truncate_pad(R, (As[i - 1].shape[2], shape[1]))
    As[i - 1] = np.einsum("pab,bx->pax", As[i - 1], R)
    As[i] = Q


def make_site_canonical(As: list, i: int) -> None:
    """Full left sweep up to i then full right sweep down to i
    (reference mps.py:183-192)."""
    for j in range(i):
        orthonormalize_left(As, j)
    for j in range(len(As) - 1, i, -1):
        orthonormalize_right(As, j)


def bond_dims(As: list) -> np.ndarray:
    """Bond dimensions as MPS.measure records them (reference
    tensor_networks/mps.py:106-109)."""
    return np.array([A.shape[1] for A in As] + [As[-1].shape[2]], dtype=float)


def _entropy_bits(rho: np.ndarray) -> float:
    """-Tr(rho log2 rho) through scipy.linalg.logm, as reference mps.py:135-140."""
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return float((-np.trace(rho @ (logm(rho) / np.log(2)))).real)


def measure_reference(As: list):
    """The reference's MPS.measure (mps.py:100-140): records bond dims, then
    moves the centre to site 0 and walks it to site N-1 with left QRs,
    reading rho at each centre.  MUTATES the gauge of ``As``.

    Returns (population, d_population, single_site_entropy, bond_dims).
    Population is Re rho[1,1], not divided by the norm.
    """
    n = len(As)
    bd = bond_dims(As)
    make_site_canonical(As, 0)
    pop = np.zeros(n)
    sse = np.zeros(n)
    for site in range(n):
        if site > 0:
            orthonormalize_left(As, site - 1)
        A = As[site]
        rho = np.einsum("pab,qab->pq", A, A.conj())
        pop[site] = np.sum(rho * P1).real
        sse[site] = _entropy_bits(rho)
    return pop, np.round(pop), sse, bd


def rho_consistent(As: list) -> np.ndarray:
    """Single-site density matrices (N, 2, 2) without moving the orthogonality
    centre: the rho the reference reads at each centre site
    (tensor_networks/mps.py:118-132).

    Requires the TDVP invariant that sites 1..N-1 are right-isometric (centre
    at site 0): then rho_i = sum E_{i-1}[a,a'] A_i[p,a,b] conj(A_i[p',a',b])
    with the left transfer E_i = sum_p A_i^T E_{i-1} conj(A_i), E_{-1} = 1.
    This equals the reference's centre-site rho for every site.
    """
    n = len(As)
    out = np.zeros((n, 2, 2), dtype=complex)
    E = np.ones((1, 1), dtype=complex)
    for site in range(n):
        A = As[site]
        EA = np.einsum("xy,pxb->pyb", E, A, optimize=True)
        out[site] = np.einsum("pyb,qyb->pq", EA, A.conj(), optimize=True)
        E = np.einsum("pyb,pyc->bc", EA, A.conj(), optimize=True)
    return out


def measure_consistent(As: list):
    """Same observables as reference tensor_networks/mps.py:100-140 without
    moving the orthogonality centre (densities from rho_consistent)."""
    n = len(As)
    bd = bond_dims(As)
    rho = rho_consistent(As)
    pop = np.array([rho[i, 1, 1].real for i in range(n)])
    sse = np.array([_entropy_bits(rho[i]) for i in range(n)])
    return pop, np.round(pop), sse, bd


def as_vector(As: list) -> np.ndarray:
    """Contract to the 2^N vector, site 0 = most significant bit
    (reference mps.py:194-208)."""
    psi = As[0]
    for A in As[1:]:
        psi = np.einsum("xab,pbc->xpac", psi, A)
        s = psi.shape
        psi = psi.reshape(s[0] * s[1], s[2], s[3])
    return np.trace(psi, axis1=1, axis2=2)


def from_vector(psi: np.ndarray) -> list[np.ndarray]:
    """Left-canonical MPS of a 2^N vector by successive reduced QRs
    (reference mps.py:54-73)."""
    out = []
    M = np.asarray(psi, dtype=complex).reshape(2, -1)
    while M.shape[1] > 1:
        Q, R = np.linalg.qr(M)
        out.append(np.transpose(Q.reshape(-1, 2, R.shape[0]), (1, 0, 2)))
        M = R.reshape(R.shape[0] * 2, -1)
    out.append(np.transpose(M.reshape(-1, 2, 1), (1, 0, 2)))
    return out
