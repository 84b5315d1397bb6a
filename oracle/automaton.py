"""Rule Hamiltonian: MPO from a counting automaton, and the closed form (oracle).

The reference builds H = sum_i sigma^x_i * Pi_i, where Pi_i projects onto the
configurations in which the number of alive cells among the neighbours within
``distance`` of cell i lies in ``activation_interval`` (a Python ``range``).
Cells beyond the chain ends count as dead.

Restated from reference tensor_networks/mpo.py:54-200 (automaton and MPO) and
mpo.py:202-228 (dense matrix), with index conventions of constants.py:3-17.
"""
from __future__ import annotations

import numpy as np

P0 = np.array([[1.0, 0.0], [0.0, 0.0]])
P1 = np.array([[0.0, 0.0], [0.0, 1.0]])
SX = np.array([[0.0, 1.0], [1.0, 0.0]])
ID2 = np.eye(2)


def automaton_states(distance: int, interval: range):
    """Enumerate automaton states and edges in the reference's insertion order.

    A state is (ket0_count, ket1_count, s_placed).  Depth-first insertion with
    children visited in the order: sigma^x edge (only when exactly ``distance``
    projectors precede it), identity self-loop (initial state only), P1 edge
    (while ket1 + 1 < interval.stop), P0 edge (while interval.start is still
    reachable).  States whose path already holds 2*distance+1 operators map to
    the single final state, appended last.  (reference mpo.py:59-149)

    Returns (states, edges) with edges as (parent_index, child_index, op)
    where child_index -1 denotes the final state.
    """
    ops_per_path = 2 * distance
    states: list[tuple[int, int, bool]] = []
    edges: list[tuple[int, int, np.ndarray]] = []

    def visit(k0: int, k1: int, s: bool) -> int:
        """Index of state (k0, k1, s), inserted depth first on first visit
        (reference tensor_networks/mpo.py:90-150, find_or_insert_recursive)."""
        key = (k0, k1, s)
        if key in states:
            return states.index(key)
        if k0 + k1 + int(s) == ops_per_path + 1:
            return -1
        me = len(states)
        states.append(key)
        n_ops = k0 + k1
        if n_ops == distance and not s:
            edges.append((me, visit(k0, k1, True), SX))
            return me
        if n_ops == 0 and not s:
            edges.append((me, 0, ID2))
        if k1 + 1 < interval.stop:
            edges.append((me, visit(k0, k1 + 1, s), P1))
        if k1 + (ops_per_path - n_ops - 1) >= interval.start:
            edges.append((me, visit(k0 + 1, k1, s), P0))
        return me

    visit(0, 0, False)
    final = len(states)
    states.append((-1, -1, True))
    edges.append((final, -1, ID2))
    return states, edges


def bulk_tensor(distance: int, interval: range) -> np.ndarray:
    """Bulk MPO tensor W[p, p', w_left, w_right] (reference mpo.py:158-166).

    An automaton edge parent -> child contributes its 2x2 operator to
    W[:, :, child, parent]; the final state has the largest index.
    """
    states, edges = automaton_states(distance, interval)
    D = len(states)
    W = np.zeros((2, 2, D, D), dtype=complex)
    for parent, child, op in edges:
        W[:, :, child % D, parent] += op
    return W


def mpo_tensors(ncells: int, distance: int, interval: range) -> list[np.ndarray]:
    """Open-chain MPO: list of (2,2,Dl,Dr) tensors (reference mpo.py:158-200).

    The boundary vectors emulate ``distance`` dead cells beyond each end:
    left = e_final^T B^d and right = B^d e_initial with B = <0|W|0>.
    For ncells == 1 the reference assigns both boundaries to index 0
    (mpo.py:196-198); this restatement requires ncells >= 2.
    """
    if ncells < 2:
        raise ValueError("the rule Hamiltonian needs at least 2 cells")
    W = bulk_tensor(distance, interval)
    D = W.shape[2]
    B = W[0, 0, :, :]
    left = np.zeros((1, D))
    left[0, D - 1] = 1.0
    right = np.zeros((D, 1))
    right[0, 0] = 1.0
    for _ in range(distance):
        left = left @ B
        right = B @ right
    first = np.einsum("xw,pqwv->pqxv", left, W)
    last = np.einsum("pqwv,vy->pqwy", W, right)
    out = [W] * ncells
    out = list(out)
    out[0] = first
    out[-1] = last
    return out


def mpo_dense(Ws: list[np.ndarray]) -> np.ndarray:
    """Contract an MPO to its 2^N x 2^N matrix, site 0 most significant
    (reference mpo.py:202-228)."""
    H = Ws[0]
    for W in Ws[1:]:
        H = np.einsum("abxy,cdyz->acbdxz", H, W)
        s = H.shape
        H = H.reshape(s[0] * s[1], s[2] * s[3], s[4], s[5])
    return np.trace(H, axis1=2, axis2=3)


def neighbour_masks(ncells: int, distance: int) -> list[int]:
    """Bit mask of the neighbours of every cell; cell i is bit (N-1-i).

    Neighbours are the cells within ``distance`` on the open chain
    (reference parameters/rules.py:2-6, the rule the MPO builder
    tensor_networks/mpo.py:159-200 encodes)."""
    masks = []
    for i in range(ncells):
        m = 0
        for j in range(max(0, i - distance), min(ncells, i + distance + 1)):
            if j != i:
                m |= 1 << (ncells - 1 - j)
        masks.append(m)
    return masks


def apply_h(psi: np.ndarray, ncells: int, distance: int, interval: range) -> np.ndarray:
    """Closed-form (H psi)[x] = sum_i [popcount(x & nbr_i) in interval] psi[x ^ bit_i].

    The operator the rule MPO of reference tensor_networks/mpo.py:159-200
    encodes, which the reference's exact path densifies with MPO.as_matrix
    (mpo.py:219-228, algorithms/exact.py:17).

    The flip condition does not depend on bit i itself, so H is real symmetric.
    Vectorised over the 2^N basis (cheap for N <= ~24 on the host).
    """
    dim = 1 << ncells
    assert psi.shape[-1] == dim
    x = np.arange(dim, dtype=np.int64)
    out = np.zeros_like(psi)
    for i, m in enumerate(neighbour_masks(ncells, distance)):
        cnt = _popcount(x & m)
        active = np.isin(cnt, np.array(list(interval), dtype=np.int64))
        bit = 1 << (ncells - 1 - i)
        src = x ^ bit
        out[..., active] += psi[..., src[active]]
    return out


def _popcount(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64)
    c = np.zeros_like(v)
    while np.any(v):
        c += v & np.uint64(1)
        v >>= np.uint64(1)
    return c.astype(np.int64)


def dense_h(ncells: int, distance: int, interval: range) -> np.ndarray:
    """Dense closed-form Hamiltonian (small N only): the reference's
    MPO.as_matrix (tensor_networks/mpo.py:219-228) without the MPO."""
    dim = 1 << ncells
    return apply_h(np.eye(dim, dtype=complex), ncells, distance, interval)
