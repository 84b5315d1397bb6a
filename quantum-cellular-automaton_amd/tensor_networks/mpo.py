"""Rule Hamiltonian as an MPO (reference tensor_networks/mpo.py:54-228).

H = sum_i sigma^x_i * [number of alive cells within ``distance`` of i lies in
``activation_interval``], cells beyond the ends counted dead.  The MPO is built
from a counting automaton whose states are (#P0 placed, #P1 placed, sigma^x
placed); the state order (depth-first insertion, final state last) fixes the
MPO bond index and is kept identical to the reference so that W tensors are
interchangeable.  This is host-side setup: it runs once per simulation.
"""
from __future__ import annotations

import numpy as np

from constants import IDENTITY, KET_0, PROJECTION_KET_0, PROJECTION_KET_1, S_OPERATOR
from parameters import Rules


def _automaton(distance: int, interval: range):
    """States in insertion order and edges (src, dst, op); dst None = final state."""
    total = 2 * distance
    order: dict[tuple, int] = {}
    edges: list[tuple[int, object, np.ndarray]] = []

    def visit(k0: int, k1: int, s: bool):
        key = (k0, k1, s)
        if key in order:
            return order[key]
        if k0 + k1 + s == total + 1:
            return None                       # path complete: the final state
        me = order[key] = len(order)
        placed = k0 + k1
        if placed == distance and not s:      # the flip operator sits in the middle
            edges.append((me, visit(k0, k1, True), S_OPERATOR))
            return me
        if placed == 0 and not s:
            edges.append((me, me, IDENTITY))  # initial state loops on identities
        if k1 + 1 < interval.stop:
            edges.append((me, visit(k0, k1 + 1, s), PROJECTION_KET_1))
        if k1 + total - placed - 1 >= interval.start:
            edges.append((me, visit(k0 + 1, k1, s), PROJECTION_KET_0))
        return me

    visit(0, 0, False)
    return len(order), edges


class MPO(object):
    """Tensors W[i] of shape (2, 2, Dl, Dr) (reference tensor_networks/mpo.py:152-157)."""

    W: list = []

    def __init__(self, Wlist: list) -> None:
        self.W = Wlist

    @classmethod
    def hamiltonian_from_rules(cls, rules: Rules) -> "MPO":
        """The rule Hamiltonian: the state automaton's transitions as the bulk
        tensor, closed at both ends by `distance` dead cells (reference
        mpo.py:158-200).  W tensors equal the reference's entry for entry."""
        if rules.ncells < 2:
            raise ValueError("the rule Hamiltonian needs at least two cells")
        nstates, edges = _automaton(rules.distance, rules.activation_interval)
        D = nstates + 1                       # + final state
        bulk = np.zeros((2, 2, D, D), dtype=complex)
        for src, dst, op in edges:
            bulk[:, :, D - 1 if dst is None else dst, src] += op
        bulk[:, :, D - 1, D - 1] += IDENTITY  # final state loops on identities
        # open boundaries emulate `distance` dead cells beyond each end
        B = np.einsum("p,q,pqxy->xy", KET_0, KET_0, bulk)
        left = np.zeros(D)
        left[D - 1] = 1.0
        right = np.zeros(D)
        right[0] = 1.0
        for _ in range(rules.distance):
            left = left @ B
            right = B @ right
        Ws = [bulk] * rules.ncells
        Ws[0] = np.einsum("x,pqxy->pqy", left, bulk)[:, :, None, :]
        Ws[-1] = np.einsum("pqxy,y->pqx", bulk, right)[:, :, :, None]
        return cls(Ws)

    @classmethod
    def merge_mpo_tensor_pair(cls, W0, W1) -> np.ndarray:
        """Two neighbouring tensors as one, physical indices paired (reference
        mpo.py:202-217)."""
        W = np.einsum("abxm,cdmy->acbdxy", W0, W1)
        s = W.shape
        return W.reshape(s[0] * s[1], s[2] * s[3], s[4], s[5])

    def as_matrix(self) -> np.ndarray:
        """Dense 2^N matrix, site 0 most significant (small N only; reference
        mpo.py:219-228)."""
        H = self.W[0]
        for W in self.W[1:]:
            H = self.merge_mpo_tensor_pair(H, W)
        return np.trace(H, axis1=2, axis2=3)
