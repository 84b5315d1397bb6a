"""1-site and 2-site TDVP sweeps restated from reference algorithms/tdvp.py (oracle).

``TDVPOracle`` keeps the reference's data flow: the MPS tensors, the cached
left/right environment lists (tdvp.py:28-29) and the boundary environments
``ones((1,1,1))`` (tdvp.py:190-196).  The two gauge modes and the two
exponential modes are described in the package docstring.
"""
from __future__ import annotations

import numpy as np

from . import local as loc
from . import mps as M

ONE = np.ones((1, 1, 1), dtype=complex)


class TDVPOracle:
    """1-site, 2-site and adaptive 1-site TDVP of reference
    algorithms/tdvp.py:10-63 (construction, do_time_step) and its sweeps
    (tdvp.py:65-145), in the reference's or the gauge-consistent operation order."""

    def __init__(self, As, Ws, algorithm: str, dt: float, chi_max: int, eps: float,
                 gauge: str = "reference", expm: str = "dense", krylov_tol: float = 1e-14):
        if algorithm not in ("1tdvp", "2tdvp", "a1tdvp"):
            raise ValueError(algorithm)
        if gauge not in ("reference", "consistent"):
            raise ValueError(gauge)
        if expm not in ("dense", "lanczos"):
            raise ValueError(expm)
        self.A = [np.array(a, dtype=complex) for a in As]
        self.W = list(Ws)
        self.algorithm = algorithm
        self.dt = dt
        self.chi_max = chi_max
        self.eps = eps
        self.gauge = gauge
        self.expm = expm
        self.tol = krylov_tol
        self.krylov_iters: list[int] = []
        n = len(self.A)
        # tdvp.py:24-26
        M.make_site_canonical(self.A, n - 1)
        M.make_site_canonical(self.A, 0)
        self.L = [None] * n
        self.R = [None] * n
        # tdvp.py:31-35 (2^i saturates in Python; min() keeps it bounded)
        self.max_bond = [min(2 ** min(i, 62), 2 ** min(n - i, 62), chi_max) for i in range(n + 1)]
        self.target = list(self.max_bond)
        # tdvp.py:37-39: the reference re-canonicalises before every env build
        for site in range(n - 1, 0, -1):
            if gauge == "reference":
                M.make_site_canonical(self.A, site - 1)
            self._update_env("right", site)

    # ------------------------------------------------------------ environments
    def _get_env(self, side: str, site: int) -> np.ndarray:
        n = len(self.A)
        if (side == "left" and site < 0) or (side == "right" and site >= n):
            return ONE
        return self.L[site] if side == "left" else self.R[site]

    def _update_env(self, side: str, site: int) -> None:
        """tdvp.py:198-210"""
        if side == "left":
            self.L[site] = loc.env_left(self._get_env("left", site - 1), self.A[site], self.W[site])
        else:
            self.R[site] = loc.env_right(self._get_env("right", site + 1), self.A[site], self.W[site])

    # ------------------------------------------------------------ local evolution
    def _evolve(self, x: np.ndarray, L, R, W, delta: float) -> np.ndarray:
        """tdvp.py:349-365 (dense) or the matrix-free Krylov equivalent."""
        if self.expm == "dense":
            return loc.expm_dense(loc.dense_heff(L, R, W), x.reshape(-1), delta).reshape(x.shape)
        shape = x.shape
        y, it = loc.expm_krylov(lambda v: loc.heff_apply(v.reshape(shape), L, R, W).reshape(-1),
                                x.reshape(-1), delta, tol=self.tol)
        self.krylov_iters.append(it)
        return y.reshape(shape)

    def _evolve_bond_tensor(self, C: np.ndarray, L, R, delta: float) -> np.ndarray:
        """tdvp.py:164-188"""
        if self.expm == "dense":
            return loc.expm_dense(loc.dense_keff(L, R), C.reshape(-1), delta).reshape(C.shape)
        shape = C.shape
        y, it = loc.expm_krylov(lambda v: loc.bond_apply(v.reshape(shape), L, R).reshape(-1),
                                C.reshape(-1), delta, tol=self.tol)
        self.krylov_iters.append(it)
        return y.reshape(shape)

    def _evolve_site(self, site: int, delta: float) -> np.ndarray:
        """tdvp.py:147-162"""
        return self._evolve(self.A[site], self._get_env("left", site - 1),
                            self._get_env("right", site + 1), self.W[site], delta)

    def _evolve_split(self, site_l: int, delta: float):
        """tdvp.py:270-296: merge, evolve, SVD split, truncate."""
        Al, Ar = self.A[site_l], self.A[site_l + 1]
        theta = loc.merge_theta(Al, Ar)
        W2 = loc.merge_w(self.W[site_l], self.W[site_l + 1])
        theta = self._evolve(theta, self._get_env("left", site_l - 1),
                             self._get_env("right", site_l + 2), W2, delta)
        return loc.split_theta(theta, Al.shape[1], Ar.shape[2], self.chi_max, self.eps)

    # ------------------------------------------------------------ sweeps
    def do_time_step(self) -> None:
        """tdvp.py:50-63"""
        if self.gauge == "reference":
            M.make_site_canonical(self.A, 0)
        if self.algorithm == "2tdvp":
            self._sweep_right_2()
            self._sweep_left_2()
        else:
            if self.algorithm == "a1tdvp":
                self._new_target_bond_dims()
            self._sweep_right_1()
            self._sweep_left_1()

    def _new_target_bond_dims(self) -> None:
        """tdvp.py:212-266 with the 'exact_exponential' propagator
        exp(-i (pi/2) (dt/2) H) (tdvp.py:256-257).  Reference mode replays the
        reference's order (make_site_canonical before each environment layer,
        right environments not rebuilt while the centre walks left); consistent
        mode moves the centre with reduced QRs and refreshes the environment it
        uses.  Targets of bonds 1..N-1: max(1, min(k, min(max_bond, len(s))))
        with k the first i with ||s[i:]|| < eps (tdvp.py:261-265)."""
        n = len(self.A)
        for site in range(n - 1):
            if self.gauge == "reference":
                M.make_site_canonical(self.A, site + 1)
            else:
                M.orthonormalize_left(self.A, site)
            self._update_env("left", site)
        for bond in range(n - 1, 0, -1):
            Al, Ar = self.A[bond - 1], self.A[bond]
            theta = loc.merge_theta(Al, Ar)
            W2 = loc.merge_w(self.W[bond - 1], self.W[bond])
            theta = self._evolve(theta, self._get_env("left", bond - 2), self._get_env("right", bond + 1),
                                 W2, self.dt / 2)
            chil, chir = Al.shape[1], Ar.shape[2]
            Mt = theta.reshape(2, 2, chil, chir).transpose(0, 2, 1, 3).reshape(2 * chil, 2 * chir)
            s = np.linalg.svd(Mt, compute_uv=False)
            cap = min(self.max_bond[bond], len(s))
            k = next((i for i in range(len(s)) if np.linalg.norm(s[i:]) < self.eps), cap)
            self.target[bond] = max(1, min(k, cap))
            if self.gauge == "reference":
                M.make_site_canonical(self.A, bond - 1)
            else:
                M.orthonormalize_right(self.A, bond)
                self._update_env("right", bond)

    def _sweep_right_2(self) -> None:
        """tdvp.py:107-125"""
        n = len(self.A)
        h = self.dt / 2
        for site in range(n - 1):
            u, s, v = self._evolve_split(site, h)
            self.A[site] = u
            self.A[site + 1] = s[None, :, None] * v
            if site < n - 2:
                self._update_env("left", site)
                self.A[site + 1] = self._evolve_site(site + 1, -h)

    def _sweep_left_2(self) -> None:
        """tdvp.py:127-145"""
        n = len(self.A)
        h = self.dt / 2
        for site in range(n - 1, 0, -1):
            u, s, v = self._evolve_split(site - 1, h)
            self.A[site] = v
            self.A[site - 1] = u * s[None, None, :]
            if site > 1:
                self._update_env("right", site)
                self.A[site - 1] = self._evolve_site(site - 1, -h)

    def _sweep_right_1(self) -> None:
        """tdvp.py:65-84"""
        n = len(self.A)
        h = self.dt / 2
        for site in range(n):
            shape = (2, self.target[site], self.target[site + 1])
            new_A = self._evolve_site(site, h)
            if site == n - 1:
                self.A[site] = new_A
                continue
            Q, C = M.left_qr(new_A, reduced=False)
            self.A[site] = M.truncate_pad(Q, shape)
            self._update_env("left", site)
            C = M.truncate_pad(C, (shape[2], C.shape[1]))
            C = self._evolve_bond_tensor(C, self._get_env("left", site),
                                         self._get_env("right", site + 1), -h)
            self.A[site + 1] = np.einsum("xb,pbc->pxc", C, self.A[site + 1])

    def _sweep_left_1(self) -> None:
        """tdvp.py:86-105"""
        h = self.dt / 2
        for site in range(len(self.A) - 1, -1, -1):
            shape = (2, self.target[site], self.target[site + 1])
            new_A = self._evolve_site(site, h)
            if site == 0:
                self.A[site] = new_A
                continue
            Q, C = M.right_qr(new_A, reduced=False)
            self.A[site] = M.truncate_pad(Q, shape)
            self._update_env("right", site)
            C = M.truncate_pad(C, (C.shape[0], shape[1]))
            C = self._evolve_bond_tensor(C, self._get_env("left", site - 1),
                                         self._get_env("right", site), -h)
            self.A[site - 1] = np.einsum("pab,bx->pax", self.A[site - 1], C)

    # ------------------------------------------------------------ observables
    def measure(self):
        """algorithm.py:58-63 -> mps.py:100-140.  Reference mode mutates the gauge."""
        if self.gauge == "reference":
            return M.measure_reference(self.A)
        return M.measure_consistent(self.A)


def run_trajectory(algo, steps: int, plot_interval: int):
    """Measure-then-step cadence of quantum_game.py:82-119 (TDVP branch):
    measure at steps 0, interval, 2*interval, ... before stepping; the state
    after the final step is not measured."""
    rows = steps // plot_interval + (1 if steps % plot_interval else 0)
    n = len(algo.A)
    pop = np.zeros((rows, n))
    dpop = np.zeros((rows, n))
    sse = np.zeros((rows, n))
    bd = np.zeros((rows, n + 1))
    for step in range(steps):
        if step % plot_interval == 0:
            r = step // plot_interval
            pop[r], dpop[r], sse[r], bd[r] = algo.measure()
        algo.do_time_step()
    return dict(population=pop, d_population=dpop, single_site_entropy=sse, bond_dims=bd)
