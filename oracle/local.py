"""Local TDVP operations restated from reference algorithms/tdvp.py and
lautils/lautils.py (oracle).

Index conventions (reference tdvp.py:298-347):
  A[p, a, b]          MPS site tensor (physical, left bond, right bond)
  W[p, p', wl, wr]    MPO tensor
  L[a, w, a*]         left environment (ket, MPO, bra)
  R[b, w, b*]         right environment (ket, MPO, bra)
Matrix-free operators (what the reference's dense H_eff applies through
``tensordot(psi, U, (0, 0))``, i.e. with the transpose of its (ket, bra)
matrix, lautils.py:80-81):
  1-site  y[p',a*,b*] = sum W[p,p',w,w'] L[a,w,a*] R[b,w',b*] x[p,a,b]
  2-site  the same with the merged W2 and x = Theta[(p1 p2), a, c]
  bond    y[a*,b*]    = sum L[a,w,a*] R[b,w,b*] C[a,b]
Every propagator is exp(-i (pi/2) delta H) (lautils.py:44-54).
"""
from __future__ import annotations

import numpy as np

HALF_PI = np.pi / 2


# ----------------------------------------------------------------- contractions

def merge_theta(Al: np.ndarray, Ar: np.ndarray) -> np.ndarray:
    """Theta[(p1 p2), a, c] = sum_b Al[p1,a,b] Ar[p2,b,c] (reference tdvp.py:278-280)."""
    T = np.einsum("pab,qbc->pqac", Al, Ar, optimize=True)
    return T.reshape(4, Al.shape[1], Ar.shape[2])


def merge_w(Wl: np.ndarray, Wr: np.ndarray) -> np.ndarray:
    """W2[(p1 p2), (p1' p2'), wl, wr] = sum_m Wl[p1,p1',wl,m] Wr[p2,p2',m,wr]
    (reference tdvp.py:282-284)."""
    T = np.einsum("abxm,cdmy->acbdxy", Wl, Wr)
    return T.reshape(4, 4, Wl.shape[2], Wr.shape[3])


def heff_apply(x: np.ndarray, L: np.ndarray, R: np.ndarray, W: np.ndarray) -> np.ndarray:
    """Matrix-free effective Hamiltonian (1-site or 2-site by the size of W):
    the operator reference algorithms/tdvp.py:298-311 (_assemble_H_eff)
    assembles and tdvp.py:350-365 (_evolve_A) applies, never formed densely."""
    t = np.einsum("awx,pac->pwxc", L, x, optimize=True)          # L first
    t = np.einsum("pqwv,pwxc->qvxc", W, t, optimize=True)        # then W
    return np.einsum("qvxc,cvy->qxy", t, R, optimize=True)       # then R


def bond_apply(C: np.ndarray, L: np.ndarray, R: np.ndarray) -> np.ndarray:
    """Bond effective Hamiltonian applied to C, matrix-free (reference
    algorithms/tdvp.py:312-326, _assemble_K_eff, used by _evolve_bond 164-188)."""
    return np.einsum("awx,bwy,ab->xy", L, R, C, optimize=True)


def env_left(Lprev: np.ndarray, A: np.ndarray, W: np.ndarray) -> np.ndarray:
    """L_new[b,w',b*] = sum A[p,a,b] W[p,p',w,w'] conj(A[p',a*,b*]) L[a,w,a*]
    (reference tdvp.py:328-347, side='left')."""
    t = np.einsum("awx,pab->pwxb", Lprev, A, optimize=True)
    t = np.einsum("pqwv,pwxb->qvxb", W, t, optimize=True)
    return np.einsum("qvxb,qxy->bvy", t, A.conj(), optimize=True)


def env_right(Rprev: np.ndarray, A: np.ndarray, W: np.ndarray) -> np.ndarray:
    """R_new[a,w,a*] = sum A[p,a,b] W[p,p',w,w'] conj(A[p',a*,b*]) R[b,w',b*]
    (reference tdvp.py:328-347, side='right')."""
    t = np.einsum("bvy,pab->pavy", Rprev, A, optimize=True)
    t = np.einsum("pqwv,pavy->qawy", W, t, optimize=True)
    return np.einsum("qawy,qxy->awx", t, A.conj(), optimize=True)


def dense_heff(L: np.ndarray, R: np.ndarray, W: np.ndarray) -> np.ndarray:
    """The reference's dense H_eff[(p,a,b),(p',a*,b*)] (tdvp.py:298-310, 349-359)."""
    H = np.einsum("pqwv,awx,bvy->pabqxy", W, L, R)
    n = W.shape[0] * L.shape[0] * R.shape[0]
    return H.reshape(n, n)


def dense_keff(L: np.ndarray, R: np.ndarray) -> np.ndarray:
    """The reference's dense bond K_eff[(a,b),(a*,b*)] (tdvp.py:312-326)."""
    K = np.einsum("awx,bwy->abxy", L, R)
    n = L.shape[0] * R.shape[0]
    return K.reshape(n, n)


# ----------------------------------------------------------------- exponentials

def expm_dense(Hmat: np.ndarray, v: np.ndarray, delta: float) -> np.ndarray:
    """Reference ``timestep`` (lautils.py:44-54, 57-81): U from eigh, result U^T v."""
    w, V = np.linalg.eigh(Hmat)
    U = (np.exp(-1j * HALF_PI * delta * w) * V) @ V.conj().T
    return np.tensordot(v, U, (0, 0))


def expm_krylov(apply, v: np.ndarray, delta: float, tol: float = 1e-14,
                max_iter: int = 80):
    """exp(-i (pi/2) delta H) v by Lanczos with full reorthogonalisation (FP64).

    The propagator of reference lautils/lautils.py:44-81 (calculate_U and
    timestep form exp(-i (pi/2) delta H) densely); here applied in a Krylov space.

    ``apply`` computes H x for flattened x.  Stops when the a-posteriori error
    beta_0 * beta_m * |[exp(-i tau T_m)]_{m,0}| drops below ``tol`` * |v|, or on
    an invariant subspace.  Returns (result, iterations).
    """
    tau = HALF_PI * delta
    beta0 = np.linalg.norm(v)
    if beta0 == 0.0:
        return v.copy(), 0
    V = [v / beta0]
    alpha, beta = [], []
    for j in range(max_iter):
        w = apply(V[j])
        a = np.vdot(V[j], w).real
        alpha.append(a)
        w = w - a * V[j] - (beta[-1] * V[j - 1] if j > 0 else 0)
        for u in V:                       # full reorthogonalisation, twice is enough
            w = w - np.vdot(u, w) * u
        for u in V:
            w = w - np.vdot(u, w) * u
        b = np.linalg.norm(w)
        m = j + 1
        T = np.diag(alpha) + np.diag(beta, 1) + np.diag(beta, -1) if m > 1 else np.array([[alpha[0]]])
        ev, S = np.linalg.eigh(T)
        c = S @ (np.exp(-1j * tau * ev) * S[0, :].conj())
        if b <= 1e-13 * max(1.0, abs(a)) or beta0 * b * abs(c[-1]) < tol * beta0:
            out = np.zeros_like(v)
            for k in range(m):
                out = out + c[k] * V[k]
            return beta0 * out, m
        beta.append(b)
        V.append(w / b)
    raise RuntimeError("Lanczos expm did not converge")


# ----------------------------------------------------------------- split

def truncation_rank(s: np.ndarray, chi_max: int, eps: float) -> int:
    """First i with ||s[i:]||_2 < eps, else min(chi_max, len(s)); then capped
    (reference tdvp.py:290-292)."""
    cap = min(chi_max, len(s))
    k = next((i for i in range(len(s)) if np.linalg.norm(s[i:]) < eps), cap)
    return min(k, cap)


def split_theta(theta: np.ndarray, chil: int, chir: int, chi_max: int, eps: float):
    """SVD split of an evolved Theta[(p1 p2), a, c] (reference tdvp.py:286-296).

    Returns (Al[p1,a,k], s_normalised[k], Vr[p2,k,c]).
    """
    M = theta.reshape(2, 2, chil, chir).transpose(0, 2, 1, 3).reshape(2 * chil, 2 * chir)
    u, s, vh = np.linalg.svd(M, full_matrices=False)
    k = truncation_rank(s, chi_max, eps)
    Al = u.reshape(2, chil, -1)[:, :, :k]
    Vr = vh.reshape(-1, 2, chir).transpose(1, 0, 2)[:, :k, :]
    sk = s[:k]
    return Al, sk / np.linalg.norm(sk), Vr
