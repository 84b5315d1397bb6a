"""GPU parity of the C-ABI primitives against the oracle (tests/../oracle).

Tolerances: complex128 paths to 1e-12 relative (contraction rounding only);
complex64 (mixed) contractions to 2e-5 relative, the FP32 round-off of sums of
up to a few thousand terms.
"""
import ctypes

import numpy as np
import pytest

from conftest import crand

import qca_native as qn
from oracle import automaton as AU, local as LOC

pytestmark = pytest.mark.gpu


def dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).cuda()


def herm_env(rng, chi, D):
    E = crand(rng, chi, D, chi)
    return 0.5 * (E + np.conj(np.transpose(E, (2, 1, 0))))


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


W_CASES = [(1, range(1, 2)), (2, range(2, 5)), (2, range(1, 5)), (1, range(2, 3))]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("chil,chir,rule", [(1, 1, 0), (3, 5, 1), (16, 16, 0), (37, 29, 2),
                                            (64, 64, 1), (128, 96, 3)])
def test_heff_apply(gpu, prec, kind, chil, chir, rule):
    import torch
    rng = np.random.default_rng(chil * 131 + chir * 7 + kind)
    d, iv = W_CASES[rule]
    W = AU.bulk_tensor(d, iv)
    D = W.shape[2]
    if kind == 2:
        chir = chil if kind == 2 and rule % 2 == 0 else chir
        x = crand(rng, chil, chir)
        L, R = herm_env(rng, chil, D), herm_env(rng, chir, D)
        ref = LOC.bond_apply(x, L, R)
        Wd = None
    elif kind == 1:
        x = crand(rng, 2, chil, chir)
        L, R = herm_env(rng, chil, D), herm_env(rng, chir, D)
        ref = LOC.heff_apply(x, L, R, W)
        Wd = W
    else:
        x = crand(rng, 4, chil, chir)
        L, R = herm_env(rng, chil, D), herm_env(rng, chir, D)
        Wd = LOC.merge_w(W, W)
        ref = LOC.heff_apply(x, L, R, Wd)
    dt = np.complex128 if prec == 0 else np.complex64
    xd, Ld, Rd = dev(torch, x, dt), dev(torch, L, dt), dev(torch, R, dt)
    yd = torch.zeros_like(xd)
    Wh = np.ascontiguousarray(Wd if Wd is not None else np.zeros(1), dtype=np.complex128)
    qn.call("qca_heff_apply", prec, kind, qn.ptr(xd), qn.ptr(yd), qn.ptr(Ld), qn.ptr(Rd),
            Wh.ctypes.data if Wd is not None else None, D, D, chil, chir, None)
    got = yd.cpu().numpy().astype(np.complex128)
    assert rel(got, ref) < (1e-12 if prec == 0 else 2e-5)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("chil,chir", [(1, 2), (5, 7), (32, 64), (128, 128)])
def test_env_update(gpu, prec, side, chil, chir):
    import torch
    rng = np.random.default_rng(chil + 17 * chir + side)
    Ws = AU.mpo_tensors(6, 2, range(2, 5))
    W = Ws[2]
    Dl, Dr = W.shape[2], W.shape[3]
    A = crand(rng, 2, chil, chir)
    dt = np.complex128 if prec == 0 else np.complex64
    if side == 0:
        E = herm_env(rng, chil, Dl)
        ref = LOC.env_left(E, A, W)
        out = torch.zeros((chir, Dr, chir), dtype=torch.complex128 if prec == 0 else torch.complex64, device="cuda")
    else:
        E = herm_env(rng, chir, Dr)
        ref = LOC.env_right(E, A, W)
        out = torch.zeros((chil, Dl, chil), dtype=torch.complex128 if prec == 0 else torch.complex64, device="cuda")
    Ad, Ed = dev(torch, A, dt), dev(torch, E, dt)
    Wh = np.ascontiguousarray(W, dtype=np.complex128)
    qn.call("qca_env_update", prec, side, qn.ptr(Ad), qn.ptr(Ed), qn.ptr(out), Wh.ctypes.data,
            Dl, Dr, chil, chir, None)
    got = out.cpu().numpy().astype(np.complex128)
    assert rel(got, ref) < (1e-12 if prec == 0 else 2e-5)


def test_env_update_golden(gpu):
    """Reference _assemble_new_layer_H_eff outputs (golden, reference tdvp.py:328-347)."""
    import torch
    from conftest import golden
    g = golden("units")
    W = g["W"]
    A, Lp, Rp = g["env_A"], g["env_Lprev"], g["env_Rprev"]
    chil, chir = A.shape[1], A.shape[2]
    Dl = Dr = W.shape[2]
    Wh = np.ascontiguousarray(W)
    for side, E, ref, shape in ((0, Lp, g["env_Lnew"], (chir, Dr, chir)), (1, Rp, g["env_Rnew"], (chil, Dl, chil))):
        out = torch.zeros(shape, dtype=torch.complex128, device="cuda")
        Ad, Ed = dev(torch, A, np.complex128), dev(torch, E, np.complex128)  # keep both alive
        qn.call("qca_env_update", 0, side, qn.ptr(Ad), qn.ptr(Ed), qn.ptr(out), Wh.ctypes.data,
                Dl, Dr, chil, chir, None)
        assert rel(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("kind,chil,chir,delta", [(1, 4, 6, 0.05), (1, 4, 6, -0.05), (0, 3, 4, 0.05),
                                                 (2, 5, 5, -0.05), (0, 24, 24, 0.05), (1, 48, 40, -0.05)])
def test_krylov_expm(gpu, prec, kind, chil, chir, delta):
    import torch
    rng = np.random.default_rng(kind * 1000 + chil)
    W = AU.bulk_tensor(2, range(2, 5))
    D = W.shape[2]
    L, R = herm_env(rng, chil, D), herm_env(rng, chir, D)
    if kind == 0:
        Wk = LOC.merge_w(W, W)
        x = crand(rng, 4, chil, chir)
        Hm = LOC.dense_heff(L, R, Wk)
    elif kind == 1:
        Wk = W
        x = crand(rng, 2, chil, chir)
        Hm = LOC.dense_heff(L, R, Wk)
    else:
        Wk = None
        x = crand(rng, chil, chir)
        Hm = LOC.dense_keff(L, R)
    x /= np.linalg.norm(x)
    ref = LOC.expm_dense(Hm, x.reshape(-1), delta).reshape(x.shape)
    dt = np.complex128 if prec == 0 else np.complex64
    xd = dev(torch, x, np.complex128)
    Ld, Rd = dev(torch, L, dt), dev(torch, R, dt)
    Wh = np.ascontiguousarray(Wk if Wk is not None else np.zeros(1), dtype=np.complex128)
    it = ctypes.c_int(0)
    qn.call("qca_krylov_expm", prec, kind, qn.ptr(xd), qn.ptr(Ld), qn.ptr(Rd),
            Wh.ctypes.data if Wk is not None else None, D, D, chil, chir, delta, 0.0,
            ctypes.byref(it), None)
    got = xd.cpu().numpy()
    assert it.value >= 2
    assert np.abs(got - ref).max() < (1e-12 if prec == 0 else 2e-5)
    assert abs(np.linalg.norm(got) - 1.0) < (1e-12 if prec == 0 else 1e-5)


@pytest.mark.parametrize("right", [1, 0])
@pytest.mark.parametrize("chil,chir,chi_max,eps", [(1, 1, 4, 1e-10), (2, 4, 8, 1e-10), (8, 8, 16, 1e-10),
                                                   (8, 8, 5, 1e-10), (16, 32, 64, 1e-3),
                                                   # one workgroup with four 16-column blocks,
                                                   # full and ragged (34..48 columns)
                                                   (32, 32, 64, 1e-10), (24, 20, 64, 1e-10),
                                                   (17, 20, 64, 1e-10),
                                                   (64, 64, 128, 1e-10), (128, 128, 128, 1e-10),
                                                   # the widest split (512 columns) and a
                                                   # ragged 400 x 272 one capped at 128
                                                   (256, 256, 256, 1e-10), (200, 136, 128, 1e-10)])
def test_svd_split(gpu, right, chil, chir, chi_max, eps):
    import torch
    rng = np.random.default_rng(chil * 3 + chir + right)
    th = crand(rng, 4, chil, chir)
    # graded spectrum so that truncation tails span many orders of magnitude
    M = th.reshape(2, 2, chil, chir).transpose(0, 2, 1, 3).reshape(2 * chil, 2 * chir)
    u, s, vh = np.linalg.svd(M, full_matrices=False)
    s = np.logspace(0, -14, len(s))
    M = (u * s) @ vh
    M /= np.linalg.norm(M)
    th = M.reshape(2, chil, 2, chir).transpose(0, 2, 1, 3).reshape(4, chil, chir)
    refAl, refs, refAr = LOC.split_theta(th, chil, chir, chi_max, eps)
    k_ref = len(refs)
    thd = dev(torch, th, np.complex128)
    kmax = min(chi_max, 2 * chil, 2 * chir)
    Al = torch.zeros((2, chil, kmax), dtype=torch.complex128, device="cuda")
    Ar = torch.zeros((2, kmax, chir), dtype=torch.complex128, device="cuda")
    k = ctypes.c_int(0)
    sig = np.zeros(min(2 * chil, 2 * chir))
    qn.call("qca_svd_split", qn.ptr(thd), qn.ptr(Al), qn.ptr(Ar), chil, chir, chi_max, eps, right,
            ctypes.byref(k), sig.ctypes.data, None)
    assert k.value == k_ref
    kk = k.value
    Al = Al.flatten()[: 2 * chil * kk].reshape(2, chil, kk).cpu().numpy()
    Ar = Ar.flatten()[: 2 * kk * chir].reshape(2, kk, chir).cpu().numpy()
    s_np = np.linalg.svd(M, compute_uv=False)
    assert np.abs(sig - s_np).max() < 1e-13
    # truncated state and isometry, gauge independent
    ref_th = np.einsum("pak,k,qkc->pqac", refAl, refs, refAr)
    got_th = np.einsum("pak,qkc->pqac", Al, Ar)
    assert np.abs(got_th - ref_th).max() < 1e-12
    if right:
        I = np.einsum("pak,pal->kl", Al.conj(), Al)
    else:
        I = np.einsum("pkc,plc->kl", Ar, Ar.conj())
    # Householder products and accumulated rotations lose orthogonality like
    # n * eps in the isometry's length n (512 at the widest split)
    n = 2 * chil if right else 2 * chir
    assert np.abs(I - np.eye(kk)).max() < max(1e-13, 2 * n * np.finfo(float).eps)


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("chil,chir,target", [(1, 1, 2), (2, 2, 4), (4, 8, 8), (8, 8, 8), (32, 32, 32),
                                              (8, 4, 4), (16, 64, 32), (128, 256, 256)])
def test_qr_site_matches_lapack(gpu, side, chil, chir, target):
    """Complete QR with LAPACK's conventions: Q and R equal numpy.linalg.qr's,
    including the complement columns (reference mps.py:83-98, tdvp.py:74-78, 95-99)."""
    import torch
    from oracle import mps as OM
    if side == 1:
        target = min(target, 2 * chir)
    else:
        target = min(target, 2 * chil)
    rng = np.random.default_rng(chil + chir + target)
    A = crand(rng, 2, chil, chir)
    if side == 0:
        Q, R = OM.left_qr(A, reduced=False)
        refA = OM.truncate_pad(Q, (2, chil, target))
        refC = OM.truncate_pad(R, (target, chir))
        An = torch.zeros((2, chil, target), dtype=torch.complex128, device="cuda")
        C = torch.zeros((target, chir), dtype=torch.complex128, device="cuda")
    else:
        Q, R = OM.right_qr(A, reduced=False)
        refA = OM.truncate_pad(Q, (2, target, chir))
        refC = OM.truncate_pad(R, (chil, target))
        An = torch.zeros((2, target, chir), dtype=torch.complex128, device="cuda")
        C = torch.zeros((chil, target), dtype=torch.complex128, device="cuda")
    qn.call("qca_qr_site", side, qn.ptr(dev(torch, A, np.complex128)), chil, chir, target,
            qn.ptr(An), qn.ptr(C), None)
    assert np.abs(An.cpu().numpy() - refA).max() < 1e-12
    assert np.abs(C.cpu().numpy() - refC).max() < 1e-12


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n,d,iv", [(6, 1, range(1, 2)), (9, 2, range(2, 5)), (12, 2, range(1, 3)),
                                    (16, 1, range(2, 3))])
def test_exact_apply(gpu, prec, n, d, iv):
    import torch
    rng = np.random.default_rng(n)
    psi = crand(rng, 1 << n)
    ref = AU.apply_h(psi, n, d, iv)
    dt = np.complex128 if prec == 0 else np.complex64
    pd = dev(torch, psi, dt)
    out = torch.zeros_like(pd)
    qn.call("qca_exact_apply_h", prec, qn.ptr(pd), qn.ptr(out), n, d, iv.start, iv.stop, None)
    assert rel(out.cpu().numpy().astype(np.complex128), ref) < (1e-15 if prec == 0 else 1e-6)


def test_exact_expm_and_rho(gpu):
    import torch
    from oracle import exact as OE, mps as OM
    n, d, iv, dt = 10, 2, range(2, 5), 0.2
    pl = [0.0] * n
    pl[4] = pl[5] = pl[6] = 1.0
    ex = OE.ExactOracle(OM.product_state(pl), n, d, iv, dt)
    psi = dev(torch, OM.as_vector(OM.product_state(pl)), np.complex128)
    for _ in range(5):
        it = ctypes.c_int(0)
        qn.call("qca_exact_expm", qn.ptr(psi), n, d, iv.start, iv.stop, dt, 0.0, ctypes.byref(it), None)
        ex.do_time_step()
        assert np.abs(psi.cpu().numpy() - ex.psi).max() < 1e-12
    rho = np.zeros((n, 2, 2), dtype=np.complex128)
    qn.call("qca_exact_rho", qn.ptr(psi), n, rho.ctypes.data, None)
    assert np.abs(rho[:, 1, 1].real - ex.populations()).max() < 1e-13
    assert np.abs(np.trace(rho, axis1=1, axis2=2) - 1).max() < 1e-12
