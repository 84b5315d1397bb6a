"""qca_heff_apply_noalloc (the effective-Hamiltonian application without
device allocation) against qca_heff_apply on the same inputs: identical
outputs (both run the same GEMM plans on same-sized workspaces), on two calls
in a row with the second one smaller, for the two-site, one-site and bond
blocks in both precisions; applications whose workspace exceeds the static
scratch are rejected."""
import ctypes

import numpy as np
import pytest

from conftest import crand

import qca_native as qn
from oracle import automaton as AU, local as OL

pytestmark = pytest.mark.gpu


def operands(rng, kind, chil, chir, D, prec):
    import torch
    ct = torch.complex128 if prec == 0 else torch.complex64
    dq = {0: 4, 1: 2, 2: 1}[kind]
    shape = (dq, chil, chir) if kind != 2 else (chil, chir)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(ct).cuda()
    return dev(crand(rng, *shape)), dev(crand(rng, chil, D, chil)), dev(crand(rng, chir, D, chir)), shape, ct


def apply(name, prec, kind, x, L, R, W, D, chil, chir):
    import torch
    y = torch.empty_like(x)
    Wp = W.ctypes.data if W is not None else None
    torch.cuda.synchronize()
    qn.call(name, prec, kind, qn.ptr(x), qn.ptr(y), qn.ptr(L), qn.ptr(R), Wp, D, D, chil, chir, None)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_noalloc_equals_heff_apply(gpu, prec, kind):
    rng = np.random.default_rng(20 + 3 * kind + prec)
    Wsite = AU.mpo_tensors(12, 2, range(2, 5))[6]
    D = Wsite.shape[2]
    W = {0: np.ascontiguousarray(OL.merge_w(Wsite, Wsite)), 1: np.ascontiguousarray(Wsite), 2: None}[kind]
    for chil, chir in ((64, 48), (20, 12)):           # second call smaller than the first
        x, L, R, shape, _ = operands(rng, kind, chil, chir, D, prec)
        a = apply("qca_heff_apply", prec, kind, x, L, R, W, D, chil, chir)
        b = apply("qca_heff_apply_noalloc", prec, kind, x, L, R, W, D, chil, chir)
        assert a.shape == shape
        assert np.array_equal(a, b), (kind, prec, chil, chir)
        # and the result is the effective Hamiltonian (oracle contraction)
        xn, Ln, Rn = (t.cpu().numpy().astype(np.complex128) for t in (x, L, R))
        ref = OL.bond_apply(xn, Ln, Rn) if kind == 2 else OL.heff_apply(xn, Ln, Rn, W)
        tol = 1e-12 if prec == 0 else 2e-5
        assert np.abs(b - ref).max() <= tol * np.abs(ref).max()


def test_noalloc_rejects_oversized_workspace(gpu):
    import torch
    rng = np.random.default_rng(5)
    Wsite = AU.mpo_tensors(12, 2, range(2, 5))[6]
    D = Wsite.shape[2]
    W = np.ascontiguousarray(OL.merge_w(Wsite, Wsite))
    x, L, R, _, _ = operands(rng, 0, 128, 128, D, 0)
    y = torch.empty_like(x)
    with pytest.raises(qn.QcaError, match="64 MiB"):
        qn.call("qca_heff_apply_noalloc", 0, 0, qn.ptr(x), qn.ptr(y), qn.ptr(L), qn.ptr(R), W.ctypes.data,
                D, D, 128, 128, None)
