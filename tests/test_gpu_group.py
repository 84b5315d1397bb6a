"""Grouped lockstep stepping (qca_group_*): many 2TDVP engines advanced with one
launch per kernel per sweep position.  Every instance must end bit-identical
to the same engine stepped alone (qca_engine_step), which the oracle tests pin
to the reference's algorithm."""
import numpy as np
import pytest

import qca_native as qn
from oracle import automaton as AU, mps as OM, tdvp as OT

pytestmark = pytest.mark.gpu


def draws_and_chain(prec):
    """Six batched-workload draws (default_rng(4), chi_max as drawn) plus the
    41-cell chi=128 chain, whose bonds pass 16 (multi-launch Krylov, Jacobi
    with QR preconditioning) within the steps taken."""
    import workloads
    insts = workloads.batch_instances(6, seed=4)
    out = [(workloads.mpo_tensors(i), workloads.initial_tensors(i), i.chi_max, i.step_size, i.svd_epsilon)
           for i in insts]
    n = 41
    pl = [0.0] * n
    pl[19] = pl[20] = pl[21] = 1.0
    out.append((AU.mpo_tensors(n, 2, range(2, 5)), OM.product_state(pl), 128, 0.1, 1e-10))
    return [qn.Engine(W, A, algorithm=2, precision=prec, chi_max=chi, dt=dt, eps=eps,
                      krylov_tol=1e-14 if prec == 0 else 1e-10) for W, A, chi, dt, eps in out]


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_group_is_bit_identical_to_engines_alone(gpu, prec):
    steps = 10
    grouped, alone = draws_and_chain(prec), draws_and_chain(prec)
    g = qn.Group(grouped)
    print(f"  group of {len(grouped)} ({'mixed' if prec else 'fp64'}): {steps} steps", flush=True)
    g.step(steps)
    for e in alone:
        e.step(steps)
    st = g.stats()
    assert st["steps"] == steps and st["host_syncs"] > 0
    assert max(int(e.bond_dims().max()) for e in grouped) > 32
    for b, a in zip(grouped, alone):
        assert np.array_equal(b.bond_dims(), a.bond_dims())
        assert np.array_equal(b.rho(), a.rho())
        for x, y in zip(b.tensors(), a.tensors()):
            assert np.array_equal(x, y)
        sb, sa = b.stats(), a.stats()
        for k in ("site_updates", "matvecs", "svd_calls", "svd_sweeps"):
            assert sb[k] == sa[k], k
    # engines stay usable alone after group steps, and the group goes on
    grouped[0].step(1)
    alone[0].step(1)
    g.step(1)
    for e in alone[1:]:
        e.step(1)
    alone[0].step(1)
    for b, a in zip(grouped, alone):
        assert np.array_equal(b.rho(), a.rho())
    g.close()
    for e in grouped + alone:
        e.close()


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_group_fused_split_gemms_equal_engines_alone(gpu, prec, monkeypatch):
    """Grouped GEMMs whose splits run fused (one workgroup per output tile
    walking all K-splits, no partial tiles or reduce; forced here on a small
    group, used by large groups) sum in the split-K launch's order: every
    instance still ends bit-identical to the engine stepped alone."""
    monkeypatch.setenv("QCA_GEMM_GROUP_FUSE_TILES", "0")
    steps = 6
    grouped, alone = draws_and_chain(prec), draws_and_chain(prec)
    g = qn.Group(grouped)
    g.step(steps)
    for e in alone:
        e.step(steps)
    assert max(int(e.bond_dims().max()) for e in grouped) > 32
    for b, a in zip(grouped, alone):
        assert np.array_equal(b.bond_dims(), a.bond_dims())
        assert np.array_equal(b.rho(), a.rho())
        for x, y in zip(b.tensors(), a.tensors()):
            assert np.array_equal(x, y)
    g.close()
    for e in grouped + alone:
        e.close()


def test_group_small_instances_match_oracle(gpu):
    """Grouped FP64 instances of the batched workload with at most 20 cells
    against the oracle (1e-10, identical bonds)."""
    import workloads
    insts = [i for i in workloads.batch_instances(40, seed=4) if i.ncells <= 20][:4]
    assert len(insts) >= 2
    engines = [qn.Engine(workloads.mpo_tensors(i), workloads.initial_tensors(i), algorithm=2, precision=0,
                         chi_max=i.chi_max, dt=i.step_size, eps=i.svd_epsilon, krylov_tol=1e-14) for i in insts]
    g = qn.Group(engines)
    g.step(4)
    for i, e in zip(insts, engines):
        o = OT.TDVPOracle(OM.product_state(i.plist), AU.mpo_tensors(i.ncells, i.distance, i.interval), "2tdvp",
                          0.1, i.chi_max, 1e-10, gauge="consistent", expm="lanczos")
        for _ in range(4):
            o.do_time_step()
        assert np.array_equal(e.bond_dims(), OM.bond_dims(o.A).astype(int))
        assert np.abs(e.rho() - OM.rho_consistent(o.A)).max() < 1e-10
    g.close()


def test_group_rejects_mixed_sets(gpu):
    n = 8
    pl = [0.0] * n
    pl[4] = 1.0
    Ws = AU.mpo_tensors(n, 1, range(1, 2))
    e2 = qn.Engine(Ws, OM.product_state(pl), algorithm=2, precision=0, chi_max=8, dt=0.1, eps=1e-10)
    e1 = qn.Engine(Ws, OM.product_state(pl), algorithm=1, precision=0, chi_max=8, dt=0.1, eps=1e-10)
    em = qn.Engine(Ws, OM.product_state(pl), algorithm=2, precision=1, chi_max=8, dt=0.1, eps=1e-10)
    with pytest.raises(qn.QcaError, match="2TDVP"):
        qn.Group([e2, e1])
    with pytest.raises(qn.QcaError, match="precision"):
        qn.Group([e2, em])
    with pytest.raises(qn.QcaError, match="twice"):
        qn.Group([e2, e2])
