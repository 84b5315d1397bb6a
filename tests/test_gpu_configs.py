"""GPU engine at the benchmark workload sizes (41-cell chi=128 2TDVP, 101-cell
chi=256 1TDVP, batched draws with chi_max up to 128, a served window with a
large request), observables beyond <n_i>, and the error paths.

Where the oracle covers the size (tests/golden/oracle_*.npz, or a live oracle
step of a few tens of seconds) the plain FP64 path must match it to 1e-10 with
identical bond dimensions; at the full sizes the checks are size-independent
properties (norm and energy conservation of projector-splitting 1TDVP, mixed
vs FP64 within the 1e-6 gate, bit-identical batched vs alone).
"""
import numpy as np
import pytest

from conftest import crand, golden

import qca_native as qn
from oracle import automaton as AU, exact as OE, local as OL, mps as OM, tdvp as OT

pytestmark = pytest.mark.gpu


def engine(As, Ws, algo, chi, dt, eps, prec=0, tol=0.0, **kw):
    return qn.Engine(Ws, As, algorithm=algo, precision=prec, chi_max=chi, dt=dt, eps=eps,
                     krylov_tol=tol or (1e-14 if prec == 0 else 1e-10), **kw)


def forty_one_cells():
    n = 41
    pl = [0.0] * n
    pl[19] = pl[20] = pl[21] = 1.0
    return n, pl, AU.mpo_tensors(n, 2, range(2, 5))


# ------------------------------------------------------------------ 41 cells, chi 128

def test_41cell_chi128_trajectory_matches_oracle(gpu):
    """2TDVP, 41 cells, distance 2, interval [2,4], cells 19..21 alive, dt 0.1,
    chi_max 128, eps 1e-10 (reference tdvp.py:107-145): the first 14 steps, in
    which the bonds grow to 128, against the oracle's trajectory; also the
    device entropy and the full single-site density matrices (off-diagonals
    included, reference mps.py:118-140)."""
    g = golden("oracle_41cell_2tdvp")
    n, pl, Ws = forty_one_cells()
    e = engine(OM.product_state(pl), Ws, 2, 128, 0.1, 1e-10)
    steps = int(g["params"][-1])
    for s in range(steps + 1):
        rho = e.rho()
        pop, dpop, sse, bd = e.observables()
        assert np.array_equal(bd, g["bond_dims"][s]), s
        assert np.abs(rho - g["rho"][s]).max() < 1e-10, s
        assert np.abs(pop - g["population"][s]).max() < 1e-10, s
        assert np.array_equal(dpop, np.round(g["population"][s])), s
        assert np.abs(sse - g["single_site_entropy"][s]).max() < 1e-8, s
        if s < steps:
            e.step(1)
    assert int(g["bond_dims"][-1].max()) == 128


def test_41cell_saturated_step_matches_oracle_and_mixed_gate(gpu):
    """From an FP64 engine state with at least 20 of the 40 bonds at chi_max
    = 128: one full 2TDVP step against the oracle restarted from the engine's
    own tensors (<n_i> and rho within 1e-10, identical bonds); then five steps
    of the mixed path from the same tensors within 1e-6 of FP64."""
    n, pl, Ws = forty_one_cells()
    import time
    from threadpoolctl import threadpool_limits
    t0 = time.perf_counter()
    e64 = engine(OM.product_state(pl), Ws, 2, 128, 0.1, 1e-10)
    e64.step(24)
    bd = e64.bond_dims()
    assert (bd == 128).sum() >= 20, bd
    As = e64.tensors()
    print(f"  41-cell: 24 FP64 steps, {(bd == 128).sum()} bonds at 128 ({time.perf_counter() - t0:.1f} s); "
          "oracle step", flush=True)
    with threadpool_limits(16):
        o = OT.TDVPOracle([a.copy() for a in As], Ws, "2tdvp", 0.1, 128, 1e-10, gauge="consistent",
                          expm="lanczos")
        o.do_time_step()
    print(f"  41-cell: oracle step done ({time.perf_counter() - t0:.1f} s)", flush=True)
    e32 = engine([a.copy() for a in As], Ws, 2, 128, 0.1, 1e-10, prec=1)
    e64.step(1)
    assert np.array_equal(e64.bond_dims(), OM.bond_dims(o.A).astype(int))
    assert np.abs(e64.rho() - OM.rho_consistent(o.A)).max() < 1e-10
    worst = 0.0
    e32.step(1)
    for _ in range(5):
        worst = max(worst, np.abs(e32.rho()[:, 1, 1].real - e64.rho()[:, 1, 1].real).max())
        e32.step(1)
        e64.step(1)
    worst = max(worst, np.abs(e32.rho()[:, 1, 1].real - e64.rho()[:, 1, 1].real).max())
    assert worst < 1e-6, worst


# ------------------------------------------------------------------ 1TDVP at chi 256

def padded_state(n, chi=256, seed=3):
    import states
    return states.padded(states.rand_plist(n, 0.3, seed=seed), chi, 1e-8, seed=seed).A


def test_17cell_1tdvp_chi256_matches_oracle(gpu):
    """1TDVP at fixed chi 256 (reference tdvp.py:65-105): 17 cells, distance 2,
    interval [2,3], the padded p=0.3 product state (seed 3), dt 0.05; bonds 8
    and 9 at 256 (512x256 complete QRs, 256x256 bond matrices).  Two steps
    against the oracle's trajectory."""
    g = golden("oracle_17cell_1tdvp_chi256")
    n = 17
    e = engine(padded_state(n), AU.mpo_tensors(n, 2, range(2, 4)), 1, 256, 0.05, 0.0)
    for s in range(int(g["params"][-1]) + 1):
        assert np.array_equal(e.bond_dims(), g["bond_dims"][s])
        assert np.abs(e.rho() - g["rho"][s]).max() < 1e-10, s
        if s < int(g["params"][-1]):
            e.step(1)
    assert int(g["bond_dims"][0].max()) == 256


def mpo_energy(As, Ws):
    """<psi|H|psi> by left environments (the oracle's environment layer,
    reference tdvp.py:328-347), on all host cores."""
    from threadpoolctl import threadpool_limits
    L = np.ones((1, 1, 1), dtype=complex)
    with threadpool_limits(16):
        for A, W in zip(As, Ws):
            L = OL.env_left(L, A, W)
    return L[0, 0, 0].real


def test_101cell_1tdvp_chi256_conserves_norm_and_energy(gpu):
    """The full 1TDVP workload (101 cells, chi 256, distance 2, interval [2,3],
    padded p=0.3 state, dt 0.05): projector-splitting 1TDVP at fixed bonds
    conserves the norm and <H> exactly up to the Krylov tolerance; the mixed
    path stays within 1e-6 of FP64 in <n_i> and its sites remain isometries."""
    n = 101
    Ws = AU.mpo_tensors(n, 2, range(2, 4))
    As = padded_state(n)
    import time
    t0 = time.perf_counter()
    e64 = engine([a.copy() for a in As], Ws, 1, 256, 0.05, 0.0)
    e32 = engine([a.copy() for a in As], Ws, 1, 256, 0.05, 0.0, prec=1)
    N0 = np.trace(e64.rho()[0]).real          # <psi|psi> (every site's rho has this trace)
    E0 = mpo_energy(e64.tensors(), Ws)
    print(f"  101-cell 1TDVP: engines built, <H> = {E0 / N0:.12f} ({time.perf_counter() - t0:.1f} s)",
          flush=True)
    for _ in range(2):
        e64.step(1)
        e32.step(1)
    N1 = np.trace(e64.rho()[0]).real
    E1 = mpo_energy(e64.tensors(), Ws)
    print(f"  101-cell 1TDVP: two steps, <H> = {E1 / N1:.12f}, norm drift {N1 - N0:.2e} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    assert abs(N1 - N0) < 1e-10 * N0
    assert abs(E1 / N1 - E0 / N0) < 1e-9 * max(1.0, abs(E0 / N0))
    assert np.array_equal(e64.bond_dims(), e32.bond_dims())
    assert np.abs(e32.rho()[:, 1, 1].real - e64.rho()[:, 1, 1].real).max() < 1e-6
    # right-isometric sites 1..N-1 (centre at site 0 after the left sweep)
    for A in e32.tensors()[1:]:
        M = A.transpose(1, 0, 2).reshape(A.shape[1], -1)
        assert np.abs(M @ M.conj().T - np.eye(A.shape[1])).max() < 1e-12


# ------------------------------------------------------------------ batched draws

def test_batch_with_chi128_draws_equals_alone_and_oracle(gpu):
    """The first 8 draws of the batched workload (default_rng(4); four with
    chi_max = 128, uncapped), both precisions: qca_batch_step over 4 workers
    gives bit-identical states to stepping each engine alone; the 15-cell
    draws (chi_max 32 and 128) match the oracle in FP64."""
    import workloads
    insts = workloads.batch_instances(8, seed=4)
    assert sum(i.chi_max == 128 for i in insts) >= 4
    steps = 4

    def engines(prec):
        return [engine(workloads.initial_tensors(i), workloads.mpo_tensors(i), 2, i.chi_max, i.step_size,
                       i.svd_epsilon, prec=prec) for i in insts]
    for prec in (0, 1):
        batched, alone = engines(prec), engines(prec)
        qn.batch_step(batched, steps, nthreads=4)
        for e in alone:
            e.step(steps)
        for b, a in zip(batched, alone):
            assert np.array_equal(b.bond_dims(), a.bond_dims())
            assert np.array_equal(b.rho(), a.rho())
        if prec == 0:
            for i, e in zip(insts, batched):
                if i.ncells > 15:
                    continue
                o = OT.TDVPOracle(OM.product_state(i.plist), AU.mpo_tensors(i.ncells, i.distance, i.interval),
                                  "2tdvp", 0.1, i.chi_max, 1e-10, gauge="consistent", expm="lanczos")
                for _ in range(steps):
                    o.do_time_step()
                assert np.array_equal(e.bond_dims(), OM.bond_dims(o.A).astype(int))
                assert np.abs(e.rho() - OM.rho_consistent(o.A)).max() < 1e-10
        for e in batched + alone:
            e.close()


# ------------------------------------------------------------------ served window

def test_served_window_with_large_request_equals_alone(gpu):
    """A served window at lambda = 50/s (default_rng(5)) with one large request
    (101 cells, chi 256, distance 2, interval [2,3]) among the batched-workload
    draws, steps capped at 12, <n_i> returned every 10 steps, 4 workers, mixed
    precision: every request reports exactly what the same instance reports
    when stepped alone."""
    import serve
    import workloads
    arrivals = serve.poisson_arrivals(rate=50.0, window_s=0.2, seed=5, big_fraction=0.0, max_steps=12)
    assert len(arrivals) >= 4
    rng = np.random.default_rng(7)
    big = workloads.Instance(101, 256, 2, range(2, 4), [1.0 if x < 0.3 else 0.0 for x in rng.random(101)])
    arrivals[1][1].inst = big
    make = serve.gpu_engine_factory(precision=1)
    out = serve.Server(make, workers=4).run(arrivals)
    assert out["requests"] == len(arrivals)
    for _, r in arrivals:
        e = make(r.inst)
        assert len(r.reports) == -(-r.steps // 10)
        for k, rep in enumerate(r.reports):
            e.step(min(10, r.steps - 10 * k))
            assert np.array_equal(e.rho()[:, 1, 1].real, rep)
        e.close()


def test_group_server_window_with_large_request_equals_alone(gpu):
    """The same served window through grouped continuous batching
    (serve.GroupServer: one qca_group_step per tick for all admitted
    requests): every request reports exactly what it reports alone."""
    import serve
    import workloads
    arrivals = serve.poisson_arrivals(rate=50.0, window_s=0.2, seed=5, big_fraction=0.0, max_steps=12)
    rng = np.random.default_rng(7)
    big = workloads.Instance(101, 256, 2, range(2, 4), [1.0 if x < 0.3 else 0.0 for x in rng.random(101)])
    arrivals[1][1].inst = big
    make = serve.gpu_engine_factory(precision=1)
    srv = serve.GroupServer(make)
    out = srv.run(arrivals)
    assert out["requests"] == len(arrivals) and out["group_builds"] >= 1
    assert 0 < srv.peak_in_use <= srv.budget
    for _, r in arrivals:
        e = make(r.inst)
        assert len(r.reports) == -(-r.steps // 10)
        for k, rep in enumerate(r.reports):
            e.step(min(10, r.steps - 10 * k))
            assert np.array_equal(e.rho()[:, 1, 1].real, rep)
        e.close()


# ------------------------------------------------------------------ observables

def test_observables_match_oracle_and_reference_golden(gpu):
    """Device observables (qca_engine_observables, reference mps.py:100-140):
    along a 12-cell trajectory, the full rho (off-diagonals included), the
    rounded population and the single-site entropy equal the oracle's
    (measure_consistent / rho_consistent); on the 15-cell single-cell run the
    first two rows equal the reference's own entropies and populations
    (golden config-1 fixture, before its stale environments act)."""
    n = 12
    pl = [0.0] * n
    pl[5] = pl[6] = pl[7] = 1.0
    Ws = AU.mpo_tensors(n, 2, range(2, 5))
    e = engine(OM.product_state(pl), Ws, 2, 16, 0.1, 1e-10)
    o = OT.TDVPOracle(OM.product_state(pl), Ws, "2tdvp", 0.1, 16, 1e-10, gauge="consistent", expm="lanczos")
    offdiag = 0.0
    for s in range(5):
        pop, dpop, sse, bd = e.observables()
        opop, odpop, osse, obd = OM.measure_consistent(o.A)
        orho = OM.rho_consistent(o.A)
        assert np.abs(e.rho() - orho).max() < 1e-10
        assert np.abs(pop - opop).max() < 1e-10
        assert np.array_equal(dpop, odpop) and np.array_equal(bd, obd)
        assert np.abs(sse - osse).max() < 1e-8
        offdiag = max(offdiag, np.abs(orho[:, 0, 1]).max())
        e.step(1)
        o.do_time_step()
    assert offdiag > 1e-3          # the off-diagonals are exercised
    g = golden("config1_2tdvp")
    n = 15
    e = engine(OM.product_state([1.0 if i == 7 else 0.0 for i in range(n)]), AU.mpo_tensors(n, 1, range(1, 2)),
               2, 16, 0.1, 1e-10)
    for s in range(2):
        pop, dpop, sse, bd = e.observables()
        assert np.abs(pop - g["population"][s]).max() < 1e-10
        assert np.array_equal(dpop, g["d_population"][s])
        assert np.abs(sse - g["single_site_entropy"][s]).max() < 1e-8
        assert np.array_equal(bd, g["bond_dims"][s])
        e.step(1)


# ------------------------------------------------------------------ error paths

def test_failed_creation_releases_device_memory(gpu, monkeypatch):
    """An engine whose construction fails after its allocations (here: an
    invalid QCA_MIXED_SVD_TOL, read after the MPS, environments and
    workspaces of the 41-cell chi=128 engine are allocated) frees them."""
    import torch
    n, pl, Ws = forty_one_cells()
    monkeypatch.setenv("QCA_MIXED_SVD_TOL", "not-a-number")
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(3):
        with pytest.raises(qn.QcaError, match="QCA_MIXED_SVD_TOL"):
            engine(OM.product_state(pl), Ws, 2, 128, 0.1, 1e-10, prec=1)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 < 64 << 20, (free0, free1)
    monkeypatch.delenv("QCA_MIXED_SVD_TOL")
    e = engine(OM.product_state(pl), Ws, 2, 128, 0.1, 1e-10, prec=1)
    e.step(1)
    e.close()


def test_svd_split_clustered_spectrum(gpu):
    """The split of a 256x256 two-site tensor whose spectrum has a degenerate
    top cluster, a cluster at 1e-8 (the size of the padding filler) and an
    exactly rank-deficient tail converges and matches LAPACK: singular values
    within 1e-13, and the truncation rank of the reference's rule
    (tdvp.py:289-292)."""
    import torch
    rng = np.random.default_rng(11)
    chil = chir = 128
    s = np.concatenate([np.full(8, 0.3), np.geomspace(0.1, 1e-4, 60), np.full(100, 1e-8),
                        np.zeros(88)])
    U, _ = np.linalg.qr(crand(rng, 256, 256))
    V, _ = np.linalg.qr(crand(rng, 256, 256))
    X = (U * s) @ V.conj().T
    X /= np.linalg.norm(X)
    theta = X.reshape(2, chil, 2, chir).transpose(0, 2, 1, 3).reshape(4, chil, chir)
    ref = np.linalg.svd(X, compute_uv=False)
    for chi_max, eps in ((128, 1e-10), (256, 1e-10), (256, 1e-7)):
        th = torch.from_numpy(np.ascontiguousarray(theta)).cuda()
        Al = torch.zeros((2, chil, min(chi_max, 256)), dtype=torch.complex128, device="cuda")
        Ar = torch.zeros((2, min(chi_max, 256), chir), dtype=torch.complex128, device="cuda")
        k = qn.c_int(0)
        sig = np.zeros(256)
        qn.call("qca_svd_split", qn.ptr(th), qn.ptr(Al), qn.ptr(Ar), chil, chir, chi_max, eps, 1,
                qn.ctypes.byref(k), sig.ctypes.data, None)
        assert np.abs(sig - ref).max() < 1e-13
        assert k.value == OL.truncation_rank(ref, chi_max, eps)


def test_exact_default_cadence_large_states(gpu):
    """The exact path at the reference's default cadence (step 0.005 x plot
    interval 200 = 1.0 per do_time_step, reference parser.py defaults and
    quantum_game.py:70-73): 18 cells, single and equal-superposition states
    (27 and 46 Lanczos iterations on the host), against the oracle's Krylov
    evolution on the host; 26 cells (24-vector
    basis, so the step is split into substeps) against two half steps, with
    the norm conserved."""
    import torch
    from parameters import Rules  # noqa: F401
    import ctypes
    for n, kind in ((18, "single"), (18, "equal")):
        pl = [0.5] * n if kind == "equal" else [1.0 if i == n // 2 else 0.0 for i in range(n)]
        As = OM.product_state(pl)
        psi = torch.from_numpy(OM.as_vector(As)).cuda()
        it = ctypes.c_int(0)
        qn.call("qca_exact_expm", qn.ptr(psi), n, 1, 1, 2, 1.0, 1e-14, ctypes.byref(it), None)
        print(f"  exact {n} cells ({kind}): GPU step {it.value} matvecs; host oracle", flush=True)
        ex = OE.ExactOracle(As, n, 1, range(1, 2), 1.0, mode="krylov")
        ex.do_time_step()
        assert np.abs(psi.cpu().numpy() - ex.psi).max() < 1e-10, (n, kind)
    n = 26
    pl = [0.5] * n
    v0 = OM.as_vector(OM.product_state(pl))
    a = torch.from_numpy(v0).cuda()
    b = a.clone()
    it = ctypes.c_int(0)
    qn.call("qca_exact_expm", qn.ptr(a), n, 2, 2, 5, 1.0, 1e-14, ctypes.byref(it), None)
    print(f"  exact {n} cells: one step of 1.0, {it.value} matvecs", flush=True)
    for _ in range(2):
        qn.call("qca_exact_expm", qn.ptr(b), n, 2, 2, 5, 0.5, 1e-14, ctypes.byref(it), None)
    print(f"  exact {n} cells: two steps of 0.5", flush=True)
    assert abs(torch.linalg.vector_norm(a).item() - 1.0) < 1e-12
    assert (a - b).abs().max().item() < 1e-10


def test_engine_footprint_equals_allocation(gpu):
    """qca_engine_footprint (host sizing, used for serving admission) equals
    the bytes the created engine holds, and the device's free memory drops by
    about that much (allocation granularity aside)."""
    import torch
    import workloads
    for inst in workloads.batch_instances(3, seed=4) + [workloads.Instance(101, 256, 2, range(2, 4),
                                                                         [0.0] * 101)]:
        Ws, As = workloads.mpo_tensors(inst), workloads.initial_tensors(inst)
        for prec in (0, 1):
            est = qn.engine_footprint(Ws, [1] * (inst.ncells + 1), algorithm=2, precision=prec,
                                      chi_max=inst.chi_max)
            torch.cuda.synchronize()
            free0 = qn.device_memory(0)[0]
            e = engine(As, Ws, 2, inst.chi_max, 0.1, 1e-10, prec=prec)
            free1 = qn.device_memory(0)[0]
            assert e.device_bytes() == est
            print(f"N={inst.ncells} chi={inst.chi_max} prec={prec}: footprint {est} B, "
                  f"free memory drop {free0 - free1} B")
            assert est <= (free0 - free1) + (64 << 20)
            assert free0 - free1 <= 1.25 * est + (256 << 20)
            e.close()
