"""GPU TDVP / exact engines against the oracle (gauge-consistent restatement).

The plain FP64 path must match the oracle's <n_i>(t) to 1e-10 with identical
bond dimensions; the mixed path must stay within 1e-6 of the FP64 path.
"""
import numpy as np
import pytest

from conftest import crand, golden

import qca_native as qn
from oracle import automaton as AU, exact as OE, mps as OM, tdvp as OT

pytestmark = pytest.mark.gpu


def gpu_trajectory(eng, steps, interval=1):
    pops, bds = [], []
    for step in range(steps):
        if step % interval == 0:
            pops.append(eng.rho()[:, 1, 1].real.copy())
            bds.append(eng.bond_dims())
        eng.step(1)
    return np.array(pops), np.array(bds)


def make_engine(As, Ws, algo, chi, dt, eps, prec=0, tol=0.0):
    return qn.Engine(Ws, As, algorithm=algo, precision=prec, chi_max=chi, dt=dt, eps=eps,
                     krylov_tol=tol or (1e-14 if prec == 0 else 1e-12))


CASES = [
    # alg, n, d, interval range, initial state, dt, steps, chi, eps
    ("2tdvp", 8, 1, range(1, 2), "single", 0.1, 10, 16, 1e-10),
    ("2tdvp", 12, 2, range(2, 5), "triple", 0.1, 8, 16, 1e-10),
    ("2tdvp", 15, 1, range(1, 2), "single", 0.1, 12, 16, 1e-10),
    ("2tdvp", 10, 2, range(1, 3), "rand", 0.05, 6, 8, 1e-6),
    ("2tdvp", 12, 2, range(2, 4), "padded", 0.1, 6, 16, 1e-10),
    ("1tdvp", 10, 1, range(1, 2), "padded", 0.1, 8, 8, 0.0),
    ("1tdvp", 11, 2, range(2, 4), "padded", 0.05, 6, 16, 0.0),
    ("1tdvp", 9, 2, range(2, 4), "triple", 0.1, 6, 16, 0.0),
    ("a1tdvp", 10, 1, range(1, 2), "single", 0.1, 6, 16, 1e-10),
    ("a1tdvp", 11, 2, range(2, 4), "padded", 0.05, 5, 8, 1e-6),
    ("a1tdvp", 12, 2, range(2, 5), "triple", 0.1, 5, 16, 1e-8),
    # bonds up to 64: splits of up to 128 columns (QR-preconditioned Gram Jacobi),
    # blocked QR in 1TDVP, both Krylov paths along the chain
    ("2tdvp", 14, 2, range(2, 5), "triple", 0.1, 10, 64, 1e-10),
    ("1tdvp", 14, 2, range(2, 4), "padded", 0.05, 6, 64, 0.0),
    # the shortest chains: one or two bonds, sweeps of a single block
    ("2tdvp", 2, 1, range(1, 2), "single", 0.1, 6, 4, 1e-10),
    ("2tdvp", 3, 2, range(1, 3), "triple", 0.1, 6, 4, 1e-10),
    ("1tdvp", 2, 1, range(1, 2), "padded", 0.1, 6, 4, 0.0),
    ("1tdvp", 3, 1, range(1, 2), "padded", 0.1, 6, 4, 0.0),
    ("a1tdvp", 3, 1, range(1, 2), "single", 0.1, 5, 4, 1e-10),
]

ALG_CODE = {"1tdvp": 1, "2tdvp": 2, "a1tdvp": 3}


def initial(kind, n, chi, seed=0):
    """Product states, or the bond-padded state (1e-3 filler) whose 1TDVP
    padding subspace is well conditioned."""
    if kind != "padded":
        return OM.product_state(plist_of(kind, n, seed))
    rng = np.random.default_rng(seed)
    pl = plist_of("rand", n, seed)
    dims = [min(2 ** i, 2 ** (n - i), chi) for i in range(n + 1)]
    out = []
    for i, p in enumerate(pl):
        A = 1e-3 * crand(rng, 2, dims[i], dims[i + 1])
        A[:, 0, 0] = ((1 - p) ** 0.5, p ** 0.5)
        out.append(A)
    return out


def plist_of(kind, n, seed=0):
    p = [0.0] * n
    mid = n // 2
    if kind == "single":
        p[mid] = 1.0
    elif kind == "triple":
        p[mid - 1] = p[mid] = p[mid + 1] = 1.0
    else:
        rng = np.random.default_rng(seed)
        p = list(np.where(rng.random(n) < 0.3, 1.0, 0.0))
    return p


def case_id(c):
    return f"{c[0]}-n{c[1]}-d{c[2]}-chi{c[7]}"


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_engine_matches_oracle(gpu, case):
    check_against_oracle(case)


@pytest.mark.parametrize("limit", [0, 4096], ids=["multi-launch", "one-launch"])
@pytest.mark.parametrize("case", [CASES[1], CASES[5], CASES[9]], ids=[case_id(c) for c in (CASES[1], CASES[5], CASES[9])])
def test_both_krylov_paths_match_oracle(gpu, monkeypatch, case, limit):
    """Every Krylov exponential on the multi-launch path (limit 0), or every
    one on the one-launch path (limit 4096 covers all blocks at chi <= 32)."""
    monkeypatch.setenv("QCA_FUSED_KRYLOV_MAX_N", str(limit))
    check_against_oracle(case)


@pytest.mark.parametrize("prec", [0, 1])
def test_one_launch_krylov_equals_multi_launch(gpu, monkeypatch, prec):
    """The one-launch exponential reproduces the multi-launch one on a chi=32
    2TDVP run (2-site blocks up to 4096 elements): FP64 to 1e-12 with the same
    bonds; mixed within the mixed gate of 1e-6.  Its device-side totals
    (matvecs, time) reach qca_engine_stats."""
    n, d, iv, chi = 12, 2, range(2, 5), 32
    runs = {}
    for limit in (0, 4096):
        monkeypatch.setenv("QCA_FUSED_KRYLOV_MAX_N", str(limit))
        e = make_engine(initial("triple", n, chi), AU.mpo_tensors(n, d, iv), 2, chi, 0.1, 1e-10, prec=prec)
        pops, bds = gpu_trajectory(e, 6)
        runs[limit] = (pops, bds, e.stats())
    # the same through the per-engine setter, starting from the default limit
    monkeypatch.delenv("QCA_FUSED_KRYLOV_MAX_N")
    e = make_engine(initial("triple", n, chi), AU.mpo_tensors(n, d, iv), 2, chi, 0.1, 1e-10, prec=prec)
    e.set_fused_krylov_limit(4096)
    with pytest.raises(qn.QcaError, match="0..4096"):
        e.set_fused_krylov_limit(4097)
    p2, b2 = gpu_trajectory(e, 6)
    assert np.array_equal(p2, runs[4096][0]) and np.array_equal(b2, runs[4096][1])
    (p0, b0, s0), (p1, b1, s1) = runs[0], runs[4096]
    if prec == 0:
        assert np.array_equal(b0, b1)
        assert np.abs(p0 - p1).max() < 1e-12
    else:
        assert np.abs(p0 - p1).max() < 1e-6
    assert s1["matvecs"] > 0 and s1["krylov_ms"] > 0
    assert abs(s1["matvecs"] - s0["matvecs"]) <= 0.05 * s0["matvecs"]
    assert s1["matvecs"] == s1["matvec_2site"] + s1["matvec_1site"] + s1["matvec_bond"]
    assert s1["flops"] > 0


def check_against_oracle(case):
    alg, n, d, iv, kind, dt, steps, chi, eps = case
    As = initial(kind, n, chi)
    Ws = AU.mpo_tensors(n, d, iv)
    o = OT.TDVPOracle([a.copy() for a in As], Ws, alg, dt, chi, eps, gauge="consistent", expm="lanczos")
    ref = OT.run_trajectory(o, steps, 1)
    eng = make_engine([a.copy() for a in As], Ws, ALG_CODE[alg], chi, dt, eps)
    pops, bds = gpu_trajectory(eng, steps)
    assert np.array_equal(bds, ref["bond_dims"].astype(int))
    # Tolerance per measured step: 1e-10, or ten times the trajectory's own
    # conditioning at that step when that is larger.  The conditioning is the
    # change of the oracle's populations when one input tensor is perturbed by
    # 1e-15 (relative).  1TDVP with fixed truncated bonds amplifies it ~10x per
    # step on some rules, so only late steps of those runs widen; every run's
    # first steps are held to 1e-10.
    As2 = [a.copy() for a in As]
    As2[n // 2 - 1] = As2[n // 2 - 1] * (1 + 1e-15)
    o2 = OT.TDVPOracle(As2, Ws, alg, dt, chi, eps, gauge="consistent", expm="lanczos")
    cond = np.abs(OT.run_trajectory(o2, steps, 1)["population"] - ref["population"]).max(axis=1)
    tol = np.maximum(1e-10, 10 * cond)
    assert tol.max() <= 1e-6, f"oracle trajectory too ill-conditioned to compare: {tol.max():.1e}"
    err = np.abs(pops - ref["population"]).max(axis=1)
    assert (err < tol).all(), f"per-step errors {err} against tolerances {tol}"


def test_engine_first_steps_match_reference_golden(gpu):
    """Before the reference's stale environments first matter (its first step
    from a product state, or the first steps of the padded runs) the GPU
    engine must reproduce the reference's own output (golden fixtures)."""
    g = golden("short_runs")
    for c in (0, 1, 3):
        alg = str(g[f"c{c}_alg"])
        n, d, lo, hi, dt, steps, chi, eps, itv = g[f"c{c}_params"]
        n, d, lo, hi, chi = int(n), int(d), int(lo), int(hi), int(chi)
        eng = make_engine(OM.product_state(g[f"c{c}_plist"]), AU.mpo_tensors(n, d, range(lo, hi)), 2,
                          chi, dt, eps)
        pops, bds = gpu_trajectory(eng, 2)
        assert np.abs(pops - g[f"c{c}_population"][:2]).max() < 1e-10
        assert np.array_equal(bds, g[f"c{c}_bond_dims"][:2].astype(int))
    p = golden("padded_runs")
    for c in (0, 2):
        alg = str(p[f"c{c}_alg"])
        n, d, lo, hi, dt, steps, chi, eps, itv = p[f"c{c}_params"]
        n, d, lo, hi, chi = int(n), int(d), int(lo), int(hi), int(chi)
        As = [p[f"c{c}_A{j}"] for j in range(n)]
        eng = make_engine(As, AU.mpo_tensors(n, d, range(lo, hi)), 1 if alg == "1tdvp" else 2, chi, dt, eps)
        pops, bds = gpu_trajectory(eng, 2)
        assert np.abs(pops - p[f"c{c}_population"][:2]).max() < 1e-9
        assert np.array_equal(bds, p[f"c{c}_bond_dims"][:2].astype(int))


@pytest.mark.parametrize("alg", [2, 1])
def test_full_rank_is_exact(gpu, alg):
    """At full bond dimension with a full-rank state, projector-splitting TDVP
    reproduces exact evolution (the oracle's exact 2^N propagation)."""
    n, d, iv, dt = 8, 1, range(1, 2), 0.1
    rng = np.random.default_rng(3)
    dims = [min(2 ** i, 2 ** (n - i)) for i in range(n + 1)]
    As = [crand(rng, 2, dims[i], dims[i + 1]) for i in range(n)]
    As[0] = As[0] / np.linalg.norm(OM.as_vector(As))
    eng = make_engine([a.copy() for a in As], AU.mpo_tensors(n, d, iv), alg, 16, dt, 1e-14)
    ex = OE.ExactOracle([a.copy() for a in As], n, d, iv, dt)
    for _ in range(6):
        assert np.abs(eng.rho()[:, 1, 1].real - ex.populations()).max() < 1e-11
        eng.step(1)
        ex.do_time_step()


def test_mixed_close_to_fp64(gpu):
    n, d, iv, dt, chi = 14, 2, range(2, 5), 0.1, 32
    pl = plist_of("triple", n)
    Ws = AU.mpo_tensors(n, d, iv)
    e64 = make_engine(OM.product_state(pl), Ws, 2, chi, dt, 1e-10, prec=0)
    e32 = make_engine(OM.product_state(pl), Ws, 2, chi, dt, 1e-10, prec=1)
    p64, _ = gpu_trajectory(e64, 10)
    p32, _ = gpu_trajectory(e32, 10)
    assert np.abs(p64 - p32).max() < 1e-6


def test_tdvp_algorithm_class_drop_in(gpu):
    """The Algorithm subclass with reference-style args and objects."""
    from parameters import Parser
    from tensor_networks import MPO, MPS
    from algorithms import TDVP, Exact
    args = Parser(["--num-cells", "9", "--algorithm", "2tdvp", "--step-size", "0.1",
                   "--max-bond-dim", "16", "--svd-epsilon", "1e-10", "--distance", "1"])
    H = MPO.hamiltonian_from_rules(args.rules)
    psi = MPS.from_density_distribution([0, 0, 0, 0, 1, 0, 0, 0, 0])
    alg = TDVP(psi, H, args)
    pop, dpop, sse, bd = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(10)
    ex_args = Parser(["--num-cells", "9", "--algorithm", "exact", "--step-size", "0.1"])
    ex = Exact(MPS.from_density_distribution([0, 0, 0, 0, 1, 0, 0, 0, 0]), H, ex_args)
    epop = np.zeros(9)
    for _ in range(5):
        alg.measure(pop, dpop, sse, bd)
        ex.measure(epop, np.zeros(9), np.zeros(9), np.zeros(10))
        alg.do_time_step()
        ex.do_time_step()
    assert alg.psi.is_valid_mps()
    assert abs(np.linalg.norm(alg.psi.as_vector()) - 1.0) < 1e-10
    # 2TDVP at chi 16 >= 2^(9//2) is exact after the initial (rank-one) steps only
    # up to the projection error of the first sweep, so compare loosely here; the
    # strict checks are test_engine_matches_oracle and test_full_rank_is_exact.
    assert np.abs(pop - epop).max() < 5e-2


def test_exact_class_matches_oracle(gpu):
    from parameters import Parser
    from tensor_networks import MPO, MPS
    from algorithms import Exact
    g = golden("exact_small")
    for c in range(5):
        n, d, lo, hi, dt, steps, itv = g[f"c{c}_params"]
        n, d, lo, hi, steps, itv = int(n), int(d), int(lo), int(hi), int(steps), int(itv)
        args = Parser(["--num-cells", str(n), "--distance", str(d), "--activation-interval",
                       str(lo), str(hi), "--step-size", str(dt * itv)])
        ex = Exact(MPS.from_density_distribution(g[f"c{c}_plist"]), MPO.hamiltonian_from_rules(args.rules), args)
        rows = g[f"c{c}_population"].shape[0]
        pop = np.zeros((rows, n))
        for r in range(rows):
            ex.measure(pop[r], np.zeros(n), np.zeros(n), np.zeros(n + 1))
            ex.do_time_step()
        assert np.abs(pop - g[f"c{c}_population"]).max() < 1e-10


def test_driver_cadence_and_files(gpu, tmp_path):
    """quantum_game: measure-then-step cadence, CSV rows [:plot_step] (the
    reference's off-by-one), npz checkpoint readable as an initial state."""
    from parameters import Parser
    from quantum_game import QuantumGame
    from tensor_networks import MPS
    args = Parser(["--num-cells", "10", "--algorithm", "2tdvp", "--step-size", "0.1", "--num-steps", "12",
                   "--plotting-frequency", "5", "--max-bond-dim", "16", "--svd-epsilon", "1e-10",
                   "--initial-states", "single"])
    assert args.plot_step_interval == 2 and args.plot_steps == 6
    out = QuantumGame(args).start(data_dir=str(tmp_path))[0]
    name = out["name"]
    csv = np.loadtxt(tmp_path / f"{name}-population.csv", delimiter=",")
    assert csv.shape == (5, 10)  # rows [:5] written at the last plot step
    assert np.allclose(csv, out["population"][:5])
    ck = MPS.from_file(str(tmp_path / f"{name}.npz"))
    assert ck.is_valid_mps() and abs(np.linalg.norm(ck.as_vector()) - 1) < 1e-10
    # the same run step by step through the oracle
    pl = [0.0] * 10
    pl[5] = 1.0
    o = OT.TDVPOracle(OM.product_state(pl), AU.mpo_tensors(10, 1, range(1, 2)), "2tdvp", 0.1, 16, 1e-10,
                      gauge="consistent", expm="lanczos")
    ref = OT.run_trajectory(o, 12, 2)
    assert np.abs(out["population"] - ref["population"]).max() < 1e-10


def test_batch_step_equals_sequential(gpu):
    """qca_batch_step over independent engines (one stream each, 4 host
    workers) gives bit-identical states to stepping each engine alone: the
    kernels are deterministic and engines share no device state."""
    import workloads
    insts = workloads.batch_instances(6, seed=4)
    def engines(prec):
        return [make_engine(workloads.initial_tensors(i), workloads.mpo_tensors(i), 2, min(i.chi_max, 32),
                            i.step_size, i.svd_epsilon, prec=prec) for i in insts]
    for prec in (0, 1):
        batched, alone = engines(prec), engines(prec)
        qn.batch_step(batched, 4, nthreads=4)
        for e in alone:
            e.step(4)
        for b, a in zip(batched, alone):
            assert np.array_equal(b.bond_dims(), a.bond_dims())
            assert np.array_equal(b.rho(), a.rho())
            assert b.stats()["site_updates"] == a.stats()["site_updates"]
    # one batched instance against the oracle
    i = insts[0]
    e = make_engine(workloads.initial_tensors(i), workloads.mpo_tensors(i), 2, 16, 0.1, 1e-10)
    qn.batch_step([e], 3, nthreads=2)
    o = OT.TDVPOracle(OM.product_state(i.plist), AU.mpo_tensors(i.ncells, i.distance, i.interval), "2tdvp",
                      0.1, 16, 1e-10, gauge="consistent", expm="lanczos")
    for _ in range(3):
        o.do_time_step()
    assert np.abs(e.rho()[:, 1, 1].real - OM.measure_consistent(o.A)[0]).max() < 1e-10


def test_server_on_gpu_matches_engines_run_alone(gpu):
    """Served requests (continuous batching over 3 workers, one engine and
    stream per request) report the same <n_i> every 10 steps as the same
    instances stepped alone."""
    import serve
    import workloads
    arrivals = serve.poisson_arrivals(rate=50.0, window_s=0.12, seed=5, big_fraction=0.0, max_steps=25)
    assert 2 <= len(arrivals) <= 12
    make = serve.gpu_engine_factory(precision=0, chi_cap=16)
    out = serve.Server(make, workers=3).run(arrivals)
    assert out["requests"] == len(arrivals) and out["p99_latency_s"] >= out["p50_latency_s"]
    for _, r in arrivals:
        e = make(r.inst)
        for k, rep in enumerate(r.reports):
            e.step(min(10, r.steps - 10 * k))
            assert np.array_equal(e.rho()[:, 1, 1].real, rep)
        e.close()


def test_a1tdvp_first_steps_match_reference_golden(gpu):
    """Adaptive 1TDVP: the engine takes the reference's first step (chosen
    bond targets and <n_i>) on every golden case."""
    g = golden("a1tdvp_runs")
    for c in range(4):
        n, d, lo, hi, dt, steps, chi, eps, itv = g[f"c{c}_params"]
        n = int(n)
        As = ([g[f"c{c}_A{j}"] for j in range(n)] if bool(g[f"c{c}_padded"])
              else OM.product_state(g[f"c{c}_plist"]))
        eng = make_engine([a.copy() for a in As], AU.mpo_tensors(n, int(d), range(int(lo), int(hi))), 3,
                          int(chi), dt, eps)
        pops, bds = gpu_trajectory(eng, 2)
        assert np.array_equal(bds, g[f"c{c}_bond_dims"][:2].astype(int))
        assert np.abs(pops - g[f"c{c}_population"][:2]).max() < 1e-9


def test_device_observables_match_host_post_processing(gpu):
    """qca_engine_observables (population, rounded population, single-site
    entropy, bond dims computed on the device) equals the host post-processing
    of the same density matrices (write_observables, eigvalsh-based entropy)."""
    from algorithms.algorithm import write_observables
    n = 12
    As = initial("triple", n, 16)
    eng = make_engine([a.copy() for a in As], AU.mpo_tensors(n, 2, range(2, 5)), 2, 16, 0.1, 1e-10)
    for step in range(4):
        rows = [np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n + 1)]
        eng.observables(*rows)
        ref = [np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n + 1)]
        write_observables(eng.rho(), eng.bond_dims(), *ref)
        assert np.array_equal(rows[0], ref[0])
        assert np.array_equal(rows[1], ref[1])
        assert np.abs(rows[2] - ref[2]).max() < 1e-12
        assert np.array_equal(rows[3], ref[3])
        if step:
            assert rows[2].max() > 1e-3        # entanglement has built up
        eng.step(1)
