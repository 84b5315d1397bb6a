"""Host-side modules of the drop-in (no GPU): MPO construction, parameters,
MPS conversions, observables post-processing, classical overlay."""
import numpy as np
import pytest

from conftest import crand, golden

from parameters import Parser, Rules
from tensor_networks import MPO, MPS
from algorithms.algorithm import Algorithm, entropy_bits, write_observables
from oracle import automaton as AU, mps as OM


def test_mpo_builder_matches_reference_golden():
    g = golden("mpo_tensors")
    for k in g.files:
        if not k.endswith("_bulk"):
            continue
        d, a, b = k.split("_")[:3]
        d, a, b = int(d[1:]), int(a), int(b)
        Ws = MPO.hamiltonian_from_rules(Rules(6, range(a, b), d, False)).W
        for tag, W in (("first", Ws[0]), ("bulk", Ws[1]), ("last", Ws[-1])):
            assert np.array_equal(W, g[f"d{d}_{a}_{b}_{tag}"]), (k, tag)


def test_mpo_matrix_is_closed_form():
    H = MPO.hamiltonian_from_rules(Rules(7, range(2, 5), 2, False)).as_matrix()
    assert np.abs(H - AU.dense_h(7, 2, range(2, 5))).max() == 0.0


def test_parser_defaults_and_cadence():
    p = Parser([])
    assert p.rules.ncells == 9 and p.rules.distance == 1
    assert p.rules.activation_interval == range(1, 2)
    assert p.algorithm == "exact" and p.max_bond_dim == 32 and p.svd_epsilon == 5e-5
    assert p.plot_step_interval == 200 and p.plot_steps == 50
    q = Parser(["--num-cells", "15", "--algorithm", "2tdvp", "--step-size", "0.1", "--num-steps", "100",
                "--max-bond-dim", "16", "--svd-epsilon", "1e-10", "--plotting-frequency", "10",
                "--precision", "mixed"])
    assert q.plot_step_interval == 1 and q.plot_steps == 100 and q.precision == "mixed"
    with pytest.raises(SystemExit):
        Parser(["--taylor-steps", "0"])


def test_mps_vector_roundtrip_and_file(tmp_path):
    rng = np.random.default_rng(0)
    n = 6
    psi = crand(rng, 1 << n)
    psi /= np.linalg.norm(psi)
    m = MPS.from_vector(psi)
    assert np.abs(m.as_vector() - psi).max() < 1e-13
    assert m.bond_dims() == [1, 2, 4, 8, 4, 2, 1]
    path = tmp_path / "state.npz"
    m.write_to_file(str(path))
    m2 = MPS.from_file(str(path))
    assert all(np.array_equal(a, b) for a, b in zip(m.A, m2.A))
    p = MPS.from_density_distribution([0.0, 1.0, 0.3])
    assert np.allclose(np.abs(p.as_vector()) ** 2, np.abs(OM.as_vector(OM.product_state([0.0, 1.0, 0.3]))) ** 2)


def test_observables_post_processing():
    rho = np.zeros((3, 2, 2), dtype=complex)
    rho[0] = [[1, 0], [0, 0]]
    rho[1] = [[0.5, 0], [0, 0.5]]
    rho[2] = [[0.3, 0.2j], [-0.2j, 0.7]]
    pop, dp, sse, bd = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(4)
    write_observables(rho, [1, 2, 2, 1], pop, dp, sse, bd)
    assert np.allclose(pop, [0, 0.5, 0.7]) and np.allclose(dp, [0, 0, 1])
    assert abs(sse[0]) < 1e-15 and abs(sse[1] - 1.0) < 1e-15
    w = np.linalg.eigvalsh(rho[2])
    assert abs(sse[2] + (w * np.log2(w)).sum()) < 1e-14
    assert list(bd) == [1, 2, 2, 1]
    assert entropy_bits(np.array([[0.0, 0.0], [0.0, 1.0]])) == 0.0


def test_classical_overlay():
    rules = Rules(7, range(1, 2), 1, False)
    first = np.array([0, 0, 0, 1, 0, 0, 0], dtype=float)
    h = Algorithm.classical_evolution(first, rules, 3)
    assert list(h[1]) == [0, 0, 1, 1, 1, 0, 0]
    assert list(h[2]) == [0, 1, 0, 1, 0, 1, 0]


def test_batch_workload_draws():
    """Batched workload: N ~ U{15..61}, chi ~ {16,32,64,128}, d ~ U{1,2},
    1 <= a <= b <= 2d, cells alive with probability 0.3, reproducible."""
    import workloads
    a = workloads.batch_instances(256, seed=4)
    b = workloads.batch_instances(256, seed=4)
    assert [(i.ncells, i.chi_max, i.distance, i.interval, i.plist) for i in a] == \
           [(i.ncells, i.chi_max, i.distance, i.interval, i.plist) for i in b]
    assert {i.chi_max for i in a} == {16, 32, 64, 128}
    assert {i.distance for i in a} == {1, 2}
    for i in a:
        assert 15 <= i.ncells <= 61 and len(i.plist) == i.ncells
        assert 1 <= i.interval.start <= i.interval.stop - 1 <= 2 * i.distance
        assert set(i.plist) <= {0.0, 1.0}
        assert i.site_updates_per_step == 2 * (i.ncells - 1)
    frac = np.mean([p for i in a for p in i.plist])
    assert 0.25 < frac < 0.35
    # prefixes are stable: the first k draws do not depend on the count
    assert [i.ncells for i in workloads.batch_instances(10, seed=4)] == [i.ncells for i in a[:10]]
    W = workloads.mpo_tensors(a[0])
    assert len(W) == a[0].ncells and W[0].shape[2] == 1
    A = workloads.initial_tensors(a[0])
    assert len(A) == a[0].ncells and A[0].shape == (2, 1, 1)


def test_serving_request_draws():
    import workloads
    rng = np.random.default_rng(5)
    reqs = [workloads.serving_request(rng) for _ in range(400)]
    big = [r for r, s in reqs if r.ncells == 101]
    assert 20 <= len(big) <= 60
    for r, s in reqs:
        assert 50 <= s <= 500
    assert all(r.chi_max == 256 and r.interval == range(2, 4) for r in big)


class _FakeEngine:
    """Host stand-in for a qca engine: fixed cost per step, <n_i> = steps done."""
    live = 0
    peak = 0
    lock = None

    def __init__(self, inst, cost):
        import threading
        self.n, self.cost, self.t = inst.ncells, cost, 0
        with _FakeEngine.lock:
            _FakeEngine.live += 1
            _FakeEngine.peak = max(_FakeEngine.peak, _FakeEngine.live)

    def step(self, k):
        import time
        time.sleep(self.cost * k)
        self.t += k

    def rho(self):
        r = np.zeros((self.n, 2, 2), complex)
        r[:, 1, 1] = self.t
        return r

    def close(self):
        with _FakeEngine.lock:
            _FakeEngine.live -= 1


def test_server_schedules_every_request():
    """Continuous batching on the host: every request gets <n> every 10 steps
    and its last report after exactly `steps` steps; admission keeps the
    footprint of live engines within the budget; latency statistics exist."""
    import threading
    import serve
    _FakeEngine.lock = threading.Lock()
    _FakeEngine.live = _FakeEngine.peak = 0
    arrivals = serve.poisson_arrivals(rate=200.0, window_s=0.2, seed=5, max_steps=60)
    assert len(arrivals) > 10
    sizes = sorted(serve.footprint_bytes(r.inst) for _, r in arrivals)
    budget = sizes[-1] + sizes[-2]          # at most a few requests at once
    srv = serve.Server(lambda inst: _FakeEngine(inst, 2e-5), workers=4, budget_bytes=budget)
    out = srv.run(arrivals)
    assert out["requests"] == len(arrivals)
    for _, r in arrivals:
        assert len(r.reports) == -(-r.steps // 10)
        assert r.reports[-1][0] == r.steps and r.engine is None
        assert r.finished >= r.arrival
    assert out["p99_latency_s"] >= out["p50_latency_s"] > 0
    assert out["site_updates"] == sum(r.steps * r.inst.site_updates_per_step for _, r in arrivals)
    assert _FakeEngine.live == 0
    assert srv._in_use == 0 and 0 < srv.peak_in_use <= budget


def test_server_rejects_oversized_and_reports_errors():
    import serve
    import workloads
    inst = workloads.Instance(101, 256, 2, range(2, 4), [0.0] * 101)
    srv = serve.Server(lambda i: None, workers=1, budget_bytes=1 << 20)
    with pytest.raises(ValueError):
        srv.submit(serve.Request(0, inst, 10, 0.0))

    class Boom:
        def __init__(self, inst):
            raise RuntimeError("device lost")
    small = workloads.batch_instances(2, seed=4)
    srv = serve.Server(Boom, workers=2, budget_bytes=200 << 30)
    with pytest.raises(RuntimeError, match="device lost"):
        srv.run([(0.0, serve.Request(i, s, 20, 0.0)) for i, s in enumerate(small)])


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and a
    message, prints no result line and touches no device."""
    import pathlib
    import subprocess
    import sys
    root = pathlib.Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""


def test_tdvp_new_state_keeps_precision_and_tolerances(monkeypatch):
    """Assigning TDVP.psi rebuilds the device engine with the algorithm,
    precision and Krylov tolerance chosen at construction (the engine class
    is replaced by a recorder, so no device is needed)."""
    from types import SimpleNamespace
    import algorithms.tdvp as T
    made = []

    class Recorder:
        def __init__(self, W, A, **kw):
            made.append(kw)

        def close(self):
            pass

    monkeypatch.setattr(T, "Engine", Recorder)
    args = SimpleNamespace(algorithm="2tdvp", precision="mixed", krylov_tol=0.0, max_bond_dim=8,
                           step_size=0.1, svd_epsilon=1e-10, device=0)
    t = T.TDVP(SimpleNamespace(A=[None]), SimpleNamespace(W=[None]), args)
    t.psi = SimpleNamespace(A=[None])
    assert len(made) == 2 and made[0] == made[1]
    assert made[0]["krylov_tol"] == 1e-10 and made[0]["precision"] == 1 and made[0]["algorithm"] == 2


def test_output_file_names_follow_the_reference():
    """Output file stems as reference quantum_game.py:49-58 builds them,
    including the a1tdvp method suffix."""
    from types import SimpleNamespace
    from quantum_game import QuantumGame
    g = object.__new__(QuantumGame)

    def args(alg, method="taylor"):
        rules = SimpleNamespace(ncells=15, distance=1, activation_interval=range(1, 2))
        return SimpleNamespace(rules=rules, step_size=0.1, algorithm=alg, max_bond_dim=16,
                               approximative_evolution_method=method, taylor_steps=5)

    g.args = args("2tdvp")
    assert g.file_name("single") == "single15_112_01_2tdvp_16"
    g.args = args("exact")
    assert g.file_name("single") == "single15_112_01_exact"
    g.args = args("a1tdvp")
    assert g.file_name("single") == "single15_112_01_a1tdvp_16_taylor_5"
    g.args = args("a1tdvp", "exact_exponential")
    assert g.file_name("0") == "015_112_01_a1tdvp_16_exact_exponential"


def test_named_states_match_reference_golden():
    """Every named initial state (states.NAMED, reference states.py:11-88)
    sets the same occupation probabilities as the reference's function of the
    same name, for several (ncells, distance) pairs; ``rand`` (unseeded in the
    reference) is checked for its support and its seeded reproducibility."""
    import states
    g = golden("named_states")
    names = [str(x) for x in g["names"]]
    for c, (n, d) in enumerate(g["cases"]):
        n, d = int(n), int(d)
        for name in names:
            ref = g[f"c{c}_{name}"]
            if name == "gradient_reverse":
                mps = states.gradient(n, reverse=True)
            else:
                mps = states.NAMED[name](n, d)
            got = np.array([abs(A[1, 0, 0]) ** 2 for A in mps.A])
            assert mps.A[0].shape == (2, 1, 1) and len(mps.A) == n
            if name == "rand":
                assert set(np.unique(got)) <= {0.0, 1.0} and set(np.unique(ref)) <= {0.0, 1.0}
                again = np.array([abs(A[1, 0, 0]) ** 2 for A in states.NAMED[name](n, d).A])
                assert np.array_equal(got, again)
                continue
            assert np.abs(got - ref).max() < 1e-15, (n, d, name)


def test_engine_footprint_is_host_computed_and_ordered():
    """qca_engine_footprint sizes an engine on the host (no device): it grows
    with the bond cap and the chain, mixed environments are smaller than
    complex128 ones, and the GPU serving factory admits by it."""
    import qca_native as qn
    import serve
    import workloads
    inst = workloads.batch_instances(1, seed=4)[0]
    Ws = workloads.mpo_tensors(inst)
    ch = [1] * (inst.ncells + 1)
    f = {(chi, prec): qn.engine_footprint(Ws, ch, algorithm=2, precision=prec, chi_max=chi)
         for chi in (16, 64, 128) for prec in (0, 1)}
    assert f[(16, 1)] < f[(64, 1)] < f[(128, 1)]
    assert f[(128, 1)] < f[(128, 0)]
    fac = serve.gpu_engine_factory(precision=1, chi_cap=64)
    assert fac.footprint(inst) == f[(min(inst.chi_max, 64), 1)]
    big = workloads.Instance(101, 256, 2, range(2, 4), [0.0] * 101)
    assert fac.footprint(big) < serve.gpu_engine_factory(precision=1).footprint(big)
    with pytest.raises(qn.QcaError, match="max bond"):
        qn.engine_footprint(Ws, ch, algorithm=2, precision=1, chi_max=300)


def test_group_server_schedules_every_request():
    """Grouped continuous batching on the host (stand-in engines and group):
    every request advances one step per tick of its lane while admitted,
    reports every 10 steps and after its last step, admission keeps the
    footprint within the budget, and a lane's group is rebuilt when its
    membership changes."""
    import threading
    import serve
    _FakeEngine.lock = threading.Lock()
    _FakeEngine.live = _FakeEngine.peak = 0

    class FakeGroup:
        built = []

        def __init__(self, engines):
            self.engines = list(engines)
            FakeGroup.built.append(len(self.engines))

        def step(self, k):
            for e in self.engines:
                e.step(k)

        def close(self):
            pass
    arrivals = serve.poisson_arrivals(rate=300.0, window_s=0.1, seed=5, max_steps=25)
    assert len(arrivals) > 10
    sizes = sorted(serve.footprint_bytes(r.inst) for _, r in arrivals)
    budget = sizes[-1] + sizes[-2] + sizes[-3]
    srv = serve.GroupServer(lambda inst: _FakeEngine(inst, 1e-5), budget_bytes=budget, group_factory=FakeGroup)
    out = srv.run(arrivals)
    assert out["requests"] == len(arrivals)
    for _, r in arrivals:
        assert len(r.reports) == -(-r.steps // 10)
        assert r.reports[-1][0] == r.steps and r.engine is None
    assert out["site_updates"] == sum(r.steps * r.inst.site_updates_per_step for _, r in arrivals)
    assert _FakeEngine.live == 0 and srv.in_use == 0 and 0 < srv.peak_in_use <= budget
    assert out["group_builds"] == len(FakeGroup.built) >= 2 and max(FakeGroup.built) >= 2
    # the draws span several size lanes, each stepped as its own group
    lanes = {serve.size_lane(r.inst) for _, r in arrivals}
    assert out["lanes"] == len(lanes) >= 2
