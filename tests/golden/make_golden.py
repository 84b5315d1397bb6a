"""Generate golden input/output vectors from the reference implementation.

Run in the development container only (the reference tree is not present on
GPU hosts):

    python tests/golden/make_golden.py [--only NAME ...]

The reference is imported read-only from /root/reference.  Its modules parse
``sys.argv`` at import time (reference ``tensor_networks/mps.py:11``,
``states.py:8``), so argv is replaced before the import.  Runs are driven
through ``algorithms.TDVP`` / ``algorithms.Exact`` directly, replaying the
measure-then-step cadence of ``quantum_game.py:82-119`` without its CSV/npz/plot
side effects.  Every fixture records the numpy/scipy versions, because SVD and
QR phase conventions are LAPACK-specific.

Fixtures are written as small ``.npz`` files next to this script; they hold
data only (inputs and expected outputs).
"""
import argparse
import os
import sys
import time

import numpy as np
import scipy

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
_CLI = list(sys.argv[1:])  # saved before the reference import replaces argv


def _import_reference():
    sys.argv = ["main.py", "--num-cells", "9"]
    if REF not in sys.path:
        sys.path.insert(0, REF)
    from parameters import Parser, Rules  # noqa: E402
    from tensor_networks import MPS, MPO  # noqa: E402
    from algorithms import TDVP, Exact  # noqa: E402
    import lautils  # noqa: E402
    return Parser, Rules, MPS, MPO, TDVP, Exact, lautils


Parser, Rules, MPS, MPO, TDVP, Exact, lautils = _import_reference()
ARGS = Parser.instance()


def _meta():
    return dict(numpy_version=np.__version__, scipy_version=scipy.__version__,
                openblas_threads=os.environ.get("OPENBLAS_NUM_THREADS", ""))


def padded_tensors(plist, chi, scale, seed):
    """Bond-padded product state (complex Gaussian filler of size ``scale``)."""
    n = len(plist)
    rng = np.random.default_rng(seed)
    dims = [min(2 ** i, 2 ** (n - i), chi) for i in range(n + 1)]
    out = []
    for i, p in enumerate(plist):
        shape = (2, dims[i], dims[i + 1])
        A = scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
        A[:, 0, 0] = ((1.0 - p) ** 0.5, p ** 0.5)
        out.append(A)
    return out


def run_reference(algorithm, ncells, distance, lo, hi, plist, dt, steps, chi, eps,
                  plot_interval, tensors=None):
    """Replay quantum_game.py:82-119 for one initial state.

    ``lo, hi`` is the half-open CLI activation interval (parser.py:182).
    Returns population, d_population, single_site_entropy, bond_dims arrays with
    one row per plot step, plus wall time.
    """
    rules = Rules(ncells=ncells, activation_interval=range(lo, hi), distance=distance,
                  periodic=False)
    ARGS.rules = rules
    ARGS.algorithm = algorithm
    ARGS.step_size = dt
    ARGS.max_bond_dim = chi
    ARGS.svd_epsilon = eps
    ARGS.num_steps = steps
    ARGS.plot_step_interval = plot_interval
    plot_steps = steps // plot_interval + (1 if steps % plot_interval else 0)
    H = MPO.hamiltonian_from_rules(rules)
    psi0 = (MPS.from_density_distribution(plist=list(plist)) if tensors is None
            else MPS.from_tensors([a.copy() for a in tensors]))
    pop = np.zeros((plot_steps, ncells))
    dpop = np.zeros((plot_steps, ncells))
    sse = np.zeros((plot_steps, ncells))
    bd = np.zeros((plot_steps, ncells + 1))
    if algorithm == "exact":
        ARGS.step_size = dt * plot_interval
        alg = Exact(psi_0=psi0, H=H, args=ARGS)
    else:
        alg = TDVP(psi_0=psi0, H=H, args=ARGS)
    t0 = time.perf_counter()
    step_times = []
    for step in range(steps):
        if step % plot_interval == 0:
            ps = step // plot_interval
            alg.measure(population=pop[ps], d_population=dpop[ps],
                        single_site_entropy=sse[ps], bond_dims=bd[ps])
            if algorithm == "exact":
                alg.do_time_step()
        if algorithm != "exact":
            ts = time.perf_counter()
            alg.do_time_step()
            step_times.append(time.perf_counter() - ts)
    wall = time.perf_counter() - t0
    ARGS.step_size = dt
    return dict(population=pop, d_population=dpop, single_site_entropy=sse,
                bond_dims=bd, wall_s=wall, step_times=np.array(step_times))


def single_plist(n, pos=None):
    p = [0.0] * n
    p[int(n / 2) if pos is None else pos] = 1.0
    return p


def blinker_plist(n):
    p = [0.0] * n
    mid = int(n / 2)
    p[mid - 1] = p[mid + 1] = 1.0
    return p


def _save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    meta = _meta()
    np.savez_compressed(path, **arrays, **{k: np.array(v) for k, v in meta.items()})
    print("wrote", path, flush=True)


# ---------------------------------------------------------------- fixtures

def g_mpo():
    """MPO bulk/boundary tensors for every (distance, half-open interval) pair used."""
    out = {}
    for d in (1, 2):
        for a in range(0, 2 * d + 1):
            for b in range(a + 1, 2 * d + 2):
                rules = Rules(ncells=6, activation_interval=range(a, b), distance=d,
                              periodic=False)
                W = MPO.hamiltonian_from_rules(rules).W
                key = f"d{d}_{a}_{b}"
                out[key + "_first"] = W[0]
                out[key + "_bulk"] = W[1]
                out[key + "_last"] = W[-1]
    _save("mpo_tensors", **out)


def g_exact():
    """Exact (dense) evolution for several rules and initial states, N = 6..10."""
    out = {}
    cases = [
        (6, 1, 1, 2, single_plist(6), 0.1, 12, 1),
        (8, 1, 1, 3, blinker_plist(8), 0.1, 12, 1),
        (9, 2, 2, 4, single_plist(9), 0.2, 8, 1),
        (10, 2, 2, 5, blinker_plist(10), 0.1, 20, 10),
        (7, 1, 2, 3, [0.3, 0.7, 0.0, 1.0, 0.5, 0.2, 0.9], 0.05, 10, 2),
    ]
    for i, (n, d, lo, hi, pl, dt, steps, itv) in enumerate(cases):
        r = run_reference("exact", n, d, lo, hi, pl, dt, steps, 0, 0.0, itv)
        out[f"c{i}_params"] = np.array([n, d, lo, hi, dt, steps, itv], dtype=float)
        out[f"c{i}_plist"] = np.array(pl, dtype=float)
        out[f"c{i}_population"] = r["population"]
        out[f"c{i}_single_site_entropy"] = r["single_site_entropy"]
    _save("exact_small", **out)


def g_fullchi():
    """2TDVP / 1TDVP at bond dim >= 2^(N/2) (no truncation) vs Exact, N=8.

    Tests whether the reference trajectory is gauge-consistent across steps.
    """
    out = {}
    n, d, lo, hi, dt, steps = 8, 1, 1, 2, 0.1, 20
    pl = single_plist(n)
    for itv in (1, 10):
        ex = run_reference("exact", n, d, lo, hi, pl, dt, steps, 0, 0.0, itv)
        out[f"exact_itv{itv}_population"] = ex["population"]
        for alg in ("2tdvp", "1tdvp"):
            r = run_reference(alg, n, d, lo, hi, pl, dt, steps, 16, 1e-12, itv)
            out[f"{alg}_itv{itv}_population"] = r["population"]
            out[f"{alg}_itv{itv}_bond_dims"] = r["bond_dims"]
            err = np.max(np.abs(r["population"] - ex["population"]))
            print(f"fullchi {alg} itv={itv}: max|dn| vs exact = {err:.3e}", flush=True)
    out["params"] = np.array([n, d, lo, hi, dt, steps], dtype=float)
    _save("fullchi_n8", **out)


def g_config1():
    """Config 1 in full: 2TDVP, N=15, d=1, [1,1] closed, single, dt=0.1, 100 steps,
    chi 16, eps 1e-10, measured every step (plotting frequency 10)."""
    r = run_reference("2tdvp", 15, 1, 1, 2, single_plist(15), 0.1, 100, 16, 1e-10, 1)
    print(f"config1 wall {r['wall_s']:.1f}s", flush=True)
    _save("config1_2tdvp", population=r["population"], d_population=r["d_population"],
          single_site_entropy=r["single_site_entropy"], bond_dims=r["bond_dims"],
          step_times=r["step_times"], params=np.array([15, 1, 1, 2, 0.1, 100, 16, 1e-10, 1]))


def g_short():
    """Short truncating runs at several rules (2TDVP and 1TDVP), for parity of the
    replayed reference semantics."""
    out = {}
    cases = [
        ("2tdvp", 10, 1, 1, 2, single_plist(10), 0.1, 6, 4, 1e-10, 1),
        ("2tdvp", 11, 2, 2, 5, [0, 0, 0, 0, 1, 1, 1, 0, 0, 0, 0], 0.1, 5, 8, 1e-10, 1),
        ("2tdvp", 12, 2, 2, 4, blinker_plist(12), 0.1, 4, 8, 1e-10, 2),
        ("2tdvp", 9, 1, 1, 3, [0.3, 0, 1, 0.5, 0, 0, 1, 0.1, 0], 0.05, 4, 6, 1e-8, 1),
        ("1tdvp", 10, 1, 1, 2, single_plist(10), 0.1, 5, 8, 1e-10, 1),
        ("1tdvp", 9, 2, 2, 4, blinker_plist(9), 0.1, 4, 4, 1e-10, 2),
    ]
    for i, (alg, n, d, lo, hi, pl, dt, steps, chi, eps, itv) in enumerate(cases):
        r = run_reference(alg, n, d, lo, hi, pl, dt, steps, chi, eps, itv)
        out[f"c{i}_alg"] = np.array(alg)
        out[f"c{i}_params"] = np.array([n, d, lo, hi, dt, steps, chi, eps, itv], dtype=float)
        out[f"c{i}_plist"] = np.array(pl, dtype=float)
        for k in ("population", "d_population", "single_site_entropy", "bond_dims"):
            out[f"c{i}_{k}"] = r[k]
    _save("short_runs", **out)


def g_units():
    """Single-call vectors of the local TDVP operations on seeded random inputs."""
    rng = np.random.default_rng(1234)

    def crand(*shape):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)

    def herm_env(chi, D):
        # Environment tensors are Hermitian in (ket, bra) for every MPO index
        # only when W is Hermitian; random ones suffice for the contraction tests.
        E = crand(chi, D, chi)
        return 0.5 * (E + np.conj(np.transpose(E, (2, 1, 0))))

    rules = Rules(ncells=6, activation_interval=range(2, 5), distance=2, periodic=False)
    W = MPO.hamiltonian_from_rules(rules).W[2]
    D = W.shape[2]
    out = {"W": W}
    # environment update, left and right
    chil, chir = 5, 7
    A = crand(2, chil, chir)
    Lp = herm_env(chil, D)
    Rp = herm_env(chir, D)
    out.update(env_A=A, env_Lprev=Lp, env_Rprev=Rp)
    out["env_Lnew"] = TDVP._assemble_new_layer_H_eff("left", Lp, A, W)
    out["env_Rnew"] = TDVP._assemble_new_layer_H_eff("right", Rp, A, W)
    # 1-site evolution (dense expm), both signs of the step
    A1 = crand(2, 4, 6)
    A1 /= np.linalg.norm(A1)
    L1, R1 = herm_env(4, D), herm_env(6, D)
    out.update(site_A=A1, site_L=L1, site_R=R1)
    out["site_fwd"] = TDVP._evolve_A(A1, L1, R1, W, 0.05)
    out["site_bwd"] = TDVP._evolve_A(A1, L1, R1, W, -0.05)
    # 2-site evolve + split + truncate
    Al, Ar = crand(2, 3, 5), crand(2, 5, 4)
    nrm = np.linalg.norm(np.tensordot(Al, Ar, (2, 1)))
    Al /= np.sqrt(nrm)
    Ar /= np.sqrt(nrm)
    L2, R2 = herm_env(3, D), herm_env(4, D)
    out.update(two_Al=Al, two_Ar=Ar, two_L=L2, two_R=R2)
    for chi, eps, tag in ((8, 1e-10, "a"), (3, 1e-10, "b"), (8, 1e-2, "c")):
        u, s, v = TDVP._evolve_split_and_truncate(Al, Ar, L2, R2, W, W, 0.05, chi, eps, 0)
        out[f"two_{tag}_u"], out[f"two_{tag}_s"], out[f"two_{tag}_v"] = u, s, v
        out[f"two_{tag}_chi_eps"] = np.array([chi, eps])
    # bond (K_eff) evolution: needs an instance for args.step_size; restate the call
    Lb, Rb = herm_env(5, D), herm_env(5, D)
    C = crand(5, 5)
    K = TDVP._assemble_K_eff(Lb, Rb)
    K = K.reshape(25, 25)
    out.update(bond_L=Lb, bond_R=Rb, bond_C=C)
    out["bond_out"] = lautils.timestep(K, C.reshape(-1), -0.05).reshape(5, 5)
    out["bond_step"] = np.array(-0.05)
    _save("units", **out)


def g_padded():
    """1TDVP and 2TDVP from bond-padded product states (well-conditioned
    padding subspace), measured every step and every 5 steps."""
    out = {}
    cases = [
        ("1tdvp", 10, 1, 1, 2, 0.1, 6, 8, 1e-10, 1, 11),
        ("1tdvp", 9, 2, 2, 4, 0.05, 10, 8, 1e-10, 5, 12),
        ("2tdvp", 10, 1, 1, 2, 0.1, 6, 8, 1e-10, 1, 13),
        ("2tdvp", 11, 2, 2, 5, 0.1, 10, 8, 1e-10, 5, 14),
    ]
    for i, (alg, n, d, lo, hi, dt, steps, chi, eps, itv, seed) in enumerate(cases):
        rng = np.random.default_rng(seed)
        pl = [1.0 if x < 0.3 else 0.0 for x in rng.random(n)]
        As = padded_tensors(pl, chi, 1e-3, seed)
        r = run_reference(alg, n, d, lo, hi, pl, dt, steps, chi, eps, itv, tensors=As)
        out[f"c{i}_alg"] = np.array(alg)
        out[f"c{i}_params"] = np.array([n, d, lo, hi, dt, steps, chi, eps, itv], dtype=float)
        for j, A in enumerate(As):
            out[f"c{i}_A{j}"] = A
        for k in ("population", "d_population", "single_site_entropy", "bond_dims"):
            out[f"c{i}_{k}"] = r[k]
    _save("padded_runs", **out)


def g_a1tdvp():
    """Adaptive 1TDVP (a1tdvp) with --approximative-evolution-method
    exact_exponential (the default 'taylor' needs numpy.math, removed in numpy 2;
    tdvp.py:247), from product states and from a bond-padded state, measured
    every step."""
    ARGS.approximative_evolution_method = "exact_exponential"
    out = {}
    cases = [
        (10, 1, 1, 2, single_plist(10), 0.1, 4, 16, 1e-10, None),
        (11, 2, 2, 4, [0, 0, 0, 1, 1, 0, 1, 0, 0, 0, 0], 0.1, 3, 8, 1e-6, None),
        (9, 1, 1, 3, [0.3, 0, 1, 0.5, 0, 0, 1, 0.1, 0], 0.05, 3, 8, 1e-8, None),
        (10, 1, 1, 2, [0, 1, 0, 0, 1, 0, 0, 0, 1, 0], 0.1, 4, 8, 1e-6, 21),
    ]
    for i, (n, d, lo, hi, pl, dt, steps, chi, eps, seed) in enumerate(cases):
        As = padded_tensors(pl, chi, 1e-3, seed) if seed is not None else None
        r = run_reference("a1tdvp", n, d, lo, hi, pl, dt, steps, chi, eps, 1, tensors=As)
        out[f"c{i}_params"] = np.array([n, d, lo, hi, dt, steps, chi, eps, 1], dtype=float)
        out[f"c{i}_plist"] = np.array(pl, dtype=float)
        out[f"c{i}_padded"] = np.array(seed is not None)
        if As is not None:
            for j, A in enumerate(As):
                out[f"c{i}_A{j}"] = A
        for k in ("population", "d_population", "single_site_entropy", "bond_dims"):
            out[f"c{i}_{k}"] = r[k]
    _save("a1tdvp_runs", **out)


def g_states():
    """Occupation probabilities of the reference's named initial states
    (states.py:11-88) for a few (ncells, distance) pairs, read back from the
    product MPS each function builds (A[i][:, 0, 0] = (sqrt(1-p), sqrt(p))).

    The reference's functions take ncells and distance from the parsed
    arguments at call time, except ``single``, whose default position is
    bound at import time to the command line's ncells (states.py:34); it is
    called here with that binding, ``single(int(ncells / 2))``, as a run with
    ``--num-cells ncells`` sees it.  ``rand`` draws from the unseeded
    ``random`` module (states.py:86-88) and is recorded for its support only."""
    import states as ref_states
    names = ["blinker", "triple_blinker", "full_blinker", "single", "single_bottom", "all_ket_0",
             "all_ket_1", "only_outer", "all_ket_1_but_outer", "equal_superposition",
             "equal_superposition_but_outer", "gradient", "gradient_reverse", "rand"]
    out = {"names": np.array(names)}
    cases = [(9, 1), (12, 2), (15, 1), (41, 2)]
    out["cases"] = np.array(cases)
    for c, (n, d) in enumerate(cases):
        ARGS.rules = Rules(ncells=n, activation_interval=range(1, 2), distance=d, periodic=False)
        for name in names:
            if name == "single":
                mps = ref_states.single(int(n / 2))
            elif name == "gradient_reverse":
                mps = ref_states.gradient(reverse=True)
            else:
                mps = getattr(ref_states, name)()
            out[f"c{c}_{name}"] = np.array([abs(A[1, 0, 0]) ** 2 for A in mps.A])
    _save("named_states", **out)


ALL = dict(padded=g_padded, mpo=g_mpo, exact=g_exact, fullchi=g_fullchi, units=g_units, short=g_short,
           config1=g_config1, a1tdvp=g_a1tdvp, states=g_states)

if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=list(ALL), choices=list(ALL))
    ns = ap.parse_args(_CLI)
    for name in ns.only:
        ALL[name]()
