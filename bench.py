"""Benchmark: 2TDVP site-updates/s at chi=128 on one MI355X, mixed vs FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the 41-cell single-instance 2TDVP case): N=41 cells, distance 2,
activation interval [2,4] closed (Python range(2,5)), cells 19..21 alive,
dt=0.1, chi_max=128, SVD cutoff 1e-10.  Preparation (fixed, independent of
the flags): a plain FP64 engine takes PREP_STEPS = 30 steps from the product
state, after which 26 of the 40 bonds sit at chi_max; a mixed engine (complex64
contractions on FP32 matrix cores) is built from that state.  Both engines then
take W untimed warmup steps, and EXACTLY K time steps of each are timed between
synchronisations.

A site-update is one two-site block update (2(N-1) per time step).  ``value``
is the mixed engine's site-updates/s on the saturated state; the FP64 rate,
their ratio, the largest |<n_i>_mixed - <n_i>_fp64| over the timed steps and
the number of bonds at chi_max at the start and end of the timed region are
reported beside it.

``batched``: B copies of the prepared chi=128 state advanced together (mixed
precision), against the plain FP64 single-instance rate above: the
batched-vs-plain ratio of the workload's target.
``served``: a window of Poisson arrivals (batched-workload draws plus 10%
101-cell chi-256 requests) served with continuous batching; p50/p99 latency.

``roofline``: the dominant contraction, the two-site effective-Hamiltonian
application (GEMM over the left environment -> sparse MPO mixing -> GEMM over
the right environment) at the bulk shape chi_l = chi_r = 128, D = 13, timed
with HIP events on its stream; achieved = its algorithmic flop count
8[4 D chi^2 (2 chi) + 16 D^2 chi^2] / average duration.

``cpu_baseline``: the oracle (NumPy complex128, matrix-free Lanczos) running
2TDVP site-updates on the same chi=128 state on the host cores for a bounded
time; it is a reported baseline, not the target.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "quantum-cellular-automaton_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

N, DIST, IV, DT, CHI, EPS = 41, 2, range(2, 5), 0.1, 128, 1e-10
# FP64 steps that bring the 41-cell state to chi_max (26 of 40 bonds at 128)
# before any warmup or timed step
PREP_STEPS = 30

# HBM bytes moved by one 2-site H_eff application at chi=128, D=13, measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes over `bench.py --roofline-only`
# (the counters scaled by a 256 MiB device copy in the same runs), recorded in
# profiles/heff2_chi128_d13_hbm_bytes.md: profile-derived, not measured by this
# run.  The algorithmic minimum is 4.46 MB (c64).
HEFF2_TRAFFIC_BYTES = {"mixed": 48.39e6, "fp64": 96.41e6}
HEFF2_TRAFFIC_SOURCE = "profiles/heff2_chi128_d13_hbm_bytes.md"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def initial_state():
    import states
    return states.three_adjacent(N)


def mpo():
    from parameters import Rules
    from tensor_networks import MPO
    return MPO.hamiltonian_from_rules(Rules(N, IV, DIST, False))


def heff2_roofline(torch, prec: int, reps: int = 50):
    """Time the 2-site H_eff application at chi=128, D=13 with HIP events."""
    import qca_native as qn
    from tensor_networks import MPO
    W = mpo().W[N // 2]
    W2 = MPO.merge_mpo_tensor_pair(W, W)
    D = W.shape[2]
    chi = CHI
    rng = np.random.default_rng(0)
    ct = torch.complex128 if prec == 0 else torch.complex64

    def rnd(*shape):
        return torch.from_numpy((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)).to(ct).cuda()

    x, L, R = rnd(4, chi, chi), rnd(chi, D, chi), rnd(chi, D, chi)
    y = torch.empty_like(x)
    Wh = np.ascontiguousarray(W2, dtype=np.complex128)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    for _ in range(3):
        qn.call("qca_heff_apply", prec, 0, qn.ptr(x), qn.ptr(y), qn.ptr(L), qn.ptr(R), Wh.ctypes.data,
                D, D, chi, chi, sp)
    torch.cuda.synchronize()
    # qca_heff_apply only queues its kernels (the sparse MPO is uploaded on the
    # first call), so `reps` back-to-back applications between two events on
    # that stream give the average device time of one application.
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        qn.call("qca_heff_apply", prec, 0, qn.ptr(x), qn.ptr(y), qn.ptr(L), qn.ptr(R), Wh.ctypes.data,
                D, D, chi, chi, sp)
    e1.record(stream)
    e1.synchronize()
    total = e0.elapsed_time(e1)
    ms = total / reps
    flops = 8.0 * (4.0 * D * chi * chi * (2 * chi) + 16.0 * D * D * chi * chi)
    return ms, flops


def batch_bench(torch, a):
    """Aggregate 2TDVP site-updates/s of the first `a.batch` instances of the
    batched workload (mixed precision), advanced one step at a time by
    `a.batch_threads` native workers, one stream per instance; the same engines
    stepped one after another on one host thread give the sequential rate,
    stepped as one group per size lane the grouped rate, and plain FP64
    engines started from the same warmed states, stepped one after another,
    the rate of the plain path."""
    import qca_native as qn
    import workloads
    insts = workloads.batch_instances(256, seed=4)[:a.batch]
    log(f"batched workload: {len(insts)} instances, building engines")
    engines = [qn.Engine(workloads.mpo_tensors(i), workloads.initial_tensors(i), algorithm=2, precision=1,
                         chi_max=i.chi_max, dt=i.step_size, eps=i.svd_epsilon, krylov_tol=1e-10)
               for i in insts]
    qn.batch_step(engines, a.batch_warmup, a.batch_threads)
    torch.cuda.synchronize()
    bonds = [int(e.bond_dims().max()) for e in engines]
    log(f"batched workload: warm after {a.batch_warmup} steps, max bond per instance "
        f"median {int(np.median(bonds))} max {max(bonds)}")
    rates = {}
    # grouped: one group per size lane (serve.size_lane), the lanes' groups
    # stepped side by side on their own host threads
    import serve
    lane_members: dict = {}
    for e, i in zip(engines, insts):
        lane_members.setdefault(serve.size_lane(i), []).append(e)

    def grouped_run():
        groups = [qn.Group(es) for es in lane_members.values()]
        errors = []

        def one(g):
            try:
                g.step(a.batch_steps)
            except BaseException as exc:  # re-raised below, after every lane has stopped
                errors.append(exc)
        threads = [threading.Thread(target=one, args=(g,)) for g in groups]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        for g in groups:
            g.close()
        if errors:
            raise errors[0]
    # one instance at a time favours latency (one-launch Krylov up to 1024
    # elements); many in flight favour throughput (up to 4096)
    for label, limit, run in (("sequential", 1024, lambda: [e.step(a.batch_steps) for e in engines]),
                              ("batched", 4096, lambda: qn.batch_step(engines, a.batch_steps, a.batch_threads)),
                              ("grouped", 4096, grouped_run)):
        for e in engines:
            e.set_fused_krylov_limit(limit)
            e.reset_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        su = sum(e.stats()["site_updates"] for e in engines)
        rates[label] = su / dt
        log(f"batched workload: {label} {su} site-updates in {dt:.2f} s")
    # the plain FP64 path on the same warmed states, one instance at a time
    plain = [qn.Engine(workloads.mpo_tensors(i), e.tensors(), algorithm=2, precision=0,
                       chi_max=i.chi_max, dt=i.step_size, eps=i.svd_epsilon)
             for i, e in zip(insts, engines)]
    for e in plain:
        e.set_fused_krylov_limit(1024)
        e.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in plain:
        e.step(a.batch_steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    su = sum(e.stats()["site_updates"] for e in plain)
    rates["fp64_sequential"] = su / dt
    log(f"batched workload: plain FP64 one after another {su} site-updates in {dt:.2f} s")
    for e in engines + plain:
        e.close()
    return {"workload": f"first {len(insts)} instances of the batched 2TDVP draw (default_rng(4)): "
                        "N~U{15..61}, chi_max~{16,32,64,128}, d~U{1,2}, 1<=a<=b<=2d, p=0.3, dt 0.1, "
                        "eps 1e-10, mixed precision",
            "instances": len(insts), "threads": a.batch_threads, "warmup_steps": a.batch_warmup,
            "timed_steps": a.batch_steps, "max_bond_median": int(np.median(bonds)),
            "gpu_max_hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
            "one_launch_krylov_limit": {"sequential": 1024, "batched": 4096},
            "site_updates_per_s": rates["batched"], "sequential_site_updates_per_s": rates["sequential"],
            "speedup_over_sequential": rates["batched"] / rates["sequential"],
            "fp64_sequential_site_updates_per_s": rates["fp64_sequential"],
            "speedup_over_fp64_sequential": rates["batched"] / rates["fp64_sequential"],
            "grouped": {"lanes": len(lane_members), "site_updates_per_s": rates["grouped"],
                        "speedup_over_fp64_sequential": rates["grouped"] / rates["fp64_sequential"],
                        "method": "one group per size lane (qca_group_step), lanes on their own threads"}}


def batch128_bench(torch, a, tensors, fp64_rate):
    """B copies of the prepared 41-cell chi=128 state (every instance starts
    with 26+ of its 40 bonds at chi_max), mixed precision, advanced together
    for `a.batch128_warmup` untimed and `a.batch128_steps` timed steps, two
    ways: as one group in lockstep with grouped launches (qca_group_step, the
    reported rate) and, for comparison, one stream per instance driven by
    `a.batch_threads` native workers (qca_batch_step).  The ratio against the
    plain FP64 single-instance rate of the headline is the batched-vs-plain
    figure of the workload's target."""
    import qca_native as qn
    Ws = mpo().W
    B = a.batch128
    out = {"workload": f"{B} instances of the 41-cell chi=128 2TDVP case, each from the prepared "
                       f"state ({PREP_STEPS} FP64 steps), mixed precision",
           "instances": B, "warmup_steps": a.batch128_warmup, "timed_steps": a.batch128_steps,
           "fp64_single_site_updates_per_s": fp64_rate}
    for method in ("grouped", "streams"):
        if method == "streams":
            B = a.batch128_streams
            if not B:
                continue
        log(f"batched chi=128 ({method}): {B} instances of the prepared state")
        engines = [qn.Engine(Ws, tensors, algorithm=2, precision=1, chi_max=CHI, dt=DT, eps=EPS,
                             krylov_tol=a.mixed_krylov_tol, fused_krylov_max_n=4096) for _ in range(B)]
        if method == "grouped":
            grp = qn.Group(engines)
            run = grp.step
        else:
            grp = None
            run = lambda k: qn.batch_step(engines, k, a.batch_threads)  # noqa: E731
        run(a.batch128_warmup)
        for e in engines:
            e.reset_stats()
        bd0 = engines[0].bond_dims()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(a.batch128_steps)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        su = sum(e.stats()["site_updates"] for e in engines)
        bd1 = engines[0].bond_dims()
        rate = su / dt
        speedup = rate / fp64_rate if fp64_rate == fp64_rate else None   # None: FP64 rate not measured
        res = {"instances": B, "site_updates": su, "seconds": dt, "site_updates_per_s": rate,
               "speedup_over_fp64_single": speedup,
               "bonds_at_chi_max": {"start": int((bd0 == CHI).sum()), "end": int((bd1 == CHI).sum())}}
        if grp is not None:
            res["group_stats"] = grp.stats()
            res["method"] = "one group, grouped launches in lockstep (qca_group_step)"
            grp.close()
        else:
            res["method"] = "one stream per instance, native worker threads (qca_batch_step)"
            res["threads"] = a.batch_threads
            res["gpu_max_hw_queues"] = os.environ.get("GPU_MAX_HW_QUEUES")
        log(f"batched chi=128 ({method}): {su} site-updates in {dt:.2f} s = {rate:.1f}/s"
            + (f" ({speedup:.2f}x the plain FP64 single instance)" if speedup is not None else ""))
        out[method] = res
        for e in engines:
            e.close()
    out["site_updates_per_s"] = out["grouped"]["site_updates_per_s"]
    out["speedup_over_fp64_single"] = out["grouped"]["speedup_over_fp64_single"]
    return out


def tdvp1_bench(torch, a):
    """1TDVP at fixed chi = 256: 101 cells, distance 2, interval [2,3], each
    cell alive with p = 0.3 (seed 3), bonds padded to min(2^i, 2^(N-i), 256)
    with complex Gaussian filler x 1e-8 (seed 3), dt 0.05.  Both precisions
    run `a.tdvp1` time steps from the same state; 2N site-updates per step."""
    import qca_native as qn
    import states
    from parameters import Rules
    from tensor_networks import MPO
    n, chi, dt = 101, 256, 0.05
    Ws = MPO.hamiltonian_from_rules(Rules(n, range(2, 4), 2, False)).W
    psi = states.padded(states.rand_plist(n, 0.3, seed=3), chi, 1e-8, seed=3)
    out, pops = {}, {}
    for name, prec, tol in (("fp64", 0, 1e-14), ("mixed", 1, a.mixed_krylov_tol)):
        log(f"1TDVP chi={chi}: building the {name} engine")
        e = qn.Engine(Ws, psi.A, algorithm=1, precision=prec, chi_max=chi, dt=dt, eps=1e-10,
                      krylov_tol=tol)
        e.reset_stats()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e.step(a.tdvp1)
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t0
        st = e.stats()
        pops[name] = e.rho()[:, 1, 1].real.copy()
        out[name] = {"s_per_step": dt_s / a.tdvp1, "site_updates_per_s": st["site_updates"] / dt_s,
                     "krylov_ms_per_step": st["krylov_ms"] / a.tdvp1, "matvecs_per_step": st["matvecs"] / a.tdvp1}
        log(f"1TDVP chi={chi} {name}: {dt_s / a.tdvp1:.3f} s/step")
        e.close()
    return {"workload": "1TDVP, 101 cells, distance 2, interval [2,3], p=0.3 product state (seed 3) "
                        "padded to chi 256 with 1e-8 filler (seed 3), dt 0.05",
            "steps": a.tdvp1, **out,
            "max_abs_dn_mixed_vs_fp64": float(np.abs(pops["mixed"] - pops["fp64"]).max())}


def exact_apply_bench(torch, ncells: int, reps: int = 20):
    """The exact path's closed-form H application (reference exact.py via
    mpo.py:202-228) on a 2^N state, distance 2, interval [2,4]: HIP-event time
    per application on its stream, and the HBM rate against its algorithmic
    bytes (psi read once, H psi written once)."""
    import qca_native as qn
    out = {"ncells": ncells, "rule": "distance 2, interval [2,4]"}
    for name, prec, ct in (("fp64", 0, torch.complex128), ("mixed", 1, torch.complex64)):
        g = torch.Generator(device="cuda").manual_seed(0)
        psi = torch.randn(1 << ncells, dtype=ct, device="cuda", generator=g)
        y = torch.empty_like(psi)
        st = torch.cuda.current_stream()
        for _ in range(2):
            qn.call("qca_exact_apply_h", prec, qn.ptr(psi), qn.ptr(y), ncells, 2, 2, 5, st.cuda_stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            qn.call("qca_exact_apply_h", prec, qn.ptr(psi), qn.ptr(y), ncells, 2, 2, 5, st.cuda_stream)
        e1.record(st)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        nbytes = 2 * psi.element_size() * (1 << ncells)
        out[name] = {"ms_per_apply": ms, "algorithmic_bytes": nbytes,
                     "achieved_GBps": nbytes / (ms * 1e-3) / 1e9,
                     "frac_of_8TBps": nbytes / (ms * 1e-3) / 8e12}
        log(f"exact H apply 2^{ncells} {name}: {ms:.3f} ms, {nbytes / (ms * 1e-3) / 1e9:.0f} GB/s")
        del psi, y
    return out


def serve_bench(a):
    """A served window: Poisson arrivals at `a.serve` requests/s for
    `a.serve_window` s (default_rng(5)), each request a batched-workload draw
    (or, with probability `a.serve_big_fraction`, the 101-cell chi-256 case),
    steps capped at `a.serve_max_steps` and chi at `a.serve_chi_cap`; <n_i>
    returned every 10 steps; mixed precision.  Admission by each engine's exact
    device footprint against the device's free memory; every admitted request
    advances one step per tick in one grouped step (serve.GroupServer), or with
    --serve-method streams on its own stream with `a.batch_threads` workers."""
    import serve
    arrivals = serve.poisson_arrivals(a.serve, a.serve_window, seed=5, big_fraction=a.serve_big_fraction,
                                      max_steps=a.serve_max_steps)
    log(f"served window ({a.serve_method}): {len(arrivals)} requests over {a.serve_window:g} s")
    cap = a.serve_chi_cap or None
    factory = serve.gpu_engine_factory(precision=1, chi_cap=cap)
    if a.serve_method == "group":
        srv = serve.GroupServer(factory)
    else:
        srv = serve.Server(factory, workers=a.batch_threads)
    out = srv.run(arrivals)
    out.update({"arrival_rate_per_s": a.serve, "window_s": a.serve_window, "max_steps": a.serve_max_steps,
                "chi_cap": a.serve_chi_cap, "big_fraction": a.serve_big_fraction, "report_every_steps": 10,
                "big_requests": sum(1 for _, r in arrivals if r.inst.ncells == 101)})
    log(f"served window: p50 {out['p50_latency_s']:.2f} s p99 {out['p99_latency_s']:.2f} s, "
        f"{out['site_updates_per_s']:.0f} site-updates/s")
    return out


def cpu_baseline(As, seconds: float):
    """Oracle 2TDVP site-updates on the host at the benchmark state (bounded)."""
    from oracle import automaton as AU, tdvp as OT
    cores = len(os.sched_getaffinity(0))
    Ws = AU.mpo_tensors(N, DIST, IV)
    o = OT.TDVPOracle([a.copy() for a in As], Ws, "2tdvp", DT, CHI, EPS, gauge="consistent",
                      expm="lanczos")
    h = DT / 2
    done = 0
    t0 = time.perf_counter()
    # right half-sweep bond by bond (reference tdvp.py:107-125) until the budget is spent
    for site in range(N - 1):
        u, s, v = o._evolve_split(site, h)
        o.A[site] = u
        o.A[site + 1] = s[None, :, None] * v
        if site < N - 2:
            o._update_env("left", site)
            o.A[site + 1] = o._evolve_site(site + 1, -h)
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle (NumPy complex128, matrix-free Lanczos) 2TDVP right half-sweep from the "
                      f"warmed chi<=128 state: {done} site-updates in {dt:.1f} s, "
                      f"BLAS threads = {os.environ.get('OPENBLAS_NUM_THREADS', 'default')}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-krylov-tol", type=float, default=1e-10,
                    help="Krylov exponential tolerance of the mixed engine (the plain one uses 1e-14); "
                         "its complex64 matvecs are accurate to ~1e-7")
    ap.add_argument("--batch128", type=int, default=64,
                    help="instances of the batched chi=128 leg, stepped as one group (0: skip)")
    ap.add_argument("--batch128-streams", type=int, default=16,
                    help="instances of the stream-per-instance comparison of that leg (0: skip)")
    ap.add_argument("--batch128-warmup", type=int, default=1)
    ap.add_argument("--batch128-steps", type=int, default=2)
    ap.add_argument("--batch", type=int, default=0,
                    help="also time the first N instances of the batched 2TDVP draws "
                         "(default 0: skip)")
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--batch-warmup", type=int, default=10)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--serve", type=float, default=50.0,
                    help="Poisson arrival rate of the served window (requests/s; 0: skip)")
    ap.add_argument("--serve-window", type=float, default=1.0)
    ap.add_argument("--serve-max-steps", type=int, default=10)
    ap.add_argument("--serve-chi-cap", type=int, default=0, help="cap on chi_max of served requests (0: none)")
    ap.add_argument("--serve-big-fraction", type=float, default=0.1)
    ap.add_argument("--serve-method", choices=["group", "streams"], default="group")
    ap.add_argument("--tdvp1", type=int, default=0,
                    help="also time N steps of the 1TDVP chi=256 workload (101 cells), both precisions")
    ap.add_argument("--exact-apply", type=int, default=0,
                    help="also time the exact path's 2^N H application at N cells")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched workload (--batch N instances) and print its object "
                         "(the command the one-launch Krylov profile is taken with)")
    ap.add_argument("--group128-only", action="store_true",
                    help="prepare the chi=128 state, run only the grouped batched leg (--batch128 instances) "
                         "and print its object (the command the grouped-step profile is taken with)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the H_eff micro-benchmark and print its roofline object "
                         "(the command the kernel profiles are taken with)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU (one MI355X): the 2TDVP sweep of one instance does not "
              "shard, and batched instances would be placed one per GPU; run with --gpus 1",
              file=sys.stderr)
        return 2
    if a.batch_only and not a.batch:
        a.batch = 32
    if a.batch or a.serve or a.batch128:
        # one hardware queue per worker stream (HIP's default is 4); must be set
        # before the HIP runtime starts
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(16, max(4, a.batch_threads)))
    import torch
    import qca_native as qn
    if not torch.cuda.is_available():
        print("no HIP device visible", file=sys.stderr)
        return 2
    try:
        return run(torch, qn, a)
    finally:
        qn.shutdown()


def run(torch, qn, a) -> int:
    if a.roofline_only:
        ms32, fl = heff2_roofline(torch, 1)
        ms64, _ = heff2_roofline(torch, 0)
        # a plain device copy of known size (256 MiB read + 256 MiB written):
        # the scale for HBM byte counters read in the same run
        src = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"roofline": {"mixed_ms_per_launch": ms32, "fp64_ms_per_launch": ms64,
                                       "flops_per_launch": fl,
                                       "mixed_tflops": fl / (ms32 * 1e-3) / 1e12,
                                       "fp64_tflops": fl / (ms64 * 1e-3) / 1e12}}), flush=True)
        return 0
    if a.batch_only:
        print(json.dumps({"batched_draws": batch_bench(torch, a)}), flush=True)
        return 0
    if a.group128_only:
        e = qn.Engine(mpo().W, initial_state().A, algorithm=2, precision=0, chi_max=CHI, dt=DT, eps=EPS,
                      krylov_tol=1e-14)
        e.step(PREP_STEPS)
        prepared = e.tensors()
        e.close()
        a.batch128_streams = 0
        print(json.dumps({"batched_chi128": batch128_bench(torch, a, prepared, float("nan"))}), flush=True)
        return 0
    Ws = mpo().W
    psi = initial_state()
    # Preparation: PREP_STEPS plain FP64 steps grow the bonds to chi_max; the
    # mixed engine then starts from that same state.
    engines = {"fp64": qn.Engine(Ws, psi.A, algorithm=2, precision=0, chi_max=CHI, dt=DT, eps=EPS,
                                 krylov_tol=1e-14)}
    log(f"engine built; preparation {PREP_STEPS} FP64 steps")
    for w in range(PREP_STEPS):
        engines["fp64"].step(1)
        if w % 5 == 4 or w == PREP_STEPS - 1:
            bd = engines["fp64"].bond_dims()
            log(f"prep {w + 1}: max bond {bd.max()}, bonds at chi_max {int((bd == CHI).sum())}")
    prepared = engines["fp64"].tensors()
    engines = {"mixed": qn.Engine(Ws, prepared, algorithm=2, precision=1, chi_max=CHI,
                                  dt=DT, eps=EPS, krylov_tol=a.mixed_krylov_tol),
               "fp64": engines["fp64"]}
    for k in range(a.warmup):
        for name, e in engines.items():
            e.step(1)
        log(f"warmup {k + 1}/{a.warmup}")
    bonds_start = {k: e.bond_dims() for k, e in engines.items()}
    maxdiff = 0.0
    times = {"mixed": 0.0, "fp64": 0.0}
    gpu_ms = {"mixed": 0.0, "fp64": 0.0}
    flops = {}
    for e in engines.values():
        e.reset_stats()
    for k in range(a.steps):
        for name, e in engines.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e.step(1)
            torch.cuda.synchronize()
            times[name] += time.perf_counter() - t0
        d = np.abs(engines["mixed"].rho()[:, 1, 1].real - engines["fp64"].rho()[:, 1, 1].real).max()
        maxdiff = max(maxdiff, d)
        log(f"timed step {k + 1}/{a.steps}: mixed {times['mixed']:.2f}s fp64 {times['fp64']:.2f}s |dn| {d:.2e}")
    for name, e in engines.items():
        gpu_ms[name] = e.gpu_ms()
        flops[name] = e.stats()
    su = a.steps * 2 * (N - 1)
    rate = {k: su / v for k, v in times.items()}
    bonds_end = {k: e.bond_dims() for k, e in engines.items()}
    log("roofline micro-benchmark of the 2-site H_eff application")
    ms32, fl = heff2_roofline(torch, 1)
    ms64, _ = heff2_roofline(torch, 0)
    ach32 = fl / (ms32 * 1e-3) / 1e12
    ach64 = fl / (ms64 * 1e-3) / 1e12
    fp64_tensors = engines["fp64"].tensors()
    for e in engines.values():
        e.close()
    batched128 = batch128_bench(torch, a, prepared, rate["fp64"]) if a.batch128 else None
    tdvp1 = tdvp1_bench(torch, a) if a.tdvp1 else None
    exact = exact_apply_bench(torch, a.exact_apply) if a.exact_apply else None
    batched = batch_bench(torch, a) if a.batch else None
    served = serve_bench(a) if a.serve else None
    cpu = None
    if not a.no_cpu_baseline:
        log(f"cpu baseline ({a.cpu_seconds:.0f} s budget)")
        cpu = cpu_baseline(fp64_tensors, a.cpu_seconds)
    out = {
        "metric": "2TDVP site-updates/s at \u03c7=128, mixed vs FP64 @max|\u0394\u27e8n\u27e9|\u22641e-6; "
                  "p50/p99 latency",
        "value": rate["mixed"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * times["mixed"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 contractions + f64 Krylov/SVD",
        "data": "synthetic (three adjacent centre cells alive, rule Hamiltonian MPO)",
        "config": {"workload": "2TDVP single instance: 41 cells, distance 2, interval [2,4], three centre "
                               "cells alive, dt 0.1, chi_max 128, eps 1e-10; timed on the state after "
                               f"{PREP_STEPS} preparation FP64 steps plus the warmup steps",
                   "ncells": N, "chi_max": CHI, "mpo_bond": int(Ws[N // 2].shape[2]),
                   "prep_steps": PREP_STEPS,
                   "bonds_at_chi_max": {k: {"start": int((bonds_start[k] == CHI).sum()),
                                            "end": int((bonds_end[k] == CHI).sum())} for k in bonds_start},
                   "bonds_total": N - 1},
        "fp64_value": rate["fp64"],
        "mixed_over_fp64": rate["mixed"] / rate["fp64"],
        "max_abs_dn_mixed_vs_fp64": maxdiff,
        "gpu_ms_per_step": {k: v / a.steps for k, v in gpu_ms.items()},
        "matvecs_per_step": {k: v["matvecs"] / a.steps for k, v in flops.items()},
        "phase_gpu_ms_per_step": {k: {"svd": v["svd_ms"] / a.steps, "krylov": v["krylov_ms"] / a.steps}
                                  for k, v in flops.items()},
        "svd_sweeps_per_call": {k: v["svd_sweeps"] / max(1, v["svd_calls"]) for k, v in flops.items()},
        "contraction_tflops_per_s": {k: flops[k]["flops"] / times[k] / 1e12 for k in flops},
        "roofline": {"bound": "mfma", "kernel": "2-site H_eff application (mixed, complex64 on FP32 MFMA)",
                     "achieved": ach32, "peak": 157.3, "unit": "TFLOP/s", "frac": ach32 / 157.3,
                     "traffic": HEFF2_TRAFFIC_BYTES["mixed"], "traffic_unit": "B per launch",
                     "traffic_source": "profile-derived (not measured in this run): rocprofv3 --pmc "
                                       "FETCH_SIZE / WRITE_SIZE of bench.py --roofline-only, scaled by a "
                                       f"256 MiB copy ({HEFF2_TRAFFIC_SOURCE})",
                     "ms_per_launch": ms32, "flops_per_launch": fl,
                     "fp64": {"achieved": ach64, "peak": 78.6, "frac": ach64 / 78.6, "ms_per_launch": ms64,
                              "traffic": HEFF2_TRAFFIC_BYTES["fp64"]}},
        "cpu_baseline": cpu,
        "batched_chi128": batched128,
        "served": served,
        "tdvp1_chi256": tdvp1,
        "exact_apply": exact,
        "batched_draws": batched,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
