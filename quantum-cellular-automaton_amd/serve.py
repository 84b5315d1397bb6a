"""Serving many independent TDVP simulations on one GPU: Poisson arrivals,
admission by HBM footprint, continuous batching, per-request latency.

A request is one simulation (an `workloads.Instance` plus a step count); the
server returns <n_i> every `report_every` steps, as the reference driver
measures every plotting interval (quantum_game.py:82-119), and the request's
latency is the time from its arrival to its last returned <n_i>.

Two schedulers share the admission rule.  GroupServer (the fast one) sorts
admitted requests into size lanes and advances each lane's requests together,
one grouped lockstep time step per tick (qca_group_step), lanes side by side
on their own threads and streams.  Server keeps admitted requests in a
round-robin queue; `workers` host threads repeatedly take the front request,
advance it by one reporting chunk on its own engine (its own HIP stream),
measure, and requeue it until done.  In both, new arrivals join while others
run (continuous batching).  A request is admitted when its device footprint
fits in the remaining budget; otherwise it waits in arrival order.  With the GPU engine factory the footprint is the exact number of
bytes the engine will allocate (qca_engine_footprint, the constructor's own
sizing) and the budget is the device's free memory at start-up, less a
reserve.  The engine factory is injectable, so the scheduler is tested on the
host with a stand-in engine.
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

import workloads


def footprint_bytes(inst: workloads.Instance, precision: int = 1) -> int:
    """Closed-form upper bound of an engine's device memory, for factories
    that do not report their own footprint: MPS (c128), environments (c64
    mixed / c128 plain), Krylov basis (64 vectors of the largest 2-site
    block) and per-engine workspaces."""
    n, chi = inst.ncells, inst.chi_max
    dims = [min(2 ** min(i, 40), 2 ** min(n - i, 40), chi) for i in range(n + 1)]
    D = 2 + 2 * inst.distance + 4 * inst.distance * inst.distance  # generous MPO bond bound
    env_elem = 8 if precision == 1 else 16
    mps = sum(2 * dims[i] * dims[i + 1] for i in range(n)) * 16
    env = 2 * sum(dims[i] ** 2 * D for i in range(n + 1)) * env_elem
    nmax = 4 * chi * chi
    krylov = 66 * nmax * 16
    work = 8 * nmax * D * 16 + (2 << 20)
    return int(mps + env + krylov + work)


@dataclass
class Request:
    rid: int
    inst: workloads.Instance
    steps: int
    arrival: float
    done_steps: int = 0
    engine: object = None
    reports: list = field(default_factory=list)
    finished: float | None = None
    footprint: int = 0


# Device memory kept free for the runtime, the primitives' workspace and
# allocation granularity when the budget comes from the free memory.
RESERVE_BYTES = 4 << 30


class Server:
    def __init__(self, make_engine, workers: int = 8, budget_bytes: int | None = None,
                 report_every: int = 10, precision: int = 1, device: int = 0):
        """``make_engine(inst)`` builds an engine; when it has a
        ``footprint(inst)`` attribute (the GPU factory), that is the admission
        size, else the closed-form bound ``footprint_bytes``.  ``budget_bytes``
        defaults to the device's free memory less RESERVE_BYTES."""
        self.make_engine = make_engine
        self.workers = workers
        if budget_bytes is None:
            import qca_native as qn
            free, _ = qn.device_memory(device)
            budget_bytes = max(0, free - RESERVE_BYTES)
        self.budget = budget_bytes
        self.report_every = report_every
        self.precision = precision
        self._footprint = getattr(make_engine, "footprint", None)
        self._cv = threading.Condition()
        self._pending = collections.deque()   # arrived, not admitted (arrival order)
        self._ready = collections.deque()     # admitted, waiting for a worker
        self._in_use = 0
        self.peak_in_use = 0
        self._closed = False
        self._error = None
        self.finished: list = []
        self.site_updates = 0

    # ------------------------------------------------------------ intake
    def submit(self, req: Request) -> None:
        req.footprint = (self._footprint(req.inst) if self._footprint is not None
                         else footprint_bytes(req.inst, self.precision))
        if req.footprint > self.budget:
            raise ValueError(f"request {req.rid} needs {req.footprint} B, above the budget")
        with self._cv:
            self._pending.append(req)
            self._admit()
            self._cv.notify_all()

    def close(self) -> None:
        """No more arrivals; workers exit when the queues drain."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()

    def _admit(self) -> None:
        while self._pending and self._in_use + self._pending[0].footprint <= self.budget:
            req = self._pending.popleft()
            self._in_use += req.footprint
            self.peak_in_use = max(self.peak_in_use, self._in_use)
            self._ready.append(req)

    # ------------------------------------------------------------ workers
    def _worker(self) -> None:
        while True:
            with self._cv:
                while not self._ready and not (self._closed and not self._pending) and self._error is None:
                    self._cv.wait()
                if self._error is not None or not self._ready:
                    return
                req = self._ready.popleft()
            try:
                if req.engine is None:
                    req.engine = self.make_engine(req.inst)
                n = min(self.report_every, req.steps - req.done_steps)
                req.engine.step(n)
                req.done_steps += n
                req.reports.append(np.asarray(req.engine.rho()[:, 1, 1].real))
                finished = req.done_steps >= req.steps
                if finished:
                    req.finished = time.perf_counter()
                    req.engine.close()
                    req.engine = None
            except Exception as exc:  # stop the server, keep the first error
                with self._cv:
                    if self._error is None:
                        self._error = exc
                    self._cv.notify_all()
                return
            with self._cv:
                self.site_updates += n * req.inst.site_updates_per_step
                if finished:
                    self.finished.append(req)
                    self._in_use -= req.footprint
                    self._admit()
                else:
                    self._ready.append(req)
                self._cv.notify_all()

    def run(self, arrivals: list) -> dict:
        """Feeds (arrival_offset_s, Request) pairs at their times, serves them
        until all finish, and returns latency and throughput statistics."""
        threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(self.workers)]
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for offset, req in arrivals:
            delay = t0 + offset - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            req.arrival = time.perf_counter()
            self.submit(req)
        self.close()
        for t in threads:
            t.join()
        if self._error is not None:
            raise self._error
        wall = time.perf_counter() - t0
        lat = np.array([r.finished - r.arrival for r in self.finished])
        return {"requests": len(self.finished), "wall_s": wall,
                "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
                "site_updates": self.site_updates,
                "site_updates_per_s": self.site_updates / wall if wall > 0 else 0.0}


def size_lane(inst: workloads.Instance) -> tuple:
    """Lane of a request in GroupServer: requests whose time steps cost about
    the same share a lockstep group.  A grouped step walks the longest chain
    of its group (2(N-1) sweep positions) at the widest bond, so chains up to
    38 cells, longer chains, and bond caps above 128 go to separate lanes."""
    return (inst.chi_max > 128, inst.ncells > 38)


class GroupServer:
    """Continuous batching through grouped lockstep steps, one group per lane.

    Admitted requests are sorted into lanes by ``lane_of(inst)`` (default
    size_lane).  Each lane runs its own tick loop on its own host thread:
    every tick, the lane's newly admitted requests join it, and all of its
    requests advance one time step in ONE grouped step (qca_group_step: each
    kernel of a sweep position runs once for all of them; a group has its own
    HIP stream, so lanes run side by side on the GPU).  A request reports
    <n_i> every ``report_every`` steps and after its last one, then leaves;
    a lane's group is rebuilt whenever its membership changes.  Requests are
    admitted in arrival order while their footprint fits the budget.
    ``group_factory(engines)`` builds a group (default qca_native.Group), so
    the scheduler is tested on the host with stand-ins."""

    def __init__(self, make_engine, budget_bytes: int | None = None, report_every: int = 10,
                 precision: int = 1, device: int = 0, group_factory=None, lane_of=size_lane):
        self.make_engine = make_engine
        if budget_bytes is None:
            import qca_native as qn
            free, _ = qn.device_memory(device)
            budget_bytes = max(0, free - RESERVE_BYTES)
        self.budget = budget_bytes
        self.report_every = report_every
        self.precision = precision
        self._footprint = getattr(make_engine, "footprint", None)
        if group_factory is None:
            import qca_native as qn
            group_factory = qn.Group
        self.group_factory = group_factory
        self.lane_of = lane_of
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_active = 0
        self.group_builds = 0
        self.ticks = 0
        self.lane_ticks: dict = {}
        self.finished: list = []
        self.site_updates = 0

    def _size(self, req) -> int:
        return (self._footprint(req.inst) if self._footprint is not None
                else footprint_bytes(req.inst, self.precision))

    def _lane(self, key, inbox: collections.deque, cond: threading.Condition, state: dict):
        """Tick loop of one lane (runs on its own thread)."""
        active: list = []
        group, members = None, ()
        try:
            while True:
                with cond:
                    while not inbox and not active and not state["stop"]:
                        cond.wait(0.05)
                    if state["error"] is not None or (not inbox and not active and state["stop"]):
                        return
                    active.extend(inbox)
                    inbox.clear()
                key_m = tuple(r.rid for r in active)
                if key_m != members:
                    if group is not None:
                        group.close()
                    group = self.group_factory([r.engine for r in active])
                    members = key_m
                    with cond:
                        self.group_builds += 1
                group.step(1)
                done, still = [], []
                for r in active:
                    r.done_steps += 1
                    last = r.done_steps >= r.steps
                    if last or r.done_steps % self.report_every == 0:
                        r.reports.append(np.asarray(r.engine.rho()[:, 1, 1].real))
                    (done if last else still).append(r)
                if done:
                    group.close()
                    group, members = None, ()
                    for r in done:
                        r.engine.close()
                        r.engine = None
                        r.finished = time.perf_counter()
                with cond:
                    self.ticks += 1
                    self.lane_ticks[key] = self.lane_ticks.get(key, 0) + 1
                    self.site_updates += sum(r.inst.site_updates_per_step for r in active)
                    for r in done:
                        self.in_use -= r.footprint
                        self.finished.append(r)
                    state["active"] -= len(done)
                    cond.notify_all()
                active = still
        except BaseException as exc:  # handed to the scheduler, which stops every lane
            with cond:
                state["error"] = state["error"] or exc
                state["stop"] = True
                cond.notify_all()
        finally:
            if group is not None:
                group.close()

    def run(self, arrivals: list) -> dict:
        arrived = collections.deque()
        feed_lock = threading.Lock()
        t0 = time.perf_counter()
        feeding = {"done": False}

        def feed():
            for offset, req in arrivals:
                delay = t0 + offset - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                req.arrival = time.perf_counter()
                with feed_lock:
                    arrived.append(req)
            feeding["done"] = True

        feeder = threading.Thread(target=feed, daemon=True)
        feeder.start()
        cond = threading.Condition()
        state = {"stop": False, "error": None, "active": 0}
        lanes: dict = {}
        pending = collections.deque()
        try:
            while True:
                with feed_lock:
                    while arrived:
                        req = arrived.popleft()
                        req.footprint = self._size(req)
                        if req.footprint > self.budget:
                            raise ValueError(f"request {req.rid} needs {req.footprint} B, above the budget")
                        pending.append(req)
                    fed = feeding["done"] and not arrived
                admitted = []
                with cond:
                    if state["error"] is not None:
                        break
                    while pending and self.in_use + pending[0].footprint <= self.budget:
                        req = pending.popleft()
                        self.in_use += req.footprint
                        self.peak_in_use = max(self.peak_in_use, self.in_use)
                        state["active"] += 1
                        admitted.append(req)
                    self.peak_active = max(self.peak_active, state["active"])
                for req in admitted:
                    req.engine = self.make_engine(req.inst)
                    key = self.lane_of(req.inst)
                    if key not in lanes:
                        inbox = collections.deque()
                        th = threading.Thread(target=self._lane, args=(key, inbox, cond, state), daemon=True)
                        lanes[key] = (inbox, th)
                        th.start()
                    with cond:
                        lanes[key][0].append(req)
                        cond.notify_all()
                with cond:
                    if fed and not pending and state["active"] == 0:
                        break
                    cond.wait(0.002)
        finally:
            with cond:
                state["stop"] = True
                cond.notify_all()
            for _, th in lanes.values():
                th.join()
            feeder.join()
        if state["error"] is not None:
            raise state["error"]
        wall = time.perf_counter() - t0
        lat = np.array([r.finished - r.arrival for r in self.finished])
        return {"requests": len(self.finished), "wall_s": wall,
                "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
                "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
                "site_updates": self.site_updates,
                "site_updates_per_s": self.site_updates / wall if wall > 0 else 0.0,
                "ticks": self.ticks, "lanes": len(lanes), "group_builds": self.group_builds,
                "peak_active": self.peak_active, "peak_footprint_bytes": self.peak_in_use,
                "budget_bytes": self.budget,
                "method": "grouped lockstep ticks (qca_group_step), one group per size lane, lanes on "
                          "their own threads and streams"}


def poisson_arrivals(rate: float, window_s: float, seed: int = 5, big_fraction: float = 0.1,
                     max_steps: int | None = None) -> list:
    """(offset_s, Request) pairs: exponential inter-arrival times of mean
    1/rate within [0, window_s); each request is workloads.serving_request
    (steps optionally capped at max_steps)."""
    rng = np.random.default_rng(seed)
    out, t, rid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= window_s:
            return out
        inst, steps = workloads.serving_request(rng, big_fraction)
        if max_steps is not None:
            steps = min(steps, max_steps)
        out.append((t, Request(rid, inst, steps, 0.0)))
        rid += 1


class GpuEngineFactory:
    """Engine factory for the GPU path: one qca_engine per request.  Requests
    share the GPU, so Krylov blocks up to ``fused_krylov_max_n`` elements use
    the one-launch exponential (throughput over single-request latency).
    ``footprint(inst)`` is the exact device size of the engine ``self(inst)``
    builds (qca_engine_footprint)."""

    def __init__(self, precision: int = 1, chi_cap: int | None = None, fused_krylov_max_n: int = 4096):
        self.precision = precision
        self.chi_cap = chi_cap
        self.fused_krylov_max_n = fused_krylov_max_n

    def _chi(self, inst: workloads.Instance) -> int:
        return inst.chi_max if self.chi_cap is None else min(inst.chi_max, self.chi_cap)

    def footprint(self, inst: workloads.Instance) -> int:
        import qca_native as qn
        return qn.engine_footprint(workloads.mpo_tensors(inst), [1] * (inst.ncells + 1), algorithm=2,
                                   precision=self.precision, chi_max=self._chi(inst))

    def __call__(self, inst: workloads.Instance):
        import qca_native as qn
        return qn.Engine(workloads.mpo_tensors(inst), workloads.initial_tensors(inst), algorithm=2,
                         precision=self.precision, chi_max=self._chi(inst), dt=inst.step_size,
                         eps=inst.svd_epsilon, krylov_tol=1e-10 if self.precision == 1 else 1e-14,
                         fused_krylov_max_n=self.fused_krylov_max_n)


def gpu_engine_factory(precision: int = 1, chi_cap: int | None = None,
                       fused_krylov_max_n: int = 4096) -> GpuEngineFactory:
    """The GPU engine factory (see GpuEngineFactory)."""
    return GpuEngineFactory(precision, chi_cap, fused_krylov_max_n)
