"""ctypes binding of libqca.so (include/qca.h).

The library is required: every call raises ``QcaError`` when it is missing or
when a call fails; there is no host fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libqca.so")

c_int, c_long, c_double, c_void_p = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
P = ctypes.POINTER


class QcaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libqca error {code}: {msg}")
        self.code = code


class EngineParams(ctypes.Structure):
    _fields_ = [("device", c_int), ("ncells", c_int), ("algorithm", c_int), ("precision", c_int),
                ("chi_max", c_int), ("reserved", c_int), ("dt", c_double), ("eps", c_double),
                ("krylov_tol", c_double)]


class Stats(ctypes.Structure):
    _fields_ = [("matvecs", c_long), ("matvec_2site", c_long), ("matvec_1site", c_long),
                ("matvec_bond", c_long), ("site_updates", c_long), ("svd_sweeps", c_long),
                ("lanczos_limit_hits", c_long), ("svd_calls", c_long), ("flops", c_double),
                ("svd_ms", c_double), ("krylov_ms", c_double)]


class GroupStats(ctypes.Structure):
    _fields_ = [("steps", c_long), ("krylov_iterations", c_long), ("host_syncs", c_long),
                ("krylov_ms", c_double), ("svd_ms", c_double), ("total_ms", c_double)]


# name -> (restype, argtypes); mirrors include/qca.h
SIGNATURES = {
    "qca_version": (c_int, []),
    "qca_last_error": (ctypes.c_char_p, []),
    "qca_device_count": (c_int, [P(c_int)]),
    "qca_engine_create": (c_int, [P(EngineParams), P(c_void_p), P(c_int), P(c_int), P(c_void_p),
                                  P(c_int), P(c_void_p)]),
    "qca_engine_destroy": (c_int, [c_void_p]),
    "qca_engine_footprint": (c_int, [P(EngineParams), P(c_void_p), P(c_int), P(c_int), P(c_int),
                                     P(ctypes.c_size_t)]),
    "qca_engine_device_bytes": (c_int, [c_void_p, P(ctypes.c_size_t)]),
    "qca_device_memory": (c_int, [c_int, P(ctypes.c_size_t), P(ctypes.c_size_t)]),
    "qca_engine_step": (c_int, [c_void_p, c_int]),
    "qca_engine_measure": (c_int, [c_void_p, c_void_p]),
    "qca_engine_bond_dims": (c_int, [c_void_p, P(c_int)]),
    "qca_engine_get_site": (c_int, [c_void_p, c_int, c_void_p]),
    "qca_engine_stats": (c_int, [c_void_p, P(Stats)]),
    "qca_engine_reset_stats": (c_int, [c_void_p]),
    "qca_batch_step": (c_int, [P(c_void_p), c_int, c_int, c_int, P(c_int)]),
    "qca_engine_observables": (c_int, [c_void_p, P(c_double), P(c_double), P(c_double), P(c_double)]),
    "qca_engine_gpu_ms": (c_int, [c_void_p, P(c_double)]),
    "qca_engine_set_fused_krylov_limit": (c_int, [c_void_p, c_long]),
    "qca_group_create": (c_int, [P(c_void_p), c_int, P(c_void_p)]),
    "qca_group_step": (c_int, [c_void_p, c_int]),
    "qca_group_get_stats": (c_int, [c_void_p, P(GroupStats)]),
    "qca_group_reset_stats": (c_int, [c_void_p]),
    "qca_group_destroy": (c_int, [c_void_p]),
    "qca_heff_apply": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_int, c_int, c_int, c_int, c_void_p]),
    "qca_heff_apply_noalloc": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_int, c_int, c_int, c_int, c_void_p]),
    "qca_env_update": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                               c_int, c_int, c_void_p]),
    "qca_krylov_expm": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                c_int, c_int, c_double, c_double, P(c_int), c_void_p]),
    "qca_svd_split": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_double, c_int,
                              P(c_int), c_void_p, c_void_p]),
    "qca_qr_site": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "qca_exact_apply_h": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p]),
    "qca_exact_expm": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double, P(c_int),
                               c_void_p]),
    "qca_exact_rho": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "qca_release_workspace": (c_int, []),
}

_lib = None
# Engines still open; closed explicitly at interpreter exit (below).
_live_engines: "weakref.WeakSet" = weakref.WeakSet()
_live_groups: "weakref.WeakSet" = weakref.WeakSet()


def shutdown() -> None:
    """Close every open engine and release the primitives' device workspace
    (qca_engine_destroy, qca_release_workspace).  Registered with ``atexit``,
    so it runs while the HIP runtime is still up: device memory, streams and
    events are never left to ``__del__`` during interpreter finalisation or to
    process-exit teardown."""
    for g in list(_live_groups):
        try:
            g.close()
        except Exception:
            pass
    for e in list(_live_engines):
        try:
            e.close()
        except Exception:
            pass
    if _lib is not None:
        _lib.qca_release_workspace()


atexit.register(shutdown)


def lib():
    """Load libqca.so once; raises QcaError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QcaError(-1, f"{LIB_PATH} not found; build it with "
                               "quantum-cellular-automaton_amd/build_lib.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise QcaError(rc, lib().qca_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def ptr(a) -> int:
    """Address of a numpy array (host) or torch tensor (device)."""
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    assert a.is_contiguous()
    return a.data_ptr()


def stream_ptr(stream=None) -> int | None:
    if stream is None:
        return None
    return stream.cuda_stream


def device_count() -> int:
    """HIP devices visible to the library (qca_device_count)."""
    n = c_int(0)
    check(lib().qca_device_count(ctypes.byref(n)))
    return n.value


def device_memory(device: int = 0) -> tuple:
    """(free, total) bytes of a HIP device (qca_device_memory)."""
    f, t = ctypes.c_size_t(0), ctypes.c_size_t(0)
    check(lib().qca_device_memory(device, ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def engine_footprint(Ws, chis, *, algorithm: int, precision: int, chi_max: int) -> int:
    """Device bytes an Engine of these MPO tensors, initial bond dimensions
    (N+1 values) and parameters would allocate (qca_engine_footprint; host
    only, no device needed)."""
    n = len(Ws)
    Wc = [np.ascontiguousarray(W, dtype=np.complex128) for W in Ws]
    if len(chis) != n + 1:
        raise ValueError("N+1 bond dimensions are required")
    prm = EngineParams(device=0, ncells=n, algorithm=algorithm, precision=precision, chi_max=chi_max,
                       reserved=0, dt=0.1, eps=1e-10, krylov_tol=0.0)
    Wp = (c_void_p * n)(*[W.ctypes.data for W in Wc])
    dl = (c_int * n)(*[W.shape[2] for W in Wc])
    dr = (c_int * n)(*[W.shape[3] for W in Wc])
    ch = (c_int * (n + 1))(*[int(c) for c in chis])
    out = ctypes.c_size_t(0)
    check(lib().qca_engine_footprint(ctypes.byref(prm), Wp, dl, dr, ch, ctypes.byref(out)))
    return out.value


class Engine:
    """Owning handle of a device-resident TDVP instance (qca_engine_*)."""

    def __init__(self, Ws, As, *, algorithm: int, precision: int, chi_max: int, dt: float,
                 eps: float, krylov_tol: float = 0.0, device: int = 0,
                 fused_krylov_max_n: int | None = None):
        n = len(As)
        if len(Ws) != n:
            raise ValueError("one MPO tensor per site is required")
        self.n = n
        self._W = [np.ascontiguousarray(W, dtype=np.complex128) for W in Ws]
        self._A = [np.ascontiguousarray(A, dtype=np.complex128) for A in As]
        chis = [A.shape[1] for A in self._A] + [self._A[-1].shape[2]]
        for i in range(n - 1):
            if self._A[i].shape[2] != self._A[i + 1].shape[1]:
                raise ValueError(f"bond mismatch between sites {i} and {i + 1}")
        prm = EngineParams(device=device, ncells=n, algorithm=algorithm, precision=precision,
                           chi_max=chi_max, reserved=0, dt=dt, eps=eps, krylov_tol=krylov_tol)
        Wp = (c_void_p * n)(*[W.ctypes.data for W in self._W])
        Ap = (c_void_p * n)(*[A.ctypes.data for A in self._A])
        dl = (c_int * n)(*[W.shape[2] for W in self._W])
        dr = (c_int * n)(*[W.shape[3] for W in self._W])
        ch = (c_int * (n + 1))(*chis)
        h = c_void_p()
        check(lib().qca_engine_create(ctypes.byref(prm), Wp, dl, dr, Ap, ch, ctypes.byref(h)))
        self._h = h
        _live_engines.add(self)
        if fused_krylov_max_n is not None:
            self.set_fused_krylov_limit(fused_krylov_max_n)

    def device_bytes(self) -> int:
        """Device bytes held by the engine (qca_engine_device_bytes)."""
        out = ctypes.c_size_t(0)
        check(lib().qca_engine_device_bytes(self._h, ctypes.byref(out)))
        return out.value

    def set_fused_krylov_limit(self, max_n: int) -> None:
        """Krylov blocks of at most ``max_n`` elements use the one-launch
        exponential (qca_engine_set_fused_krylov_limit)."""
        check(lib().qca_engine_set_fused_krylov_limit(self._h, int(max_n)))

    def close(self) -> None:
        """Release the device instance (qca_engine_destroy)."""
        if getattr(self, "_h", None):
            check(lib().qca_engine_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps: int = 1) -> None:
        """nsteps TDVP time steps on the device (qca_engine_step; reference tdvp.py:50-63 per step)."""
        check(lib().qca_engine_step(self._h, nsteps))

    def rho(self) -> np.ndarray:
        """Single-site density matrices (N, 2, 2) (qca_engine_measure)."""
        out = np.zeros((self.n, 2, 2), dtype=np.complex128)
        check(lib().qca_engine_measure(self._h, out.ctypes.data))
        return out

    def observables(self, population=None, d_population=None, single_site_entropy=None,
                    bond_dims=None) -> tuple:
        """Device-computed rows of the reference's measure(): copies them into
        the given float64 arrays (row views allowed) and returns the four rows."""
        bufs = [np.zeros(self.n), np.zeros(self.n), np.zeros(self.n), np.zeros(self.n + 1)]
        check(lib().qca_engine_observables(self._h, *[b.ctypes.data_as(P(c_double)) for b in bufs]))
        for dst, b in zip((population, d_population, single_site_entropy, bond_dims), bufs):
            if dst is not None:
                dst[...] = b
        return tuple(bufs)

    def bond_dims(self) -> np.ndarray:
        """The N+1 current bond dimensions (qca_engine_bond_dims)."""
        out = (c_int * (self.n + 1))()
        check(lib().qca_engine_bond_dims(self._h, out))
        return np.array(list(out), dtype=int)

    def site(self, i: int) -> np.ndarray:
        """Copy of site tensor i, shape (2, D[i], D[i+1]) (qca_engine_get_site)."""
        bd = self.bond_dims()
        out = np.zeros((2, bd[i], bd[i + 1]), dtype=np.complex128)
        check(lib().qca_engine_get_site(self._h, i, out.ctypes.data))
        return out

    def tensors(self) -> list:
        """Copies of every site tensor (the MPS the engine evolves)."""
        return [self.site(i) for i in range(self.n)]

    def stats(self) -> dict:
        """Running counters and device times since creation or reset (qca_engine_stats)."""
        s = Stats()
        check(lib().qca_engine_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def reset_stats(self) -> None:
        """Zero the counters (qca_engine_reset_stats)."""
        check(lib().qca_engine_reset_stats(self._h))

    def gpu_ms(self) -> float:
        """Device time of the engine's stream since creation (qca_engine_gpu_ms)."""
        v = c_double(0)
        check(lib().qca_engine_gpu_ms(self._h, ctypes.byref(v)))
        return v.value


def batch_step(engines, nsteps: int = 1, nthreads: int = 8) -> None:
    """Advance every engine by ``nsteps`` time steps with ``nthreads`` native
    worker threads, one stream per engine (qca_batch_step)."""
    engines = list(engines)
    if not engines:
        return
    hs = (c_void_p * len(engines))(*[e._h for e in engines])
    failed = c_int(-1)
    rc = lib().qca_batch_step(hs, len(engines), nsteps, nthreads, ctypes.byref(failed))
    check(rc)


class Group:
    """Lockstep grouped stepping of several 2TDVP engines (qca_group_*): one
    launch per kernel per sweep position for all of them.  Keeps references
    to its engines so that they outlive it."""

    def __init__(self, engines):
        self.engines = list(engines)
        if not self.engines:
            raise ValueError("a group needs at least one engine")
        hs = (c_void_p * len(self.engines))(*[e._h for e in self.engines])
        h = c_void_p()
        check(lib().qca_group_create(hs, len(self.engines), ctypes.byref(h)))
        self._h = h
        _live_groups.add(self)

    def step(self, nsteps: int = 1) -> None:
        """nsteps time steps of every engine (qca_group_step)."""
        check(lib().qca_group_step(self._h, nsteps))

    def stats(self) -> dict:
        s = GroupStats()
        check(lib().qca_group_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in GroupStats._fields_}

    def reset_stats(self) -> None:
        check(lib().qca_group_reset_stats(self._h))

    def close(self) -> None:
        """Release the group (qca_group_destroy); the engines stay open."""
        if getattr(self, "_h", None):
            check(lib().qca_group_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
