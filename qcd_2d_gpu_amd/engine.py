"""Batched lattices and serving (ctypes front end of su2_engine_* and
su2_server_* in include/su2mc.h).

`Engine` holds many lattices in one device arena and advances all of them
with one launch per (direction, parity); `Server` runs an engine in a
background thread with continuous admission.  Both run entirely in
libsu2mc.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


class Su2Job(C.Structure):
    _fields_ = [
        ("Lt", C.c_int), ("Lx", C.c_int), ("Ly", C.c_int),
        ("beta", C.c_float), ("epsilon", C.c_float), ("hits", C.c_int),
        ("bc", C.c_int), ("algo", C.c_int), ("start", C.c_int), ("seed", C.c_uint64),
        ("n_therm", C.c_int), ("n_meas", C.c_int), ("n_cor", C.c_int), ("or_sweeps", C.c_int),
        ("normalize_every", C.c_int), ("fp64_twin", C.c_int),
    ]


class Su2EngineConfig(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("device", C.c_int), ("max_jobs", C.c_int), ("max_or", C.c_int),
        ("budget_bytes", C.c_size_t), ("use_graph", C.c_int), ("block", C.c_int), ("grid", C.c_int),
    ]


class Su2JobResult(C.Structure):
    _fields_ = [
        ("id", C.c_longlong), ("status", C.c_int),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("E_mean", C.c_double), ("E_err", C.c_double), ("P_mean", C.c_double), ("P_err", C.c_double),
        ("accept", C.c_double), ("link_updates", C.c_longlong),
        ("P_fp64_twin", C.c_double), ("P_fp64_twin_err", C.c_double), ("P_err_binned", C.c_double),
        ("P_diff", C.c_double), ("P_diff_sigma", C.c_double),
    ]


_P = C.c_void_p
ENGINE_PROTOTYPES = {
    "su2_job_bytes": (C.c_size_t, [C.POINTER(Su2Job), C.c_int]),
    "su2_engine_create": (C.c_int, [C.POINTER(Su2EngineConfig), C.POINTER(C.c_void_p)]),
    "su2_engine_destroy": (C.c_int, [_P]),
    "su2_engine_add": (C.c_int, [_P, C.POINTER(Su2Job), C.POINTER(C.c_int)]),
    "su2_engine_tick": (C.c_int, [_P, C.c_int]),
    "su2_engine_run": (C.c_int, [_P]),
    "su2_engine_sync": (C.c_int, [_P]),
    "su2_engine_job_status": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "su2_engine_read": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(C.c_double), _P, _P]),
    "su2_engine_remove": (C.c_int, [_P, C.c_int]),
    "su2_engine_autotune": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "su2_engine_set_shape": (C.c_int, [_P, C.c_int, C.c_int]),
    "su2_engine_memory": (C.c_int, [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "su2_server_create": (C.c_int, [C.POINTER(Su2EngineConfig), C.POINTER(C.c_void_p)]),
    "su2_server_destroy": (C.c_int, [_P]),
    "su2_server_submit": (C.c_int, [_P, C.POINTER(Su2Job), C.POINTER(C.c_longlong)]),
    "su2_server_poll": (C.c_int, [_P, C.POINTER(Su2JobResult), C.c_int, C.POINTER(C.c_int)]),
    "su2_server_drain": (C.c_int, [_P]),
    "su2_server_series": (C.c_int, [_P, C.c_longlong, _P, _P, C.c_int]),
    "su2_server_now": (C.c_double, [_P]),
}
L.PROTOTYPES.update(ENGINE_PROTOTYPES)


@dataclass
class Job:
    """One run request: the reference's per-run parameters (kernel.cu:10-46)."""
    Lt: int
    Lx: int
    Ly: int
    beta: float
    bc: int = L.BC_PERIODIC
    algo: int = L.ALGO_HEATBATH
    start: int = L.START_HOT
    seed: int = 1234
    n_therm: int = 50
    n_meas: int = 200
    n_cor: int = 1
    or_sweeps: int = 0
    epsilon: float = 0.1
    hits: int = 1
    normalize_every: int = 100
    fp64_twin: int = 0  # FP32-link server: also run with FP64 links and state the plaquette difference

    def c(self) -> Su2Job:
        return Su2Job(self.Lt, self.Lx, self.Ly, self.beta, self.epsilon, self.hits, self.bc, self.algo, self.start,
                      self.seed, self.n_therm, self.n_meas, self.n_cor, self.or_sweeps, self.normalize_every,
                      self.fp64_twin)

    @property
    def V(self) -> int:
        return self.Lt * self.Lx * self.Ly

    @property
    def link_updates(self) -> int:
        return 3 * self.V * (self.n_therm + self.n_meas * self.n_cor) * (1 + self.or_sweeps)


def _cfg(precision, max_jobs, max_or, budget_bytes, use_graph, block, grid, device):
    return Su2EngineConfig(precision, device, max_jobs, max_or, budget_bytes, int(use_graph), block, grid)


class Engine:
    """Many resident lattices advanced by shared launches."""

    def __init__(self, precision: int = L.PREC_F64, max_jobs: int = 256, max_or: int = 1,
                 budget_bytes: int = 1 << 30, use_graph: bool = True, block: int = 0, grid: int = 0,
                 device: int = 0):
        self.lib = L.load()
        self.h = C.c_void_p()
        self.cfg = _cfg(precision, max_jobs, max_or, budget_bytes, use_graph, block, grid, device)
        L.check(self.lib.su2_engine_create(C.byref(self.cfg), C.byref(self.h)), "engine_create")
        self.jobs: dict[int, Job] = {}

    def add(self, job: Job) -> int:
        slot = C.c_int(-1)
        cj = job.c()
        L.check(self.lib.su2_engine_add(self.h, C.byref(cj), C.byref(slot)), "engine_add")
        self.jobs[slot.value] = job
        return slot.value

    def tick(self, n: int = 1) -> None:
        L.check(self.lib.su2_engine_tick(self.h, n), "engine_tick")

    def run(self) -> None:
        L.check(self.lib.su2_engine_run(self.h), "engine_run")

    def sync(self) -> None:
        L.check(self.lib.su2_engine_sync(self.h), "engine_sync")

    def status(self, slot: int) -> tuple[int, int]:
        d, t = C.c_int(), C.c_int()
        L.check(self.lib.su2_engine_job_status(self.h, slot, C.byref(d), C.byref(t)), "job_status")
        return d.value, t.value

    def read(self, slot: int, field: bool = False, rng: bool = False):
        job = self.jobs[slot]
        E = np.zeros(job.n_meas)
        P = np.zeros(job.n_meas)
        acc = C.c_double()
        phi = np.zeros(12 * job.V) if field else None
        st = np.zeros(6 * (job.V // 2), dtype=np.uint32) if rng else None
        L.check(self.lib.su2_engine_read(self.h, slot, E.ctypes.data, P.ctypes.data, C.byref(acc),
                                         phi.ctypes.data if field else None, st.ctypes.data if rng else None),
                "engine_read")
        out = {"E": E, "P": P, "accept": acc.value}
        if field:
            out["field"] = phi
        if rng:
            out["rng"] = st.reshape(6, job.V // 2)
        return out

    def remove(self, slot: int) -> None:
        L.check(self.lib.su2_engine_remove(self.h, slot), "engine_remove")
        self.jobs.pop(slot, None)

    def autotune(self) -> tuple[int, int]:
        b, g = C.c_int(), C.c_int()
        L.check(self.lib.su2_engine_autotune(self.h, C.byref(b), C.byref(g)), "engine_autotune")
        return b.value, g.value

    def set_shape(self, block: int, grid: int = 0) -> None:
        """Sweep launches of `block` threads (64..256), grid cap `grid` (0 keeps it)."""
        L.check(self.lib.su2_engine_set_shape(self.h, block, grid), "engine_set_shape")

    def memory(self) -> tuple[int, int]:
        u, c = C.c_size_t(), C.c_size_t()
        L.check(self.lib.su2_engine_memory(self.h, C.byref(u), C.byref(c)), "engine_memory")
        return u.value, c.value

    def close(self) -> None:
        if self.h:
            self.lib.su2_engine_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Result:
    id: int
    status: int
    t_submit: float
    t_admit: float
    t_done: float
    E_mean: float
    E_err: float
    P_mean: float
    P_err: float
    accept: float
    link_updates: int
    P_fp64_twin: float = 0.0  # jobs with fp64_twin: the FP64 twin's plaquette and binned errors
    P_fp64_twin_err: float = 0.0
    P_err_binned: float = 0.0
    P_diff: float = 0.0  # P_mean - P_fp64_twin
    P_diff_sigma: float = 0.0  # combined binned error of P_diff
    E: np.ndarray | None = field(default=None, repr=False)
    P: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class Server:
    """Continuous-batching server: submit jobs any time; results come back
    as they finish.  Ticks are enqueued as launches on the server's stream
    by default: replaying them from captured hipGraphs (use_graph=True)
    measured 4-11 us slower per tick for lone 16^3-64^3 jobs on MI355X."""

    def __init__(self, precision: int = L.PREC_F64, max_jobs: int = 256, max_or: int = 1,
                 budget_bytes: int = 1 << 30, use_graph: bool = False, block: int = 0, grid: int = 0,
                 device: int = 0):
        self.lib = L.load()
        self.h = C.c_void_p()
        self.cfg = _cfg(precision, max_jobs, max_or, budget_bytes, use_graph, block, grid, device)
        L.check(self.lib.su2_server_create(C.byref(self.cfg), C.byref(self.h)), "server_create")
        self.jobs: dict[int, Job] = {}

    def now(self) -> float:
        return self.lib.su2_server_now(self.h)

    def submit(self, job: Job) -> int:
        jid = C.c_longlong()
        cj = job.c()
        L.check(self.lib.su2_server_submit(self.h, C.byref(cj), C.byref(jid)), "server_submit")
        self.jobs[jid.value] = job
        return jid.value

    def poll(self, max_results: int = 1024) -> list[Result]:
        """Finished results since the last poll.  If the server has stopped
        on an error, raises Su2Error carrying the results it still handed
        over in its `results` attribute."""
        out = (Su2JobResult * max_results)()
        n = C.c_int()
        rc = self.lib.su2_server_poll(self.h, out, max_results, C.byref(n))
        res = [Result(r.id, r.status, r.t_submit, r.t_admit, r.t_done, r.E_mean, r.E_err, r.P_mean, r.P_err,
                      r.accept, r.link_updates, r.P_fp64_twin, r.P_fp64_twin_err, r.P_err_binned, r.P_diff,
                      r.P_diff_sigma) for r in out[: n.value]]
        if rc != L.SU2_OK:
            err = L.Su2Error(rc, "server_poll", self.lib.su2_last_error().decode(errors="replace"))
            err.results = res
            raise err
        return res

    def series(self, jid: int) -> tuple[np.ndarray, np.ndarray]:
        job = self.jobs[jid]
        E = np.zeros(job.n_meas)
        P = np.zeros(job.n_meas)
        L.check(self.lib.su2_server_series(self.h, jid, E.ctypes.data, P.ctypes.data, job.n_meas), "server_series")
        return E, P

    def drain(self) -> None:
        L.check(self.lib.su2_server_drain(self.h), "server_drain")

    def close(self) -> None:
        if self.h:
            rc = self.lib.su2_server_destroy(self.h)
            self.h = C.c_void_p()
            L.check(rc, "server_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
