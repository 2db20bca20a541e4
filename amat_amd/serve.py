"""Serving CG solves on one GPU: request stream, HBM-budget admission and the
continuous-batching loop around the native server (amx_server_* in
include/amat_amd.h).

The engine (native): every active job advances by `n_iter` CG iterations per
step, each phase of the iteration one launch over all jobs; a job's result is
bit-identical alone and in any mix.  The policy (here): requests arrive on a
clock; a request is admitted when a job slot is free and its operator plus its
Krylov workspace fit the HBM budget next to everything already resident.
Operators (element matrices, maps, BC tables) stay resident in an LRU cache
keyed by (kind, n) and are evicted, least recently used first, only when no
running job uses them and room is needed.

Request mix of the serving workload: 60% hex8 Poisson n^3 with n ~ U{16..96},
30% hex20 elasticity n ~ U{8..40} with E_e ~ U[0.5, 2), 10% jittered tet10
Poisson n ~ U{12..48}; each request has its own seeded RHS in [-1, 1)
(constrained rows 0) and is solved to 1e-10 relative residual.
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import ctypes as C
import dataclasses
import math
import time
from typing import Callable

import numpy as np

from . import _capi
from ._capi import call
from .amat import BCOP, PREC, Maps, _ptr, aMatFree
from . import mesh as meshmod, problems

MIX = (("hex8", 0.6, 16, 96), ("hex20", 0.3, 8, 40), ("tet10", 0.1, 12, 48))


class Server:
    """The native continuous-batching CG server (amx_server_*)."""

    def __init__(self, device: int = 0, max_jobs: int = 64):
        h = C.c_void_p()
        call("amx_server_create", int(device), int(max_jobs), C.byref(h))
        self._h = h
        self.max_jobs = max_jobs
        self._ops = {}  # job id -> operator (kept alive while the job runs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _capi.lib().amx_server_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @staticmethod
    def job_bytes(op: aMatFree, method: PREC) -> int:
        v = C.c_uint64()
        call("amx_server_job_bytes", op.handle, int(method), C.byref(v))
        return int(v.value)

    def submit(self, op: aMatFree, method: PREC, b: np.ndarray, rtol: float = 1e-10, inner_rtol: float = 1e-2,
               max_it: int = 20000, max_outer: int = 60) -> int:
        b = np.ascontiguousarray(b, dtype=np.float64)
        jid = C.c_int64()
        call("amx_server_submit", self._h, op.handle, int(method), _ptr(b), float(rtol), float(inner_rtol),
             int(max_it), int(max_outer), C.byref(jid))
        self._ops[jid.value] = op
        return int(jid.value)

    def step(self, n_iter: int = 8) -> int:
        k = C.c_int32()
        call("amx_server_step", self._h, int(n_iter), C.byref(k))
        return k.value

    def poll(self) -> list[int]:
        buf = (C.c_int64 * self.max_jobs)()
        k = C.c_int32()
        out = []
        while True:
            call("amx_server_poll", self._h, buf, self.max_jobs, C.byref(k))
            out.extend(buf[i] for i in range(k.value))
            if k.value < self.max_jobs:
                return out

    def result(self, job_id: int, n: int | None = None):
        """(x or None, SolveStats); releases the job."""
        st = _capi.SolveStats()
        x = None
        if n is not None:
            x = np.empty(n, dtype=np.float64)
        call("amx_server_result", self._h, C.c_int64(job_id), _ptr(x), C.byref(st))
        self._ops.pop(job_id, None)
        return x, st

    def info(self) -> _capi.ServerInfo:
        i = _capi.ServerInfo()
        call("amx_server_get_info", self._h, C.byref(i))
        return i


# ----------------------------------------------------------------- workload
@dataclasses.dataclass
class Request:
    rid: int
    kind: str
    n: int
    arrival: float  # seconds after the start of the stream
    seed: int


def make_requests(count: int, rate: float, seed: int = 7, mix=MIX) -> list[Request]:
    """Poisson arrivals at `rate` per second (exponential gaps) with the kind
    and size mix drawn from counter-based uniforms (reproducible)."""
    u_gap = meshmod.uniform(seed, 11, count, 0.0, 1.0)
    u_kind = meshmod.uniform(seed, 12, count, 0.0, 1.0)
    u_size = meshmod.uniform(seed, 13, count, 0.0, 1.0)
    t = 0.0
    out = []
    cum = np.cumsum([m[1] for m in mix])
    for i in range(count):
        t += -math.log(1.0 - u_gap[i]) / rate
        k = int(np.searchsorted(cum, u_kind[i] * cum[-1], side="right"))
        kind, _, lo, hi = mix[min(k, len(mix) - 1)]
        n = lo + min(int(u_size[i] * (hi - lo + 1)), hi - lo)
        out.append(Request(i, kind, n, t, seed * 100003 + i))
    return out


def build_problem(kind: str, n: int) -> problems.Problem:
    if kind == "hex8":
        return problems.poisson_hex8(n)
    if kind == "hex20":
        return problems.elasticity_hex20(n, seed=1000 + n)
    if kind == "tet10":
        return problems.poisson_tet10(n, seed=2000 + n, with_rhs=False)
    raise ValueError(kind)


def rhs_for(p: problems.Problem, seed: int) -> np.ndarray:
    b = meshmod.uniform(seed, 5, p.n_dofs, -1.0, 1.0)
    b[p.constrained.astype(np.int64)] = 0.0
    return b


def build_operator(p: problems.Problem, precision=PREC.FP64 | PREC.FP32 | PREC.BF16, free: bool = False) -> aMatFree:
    """The served operator: stored element matrices in the given storage
    precisions, or (free=True) method 2, K_e recomputed in-kernel (FP64 CG, or FP32
    inner solves with FP64 reliable updates)."""
    if free:
        return p.build_free_operator(aMatFree, Maps, bcop=BCOP.IDENTITY)
    return p.build_operator(aMatFree, Maps, precision=precision, bcop=BCOP.IDENTITY)


# ------------------------------------------------------------------ policy
@dataclasses.dataclass
class CachedOp:
    key: tuple
    problem: problems.Problem
    op: aMatFree
    bytes: int
    users: int = 0


class OperatorCache:
    """Resident operators, LRU order; evicts unused ones to make room."""

    def __init__(self, builder: Callable[[str, int], tuple] | None = None, free: bool = False):
        self.items: collections.OrderedDict[tuple, CachedOp] = collections.OrderedDict()
        self.free = free
        self.builder = builder or self._default_builder
        self.builds = 0
        self.build_seconds = 0.0
        self.evictions = 0

    def _default_builder(self, kind: str, n: int):
        p = build_problem(kind, n)
        op = build_operator(p, free=self.free)
        return p, op, int(op.info().device_bytes)

    def estimate(self, kind: str, n: int) -> int:
        return estimate_op_bytes(kind, n, free=self.free)

    @property
    def bytes(self) -> int:
        return sum(c.bytes for c in self.items.values())

    def get(self, key):
        c = self.items.get(key)
        if c is not None:
            self.items.move_to_end(key)
        return c

    def evict_for(self, need: int, budget: int, other: int) -> bool:
        """Evict unused operators (LRU first) until `need` more bytes fit."""
        for key in list(self.items.keys()):
            if self.bytes + other + need <= budget:
                break
            c = self.items[key]
            if c.users == 0:
                del self.items[key]
                self.evictions += 1
        return self.bytes + other + need <= budget

    def make(self, key) -> tuple:
        """Build the operator of `key` without touching the cache (safe on a
        worker thread): (problem, operator, device bytes, seconds)."""
        t0 = time.perf_counter()
        p, op, nbytes = self.builder(*key)
        return p, op, nbytes, time.perf_counter() - t0

    def add(self, key, made: tuple) -> CachedOp:
        p, op, nbytes, seconds = made
        self.build_seconds += seconds
        self.builds += 1
        c = CachedOp(key, p, op, nbytes)
        self.items[key] = c
        return c

    def build(self, key) -> CachedOp:
        return self.add(key, self.make(key))


def estimate_op_bytes(kind: str, n: int, free: bool = False) -> int:
    """Device bytes of an operator of this kind and size with FP64 + FP32 +
    BF16 storage (element matrices, maps, deterministic-scatter tables), or
    with free=True the method-2 form (coordinates in FP64 and FP32 and a factor per element
    instead of matrices), used before the operator exists."""
    if kind == "hex8":
        E, nd, N, nodes = n ** 3, 8, (n + 1) ** 3, 8
    elif kind == "hex20":
        E, nd, N, nodes = n ** 3, 60, 3 * ((n + 1) ** 3 + 3 * n * (n + 1) ** 2), 20
    else:
        E, nd = 6 * n ** 3, 10
        N, nodes = (n + 1) ** 3 + 7 * n ** 3 + 3 * n * (n + 1) ** 2, 10
    per_elem = 36 * nodes + 8 if free else nd * nd * 14
    return E * per_elem + E * nd * (4 + 4 + 4 + 8 + 4) + N * 40


def estimate_job_bytes(kind: str, n: int, method: PREC) -> int:
    if kind == "hex8":
        E, nd, N = n ** 3, 8, (n + 1) ** 3
    elif kind == "hex20":
        E, nd, N = n ** 3, 60, 3 * ((n + 1) ** 3 + 3 * n * (n + 1) ** 2)
    else:
        E, nd = 6 * n ** 3, 10
        N = (n + 1) ** 3 + 7 * n ** 3 + 3 * n * (n + 1) ** 2
    ve = E * nd * 8
    return (7 * N * 8 if int(method) == PREC.FP64 else 5 * N * 8 + 6 * N * 4) + ve + 4096 * 8


@dataclasses.dataclass
class Done:
    rid: int
    kind: str
    n: int
    arrival: float
    admitted: float
    finished: float
    iterations: int
    outer: int
    rel_residual: float
    converged: bool
    x: np.ndarray | None = None


class ServingLoop:
    """Admission by HBM budget + continuous batching on one Server.

    Each turn: admit arrived requests while a slot is free and the operator +
    workspace fit the budget; run one step of the server; retire finished
    jobs.  Latency = finish - arrival; service time = finish - admission (the
    job on the GPU); queue wait = admission - arrival.

    Admission order: "fifo" takes the oldest request (head of line: one that
    does not fit waits).  "affinity" takes the oldest request among the first
    `window` waiting ones whose operator is resident, so a request that would
    build (and evict) an operator lets those that reuse one go first; no
    request is overtaken more than `max_bypass` times (the scan stops at a
    request that has been), and with every operator resident it is FIFO.

    build_workers > 0: an operator that must be built is built on that many
    worker threads while the server keeps stepping the running jobs (its
    bytes are reserved in the budget from the start); the request waiting for
    it is admitted when it is ready.  0: built inline, the server idle
    meanwhile.  Results do not depend on it (a job's arithmetic is its own)."""

    def __init__(self, server: Server, cache: OperatorCache, method: PREC, budget_bytes: int, n_iter: int = 8,
                 rtol: float = 1e-10, inner_rtol: float = 1e-2, keep_x: bool = False,
                 clock: Callable[[], float] = time.perf_counter, admission: str = "fifo", window: int = 256,
                 max_bypass: int = 64, build_workers: int = 0):
        if admission not in ("fifo", "affinity"):
            raise ValueError(f"admission must be 'fifo' or 'affinity', not {admission!r}")
        self.server = server
        self.cache = cache
        self.method = int(method)
        self.budget = int(budget_bytes)
        self.n_iter = n_iter
        self.rtol = rtol
        self.inner_rtol = inner_rtol
        self.keep_x = keep_x
        self.clock = clock
        self.admission = admission
        self.window = max(1, int(window))
        self.max_bypass = max(0, int(max_bypass))
        self.build_workers = max(0, int(build_workers))

    def _pick(self, pending, bypassed: dict) -> int:
        """Index in `pending` of the request to admit next."""
        if self.admission == "fifo":
            return 0
        for j in range(min(self.window, len(pending))):
            r = pending[j]
            if (r.kind, r.n) in self.cache.items or bypassed.get(r.rid, 0) >= self.max_bypass:
                return j
        return 0

    def run(self, requests: list[Request], realtime: bool = True,
            progress: Callable[[float, int, int, int], None] | None = None) -> dict:
        """Serve `requests` (arrival times relative to the start); progress, if
        given, is called as progress(elapsed_s, finished, running, pending)."""
        pending = collections.deque()
        incoming = collections.deque(sorted(requests, key=lambda r: r.arrival))
        running = {}  # job id -> (request, CachedOp, admitted time, workspace bytes)
        done: list[Done] = []
        workspace = 0
        t0 = self.clock()
        peak_jobs = 0
        peak_bytes = 0
        builds0, evict0, build_s0 = self.cache.builds, self.cache.evictions, self.cache.build_seconds
        budget_waits = 0   # turns on which the chosen request waited for HBM (a slot was free)
        bypassed: dict[int, int] = {}  # request id -> times a later request was admitted first
        pool = cf.ThreadPoolExecutor(max_workers=self.build_workers) if self.build_workers else None
        building: dict[tuple, tuple[cf.Future, int]] = {}  # key -> (build, bytes reserved for it)
        reserved = 0
        try:
            while incoming or pending or running or building:
                now = self.clock() - t0
                while incoming and (incoming[0].arrival <= now or not realtime):
                    pending.append(incoming.popleft())
                # operators built in the background join the cache (on this thread)
                for key in [k for k, (f, _) in building.items() if f.done()]:
                    f, nb = building.pop(key)
                    reserved -= nb
                    self.cache.add(key, f.result())
                # admission: the picked request waits (and so does everything behind it) if it does not fit
                while pending and len(running) < self.server.max_jobs:
                    idx = self._pick(pending, bypassed)
                    r = pending[idx]
                    key = (r.kind, r.n)
                    c = self.cache.get(key)
                    if c is None and key in building:
                        break  # its operator is being built
                    need_op = 0 if c is not None else self.cache.estimate(r.kind, r.n)
                    need_job = estimate_job_bytes(r.kind, r.n, self.method)
                    if not self.cache.evict_for(need_op + need_job, self.budget, workspace + reserved):
                        if not running and not building:
                            raise RuntimeError(f"request {r.rid} ({r.kind} n={r.n}) cannot fit the HBM budget")
                        budget_waits += 1
                        break
                    if c is None:
                        if pool is not None:
                            building[key] = (pool.submit(self.cache.make, key), need_op)
                            reserved += need_op
                            break
                        c = self.cache.build(key)
                    b = rhs_for(c.problem, r.seed)
                    jid = self.server.submit(c.op, self.method, b, rtol=self.rtol, inner_rtol=self.inner_rtol)
                    c.users += 1
                    jb = self.server.job_bytes(c.op, self.method)
                    workspace += jb
                    running[jid] = (r, c, self.clock() - t0, jb)
                    for j in range(idx):
                        bypassed[pending[j].rid] = bypassed.get(pending[j].rid, 0) + 1
                    del pending[idx]
                    bypassed.pop(r.rid, None)
                peak_jobs = max(peak_jobs, len(running))
                peak_bytes = max(peak_bytes, workspace + self.cache.bytes + reserved)
                if progress is not None:
                    progress(self.clock() - t0, len(done), len(running), len(pending) + len(incoming))
                if running:
                    self.server.step(self.n_iter)
                    for jid in self.server.poll():
                        r, c, t_adm, jb = running.pop(jid)
                        x, st = self.server.result(jid, c.problem.n_dofs if self.keep_x else None)
                        c.users -= 1
                        workspace -= jb
                        done.append(Done(r.rid, r.kind, r.n, r.arrival, t_adm, self.clock() - t0, st.iterations,
                                         st.outer_iterations, st.rel_residual, bool(st.converged), x))
                elif building:
                    cf.wait([f for f, _ in building.values()], timeout=0.01, return_when=cf.FIRST_COMPLETED)
                elif incoming and realtime:
                    time.sleep(max(0.0, min(incoming[0].arrival - (self.clock() - t0), 0.01)))
        finally:
            if pool is not None:
                pool.shutdown(wait=True)
        makespan = self.clock() - t0
        lat = np.array([d.finished - d.arrival for d in done])
        svc = np.array([d.finished - d.admitted for d in done])
        wait = np.array([d.admitted - d.arrival for d in done])

        def pct(a, q):
            return float(np.percentile(a, q)) if len(a) else 0.0
        return {
            "done": done,
            "solves": len(done),
            "makespan_s": makespan,
            "solves_per_s": len(done) / makespan if makespan > 0 else 0.0,
            "p50_s": pct(lat, 50),
            "p99_s": pct(lat, 99),
            "service_p50_s": pct(svc, 50),
            "service_p99_s": pct(svc, 99),
            "queue_wait_p50_s": pct(wait, 50),
            "queue_wait_p99_s": pct(wait, 99),
            "operator_builds": self.cache.builds - builds0,
            "operator_build_s": self.cache.build_seconds - build_s0,
            "evictions": self.cache.evictions - evict0,
            "budget_waits": budget_waits,
            "admission": self.admission,
            "build_workers": self.build_workers,
            "all_converged": all(d.converged for d in done),
            "max_rel_residual": max((d.rel_residual for d in done), default=0.0),
            "peak_jobs": peak_jobs,
            "peak_resident_bytes": peak_bytes,
        }
