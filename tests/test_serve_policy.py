"""CPU tests of the serving policy (amat_amd.serve): the request stream, FIFO
admission by slots and HBM budget, LRU operator eviction, latency metrics.
The native server is replaced by a deterministic stand-in whose jobs finish
after a fixed number of steps (the engine itself is tested on the GPU)."""
import numpy as np
import pytest

from amat_amd import PREC, serve


class FakeStats:
    def __init__(self):
        self.iterations, self.outer_iterations, self.rel_residual, self.converged = 10, 0, 1e-11, 1


class FakeServer:
    def __init__(self, max_jobs, steps_per_job=3):
        self.max_jobs = max_jobs
        self.steps_per_job = steps_per_job
        self.jobs = {}
        self.finished = []
        self.next = 1
        self.active_max = 0
        self.steps = 0

    @staticmethod
    def job_bytes(op, method):
        return op["bytes_job"]

    def submit(self, op, method, b, rtol=1e-10, inner_rtol=1e-2):
        assert len(self.jobs) < self.max_jobs
        jid = self.next
        self.next += 1
        self.jobs[jid] = self.steps_per_job
        self.active_max = max(self.active_max, len(self.jobs))
        return jid

    def step(self, n_iter):
        self.steps += 1
        for j in list(self.jobs):
            self.jobs[j] -= 1
            if self.jobs[j] == 0:
                del self.jobs[j]
                self.finished.append(j)

    def poll(self):
        out, self.finished = self.finished, []
        return out

    def result(self, jid, n=None):
        return None, FakeStats()


class FakeProblem:
    n_dofs = 10
    constrained = np.zeros(0, np.uint64)


def fake_builder(log):
    def build(kind, n):
        log.append((kind, n))
        return FakeProblem(), {"bytes_job": 100}, 1000 * n
    return build


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def test_make_requests_mix_and_rate():
    reqs = serve.make_requests(4000, rate=200.0, seed=1)
    kinds = np.array([r.kind for r in reqs])
    assert abs((kinds == "hex8").mean() - 0.6) < 0.03
    assert abs((kinds == "hex20").mean() - 0.3) < 0.03
    assert abs((kinds == "tet10").mean() - 0.1) < 0.02
    for kind, _, lo, hi in serve.MIX:
        ns = [r.n for r in reqs if r.kind == kind]
        assert min(ns) >= lo and max(ns) <= hi and min(ns) <= lo + 3 and max(ns) >= hi - 3
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert abs(len(reqs) / t[-1] - 200.0) / 200.0 < 0.05
    assert serve.make_requests(10, 50.0, seed=4) == serve.make_requests(10, 50.0, seed=4)


def test_budget_slots_fifo_and_lru(monkeypatch):
    monkeypatch.setattr(serve, "rhs_for", lambda p, seed: np.zeros(p.n_dofs))
    monkeypatch.setattr(serve, "estimate_op_bytes", lambda kind, n, free=False: 1000 * n)
    monkeypatch.setattr(serve, "estimate_job_bytes", lambda kind, n, m: 100)
    log = []
    cache = serve.OperatorCache(builder=fake_builder(log))
    srv = FakeServer(max_jobs=3)
    mix = (("hex8", 1.0, 1, 9),)
    reqs = serve.make_requests(40, rate=1e9, seed=2, mix=mix)
    budget = 12_000  # at most a few operators resident at once
    loop = serve.ServingLoop(srv, cache, PREC.FP64, budget, clock=Clock())
    m = loop.run(reqs, realtime=False)
    assert m["solves"] == 40 and m["all_converged"]
    assert srv.active_max <= 3 and m["peak_jobs"] <= 3
    assert m["peak_resident_bytes"] <= budget
    assert cache.evictions > 0 and cache.builds > len(set((r.kind, r.n) for r in reqs)) - 1
    # FIFO: admission order follows arrival order
    adm = [d.admitted for d in sorted(m["done"], key=lambda d: d.rid)]
    assert all(a <= b for a, b in zip(adm, adm[1:]))
    lat = [d.finished - d.arrival for d in m["done"]]
    assert m["p50_s"] == pytest.approx(float(np.percentile(lat, 50)))
    # service time (admission -> finish) and queue wait (arrival -> admission) split the latency
    svc = [d.finished - d.admitted for d in m["done"]]
    wait = [d.admitted - d.arrival for d in m["done"]]
    assert m["service_p99_s"] == pytest.approx(float(np.percentile(svc, 99)))
    assert m["queue_wait_p50_s"] == pytest.approx(float(np.percentile(wait, 50)))
    assert all(s > 0 for s in svc) and all(w >= 0 for w in wait)
    # the budget bound inside the stream: builds, evictions and waits are counted by the run
    assert m["operator_builds"] == cache.builds and m["evictions"] == cache.evictions > 0
    assert m["budget_waits"] > 0


def test_request_that_never_fits_raises(monkeypatch):
    monkeypatch.setattr(serve, "rhs_for", lambda p, seed: np.zeros(p.n_dofs))
    monkeypatch.setattr(serve, "estimate_op_bytes", lambda kind, n, free=False: 10 ** 9)
    monkeypatch.setattr(serve, "estimate_job_bytes", lambda kind, n, m: 100)
    cache = serve.OperatorCache(builder=fake_builder([]))
    loop = serve.ServingLoop(FakeServer(2), cache, PREC.FP64, 1000, clock=Clock())
    with pytest.raises(RuntimeError):
        loop.run(serve.make_requests(2, 10.0, mix=(("hex8", 1.0, 4, 4),)), realtime=False)


def test_estimates_cover_small_meshes():
    from amat_amd import mesh
    for kind, n in (("hex8", 6), ("hex20", 3), ("tet10", 3)):
        p = serve.build_problem(kind, n)
        E, nd = p.gmap.shape
        assert serve.estimate_op_bytes(kind, n) >= E * nd * nd * 14
        assert serve.estimate_job_bytes(kind, n, PREC.FP64) >= 7 * p.n_dofs * 8


def _run_policy(monkeypatch, admission, max_bypass=8, n_req=200, budget=12_000, warm=False):
    monkeypatch.setattr(serve, "rhs_for", lambda p, seed: np.zeros(p.n_dofs))
    monkeypatch.setattr(serve, "estimate_op_bytes", lambda kind, n, free=False: 1000 * n)
    monkeypatch.setattr(serve, "estimate_job_bytes", lambda kind, n, m: 100)
    log = []
    cache = serve.OperatorCache(builder=fake_builder(log))
    srv = FakeServer(max_jobs=3)
    reqs = serve.make_requests(n_req, rate=1e9, seed=5, mix=(("hex8", 1.0, 1, 9),))
    if warm:
        for key in sorted({(r.kind, r.n) for r in reqs}):
            cache.build(key)
    loop = serve.ServingLoop(srv, cache, PREC.FP64, budget, clock=Clock(), admission=admission,
                             max_bypass=max_bypass)
    return loop.run(reqs, realtime=False), cache


def test_affinity_admission_builds_less_and_bounds_overtaking(monkeypatch):
    fifo, cf = _run_policy(monkeypatch, "fifo")
    aff, ca = _run_policy(monkeypatch, "affinity", max_bypass=8)
    assert fifo["solves"] == aff["solves"] == 200 and aff["all_converged"]
    assert aff["admission"] == "affinity" and fifo["admission"] == "fifo"
    assert aff["peak_resident_bytes"] <= 12_000 and aff["peak_jobs"] <= 3
    # requests that reuse a resident operator go first: fewer builds and evictions than FIFO
    assert ca.builds < cf.builds and ca.evictions < cf.evictions
    # no request is overtaken by more than max_bypass later requests (arrival order = rid order here)
    order = [d.rid for d in sorted(aff["done"], key=lambda d: (d.admitted, d.rid))]
    pos = {rid: i for i, rid in enumerate(order)}
    for rid in pos:
        overtook = sum(1 for o in pos if o > rid and pos[o] < pos[rid])
        assert overtook <= 8, (rid, overtook)
    # FIFO admits strictly in arrival order
    forder = [d.rid for d in sorted(fifo["done"], key=lambda d: (d.admitted, d.rid))]
    assert forder == sorted(forder)


def test_affinity_is_fifo_when_every_operator_is_resident(monkeypatch):
    fifo, _ = _run_policy(monkeypatch, "fifo", budget=10 ** 9, warm=True)
    aff, _ = _run_policy(monkeypatch, "affinity", budget=10 ** 9, warm=True)
    assert [d.rid for d in aff["done"]] == [d.rid for d in fifo["done"]]
    assert aff["operator_builds"] == fifo["operator_builds"]


def test_unknown_admission_is_rejected():
    with pytest.raises(ValueError):
        serve.ServingLoop(FakeServer(1), serve.OperatorCache(builder=fake_builder([])), PREC.FP64, 1, admission="lifo")


@pytest.mark.parametrize("admission", ["fifo", "affinity"])
def test_background_builds_overlap_the_server(monkeypatch, admission):
    """With build workers the server keeps stepping while an operator is
    built; every request is served, within the slots and the budget (the
    bytes of an operator under construction are reserved)."""
    import time as _time
    monkeypatch.setattr(serve, "rhs_for", lambda p, seed: np.zeros(p.n_dofs))
    monkeypatch.setattr(serve, "estimate_op_bytes", lambda kind, n, free=False: 1000 * n)
    monkeypatch.setattr(serve, "estimate_job_bytes", lambda kind, n, m: 100)
    class SlowServer(FakeServer):
        def step(self, n_iter):   # a step takes GPU time
            _time.sleep(0.001)
            super().step(n_iter)

    srv = SlowServer(max_jobs=3, steps_per_job=20)
    overlap = []

    def slow_build(kind, n):
        s0 = srv.steps
        _time.sleep(0.01)
        overlap.append(srv.steps - s0)
        return FakeProblem(), {"bytes_job": 100}, 1000 * n

    cache = serve.OperatorCache(builder=slow_build)
    reqs = serve.make_requests(60, rate=1e9, seed=3, mix=(("hex8", 1.0, 1, 6),))
    loop = serve.ServingLoop(srv, cache, PREC.FP64, 16_000, clock=Clock(), admission=admission, build_workers=2)
    m = loop.run(reqs, realtime=False)
    assert m["solves"] == 60 and m["all_converged"] and m["build_workers"] == 2
    assert m["peak_resident_bytes"] <= 16_000 and srv.active_max <= 3
    assert m["operator_builds"] == len(overlap) == cache.builds > 0
    assert any(k > 0 for k in overlap)   # the server stepped during some build
    if admission == "fifo":
        order = [d.rid for d in sorted(m["done"], key=lambda d: (d.admitted, d.rid))]
        assert order == sorted(order)
