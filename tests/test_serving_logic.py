"""Scheduling logic of the continuous-batching server on a simulated clock
(no GPU): FCFS admission under the HBM budget, every job runs exactly its
substep count, concurrent jobs share substeps, latency accounting."""
import numpy as np

from chysx import serving


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, s):
        self.t += max(s, 1e-4)


class FakeExecutor:
    def __init__(self, clock, cost_per_vertex=1e-8, fixed=1e-4):
        self.clock, self.c, self.fixed = clock, cost_per_vertex, fixed
        self.active = []
        self.admitted = []
        self.max_bytes = 0

    def admit(self, req, dt=None):
        req.bytes = req.spec.n * 1000
        self.active.append(req)
        self.admitted.append(req.job_id)

    def retire(self, req):
        self.active.remove(req)

    def step(self, dt):
        self.clock.t += self.fixed + self.c * sum(r.spec.n for r in self.active)
        self.max_bytes = max(self.max_bytes, sum(r.bytes for r in self.active))

    def sync(self):
        pass


class Spec:
    def __init__(self, n):
        self.n = n


def _requests(count=40, seed=0):
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    for i in range(count):
        t += rng.exponential(0.01)
        reqs.append(serving.Request(job_id=i, arrival=t, spec=Spec(int(rng.integers(1000, 20000))),
                                    substeps=int(rng.integers(5, 30))))
    return reqs


def test_all_jobs_complete_fcfs_under_budget():
    clock = FakeClock()
    ex = FakeExecutor(clock)
    budget = 60000 * serving.BYTES_PER_VERTEX_ESTIMATE
    eng = serving.ServingEngine(ex, hbm_budget_bytes=budget, clock=clock, sleep=clock.sleep)
    reqs = _requests()
    rep = eng.run(reqs, dt=1 / 600)
    assert rep.completed == len(reqs)
    assert ex.admitted == sorted(ex.admitted)  # first come, first served
    assert all(r.done == r.substeps for r in reqs)
    assert ex.max_bytes <= budget
    assert rep.vertex_substeps == sum(r.spec.n * r.substeps for r in reqs)
    assert rep.p99_latency_s >= rep.p50_latency_s > 0
    for r in reqs:
        assert r.finished - r.arrival >= r.substeps * ex.fixed - 1e-12
        assert r.admitted >= r.arrival
    assert rep.max_concurrency > 1  # jobs overlapped on the device


def test_oversized_job_runs_alone():
    clock = FakeClock()
    ex = FakeExecutor(clock)
    eng = serving.ServingEngine(ex, hbm_budget_bytes=10, clock=clock, sleep=clock.sleep)
    reqs = [serving.Request(job_id=i, arrival=0.0, spec=Spec(1000), substeps=3) for i in range(3)]
    rep = eng.run(reqs, dt=1 / 600)
    assert rep.completed == 3 and rep.max_concurrency == 1


def test_horizon_stops_and_releases():
    clock = FakeClock()
    ex = FakeExecutor(clock, fixed=1e-2)
    eng = serving.ServingEngine(ex, hbm_budget_bytes=1e12, clock=clock, sleep=clock.sleep)
    reqs = _requests(count=20)
    rep = eng.run(reqs, dt=1 / 600, horizon_s=0.1)
    assert rep.completed < len(reqs)
    assert ex.active == []


class SetupExecutor(FakeExecutor):
    """Set-up takes `setup` seconds of the simulated clock per job, on
    `prepare_workers` workers."""

    def __init__(self, clock, setup=0.05, prepare_workers=2, **kw):
        super().__init__(clock, **kw)
        self.setup, self.prepare_workers = setup, prepare_workers
        self.in_setup = 0
        self.max_in_setup = 0

    def begin(self, req, dt):
        req.ready_at = self.clock.t + self.setup
        self.in_setup += 1
        self.max_in_setup = max(self.max_in_setup, self.in_setup)

    def ready(self, req):
        return self.clock.t >= req.ready_at

    def admit(self, req, dt=None):
        self.in_setup -= 1
        super().admit(req, dt)


def test_setup_queue_is_bounded():
    """Under overload at most twice the set-up workers are preparing at
    once (the rest wait without holding HBM); jobs still run FCFS and all
    complete."""
    clock = FakeClock()
    ex = SetupExecutor(clock, setup=0.05, prepare_workers=2)
    eng = serving.ServingEngine(ex, hbm_budget_bytes=1e12, clock=clock, sleep=clock.sleep)
    reqs = _requests(count=60, seed=3)
    rep = eng.run(reqs, dt=1 / 600)
    assert rep.completed == len(reqs)
    assert ex.max_in_setup <= 4
    assert ex.admitted == sorted(ex.admitted)
    eng2 = serving.ServingEngine(SetupExecutor(FakeClock(), prepare_workers=2), hbm_budget_bytes=1e12,
                                 max_preparing=1)
    assert eng2.max_prep == 1


def test_horizon_counts_unfinished_jobs_as_censored():
    """Overload (each substep costs more than the arrivals allow) stopped at
    a horizon: the censored percentiles cover every arrived job, the
    unfinished ones at horizon - arrival, so they bound the completed-only
    percentiles from above, and every arrival is counted."""
    clock = FakeClock()
    ex = FakeExecutor(clock, cost_per_vertex=2e-7)
    reqs = _requests(count=200, seed=3)
    rep = serving.ServingEngine(ex, hbm_budget_bytes=1e12, clock=clock, sleep=clock.sleep).run(
        reqs, 1.0, horizon_s=1.0)
    assert 0 < rep.completed < rep.arrived <= rep.submitted
    unfinished = [r for r in reqs if r.arrival <= rep.wall_s and r.finished is None]
    assert len(unfinished) == rep.arrived - rep.completed
    cens = [r.finished - r.arrival for r in reqs if r.finished is not None] + \
           [rep.wall_s - r.arrival for r in unfinished]
    assert rep.p99_latency_censored_s == float(np.percentile(cens, 99))
    assert rep.p50_latency_censored_s >= rep.p50_latency_s


def test_load_shedding_rejects_jobs_that_waited_too_long():
    """With max_queue_s, a job not admitted within that time of its arrival
    is rejected: no admitted job waited longer, and rejected + completed
    account for every job once the stream drains."""
    clock = FakeClock()
    ex = FakeExecutor(clock, cost_per_vertex=2e-7)
    reqs = _requests(count=200, seed=4)
    rep = serving.ServingEngine(ex, hbm_budget_bytes=50e6, clock=clock, sleep=clock.sleep, max_queue_s=0.2).run(
        reqs, 1.0)
    assert rep.rejected > 0 and rep.rejected + rep.completed == rep.submitted
    for r in reqs:
        if r.rejected:
            assert r.admitted is None
        else:
            assert r.admitted - r.arrival <= 0.2 + 1e-3
