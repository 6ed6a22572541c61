"""Continuous-batching server of cloth-drape jobs on one GPU.

Jobs (a cloth spec and a substep count) arrive over time.  The server keeps
every admitted job's state resident in HBM (its ClothSimulator arrays and
torch-owned pos/vel/inv_mass) and advances all running jobs together, one
fused substep per loop iteration (`ClothBatch.step`: one element pass, one
gather pass, one persistent PCG launch for the whole mix).

A job moves through four states:

  waiting    arrived, not admitted yet;
  preparing  admitted first come first served while its estimated HBM
             footprint fits the budget and fewer than `max_preparing` jobs
             are in set-up (by default twice the executor's set-up workers,
             so that under overload late arrivals wait without holding HBM);
             its set-up (topology, Hessian pattern, scatter maps, device
             arrays) runs on a worker thread while the GPU keeps stepping
             the running jobs;
  running    in the batch; it joins at the next substep boundary once its
             set-up is done;
  finished   after its last substep; it leaves the batch and frees its memory.

So requests that arrive while others run join as soon as memory and their
set-up allow, and their substeps overlap on the device with everyone
else's.  Each job's result is identical alone and in any mix (the batch runs
the same kernels on the job's own arrays with per-instance reductions).

Reported: throughput (vertex-substeps/s), p50/p99 request latency (arrival
to the end of the substep that finishes the job) over the completed jobs and,
when the server stops at a horizon, over every job that had arrived by then
with the unfinished ones counted as censored at the horizon (their latency is
at least horizon - arrival, so these percentiles are lower bounds), queueing
and set-up delays, peak concurrency and peak HBM in use.  With `max_queue_s`
the server sheds load: a job still waiting for admission that long after its
arrival is rejected, and the rejection count is reported.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

# Device bytes per vertex of a cloth grid instance (matrix, gather map,
# element records, PCG vectors, external state), used for admission before a
# job's exact footprint is known; measured afterwards.
BYTES_PER_VERTEX_ESTIMATE = 1800


@dataclass
class Request:
    job_id: int
    arrival: float                 # seconds since the start of the stream
    spec: object                   # chysx.scenes.ClothSpec, or a JobSpec built on demand
    substeps: int
    admitted: float | None = None  # set-up started
    joined: float | None = None    # entered the batch
    finished: float | None = None
    done: int = 0
    handle: object = None
    bytes: int = 0
    result: object = None
    future: object = None
    rejected: bool = False         # shed at admission


@dataclass
class ServingReport:
    completed: int
    submitted: int
    vertex_substeps: float
    wall_s: float
    throughput: float
    p50_latency_s: float | None
    p99_latency_s: float | None
    p50_queue_s: float | None
    p50_setup_s: float | None
    max_concurrency: int
    peak_bytes: int
    substep_launches: int
    unfinished: int = 0
    arrived: int = 0                              # arrivals by the end of the run
    rejected: int = 0                             # shed at admission (max_queue_s)
    p50_latency_censored_s: float | None = None   # over every arrived, not rejected job;
    p99_latency_censored_s: float | None = None   # unfinished ones censored at the end
    latencies: list = field(default_factory=list)

    def as_dict(self):
        d = dict(self.__dict__)
        d.pop("latencies", None)
        return d


class GpuExecutor:
    """Runs admitted jobs on the GPU through one chysx.ClothBatch; job set-up
    runs on `prepare_workers` host threads and its own stream."""

    def __init__(self, lanes_per_row: int = 8, stream=None, keep_results: bool = False, prepare_workers: int = 2,
                 tile_rows: int | None = None, grid_blocks: int | None = None, lds_staging: bool = True,
                 keep_only=None):
        import concurrent.futures

        import torch

        import chysx

        self.torch = torch
        self.batch = chysx.ClothBatch()
        self.batch.set_lanes_per_row(lanes_per_row)
        self.lanes = lanes_per_row
        if tile_rows:
            self.batch.set_tile_rows(tile_rows)
        if grid_blocks:
            self.batch.set_grid_blocks(grid_blocks)
        self.batch.set_lds_staging(lds_staging)
        import queue

        self.stream = stream if stream is not None else torch.cuda.Stream()
        # One preparation stream per worker, so set-ups wait only for
        # themselves (never for the batch or for each other).
        self.prep_streams = queue.SimpleQueue()
        for _ in range(max(1, prepare_workers)):
            self.prep_streams.put(torch.cuda.Stream())
        self.keep_results = keep_results
        self.keep_only = keep_only  # predicate on the request: whose final state to keep
        self.prepare_workers = prepare_workers
        self.pool = concurrent.futures.ThreadPoolExecutor(max_workers=prepare_workers) if prepare_workers > 0 else None

    def _build(self, spec, dt):
        from . import scenes

        if hasattr(spec, "build"):  # a job spec built on demand
            spec = spec.build()
        st = self.prep_streams.get()
        try:
            cloth = scenes.DeviceCloth(spec, lanes_per_row=self.lanes, stream=st)
            # Set-up uploads and waits on the preparation stream only: the
            # batch keeps stepping on its own stream meanwhile.
            cloth.sim.prepare(dt, st.cuda_stream)
        finally:
            self.prep_streams.put(st)
        return cloth

    def begin(self, req: Request, dt: float):
        if self.pool is not None:
            req.future = self.pool.submit(self._build, req.spec, dt)

    def ready(self, req: Request) -> bool:
        return req.future is None or req.future.done()

    def admit(self, req: Request, dt: float):
        cloth = req.future.result() if req.future is not None else self._build(req.spec, dt)
        req.future = None
        req.handle = cloth
        req.bytes = int(cloth.sim.device_bytes() + 28 * req.spec.n)
        self.batch.add_instance(cloth.sim)

    def retire(self, req: Request):
        cloth = req.handle
        if self.keep_results and (self.keep_only is None or self.keep_only(req)):
            req.result = cloth.state()
        self.batch.remove_instance(cloth.sim)
        req.handle = None

    def step(self, dt: float):
        with self.torch.cuda.stream(self.stream):
            self.batch.step(dt, self.stream.cuda_stream)

    def sync(self):
        self.stream.synchronize()

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True)


class ServingEngine:
    def __init__(self, executor, hbm_budget_bytes: float = 200e9, max_concurrency: int = 1 << 30,
                 clock=time.perf_counter, sleep=time.sleep, max_preparing: int | None = None,
                 max_queue_s: float | None = None):
        self.ex = executor
        self.max_queue_s = max_queue_s
        self.budget = float(hbm_budget_bytes)
        self.max_conc = max_concurrency
        if max_preparing is None:
            workers = getattr(executor, "prepare_workers", None)
            max_preparing = 2 * workers if workers else 1 << 30
        self.max_prep = max(1, int(max_preparing))
        self.clock = clock
        self.sleep = sleep

    def _begin(self, req, dt):
        begin = getattr(self.ex, "begin", None)
        if begin is not None:
            begin(req, dt)

    def _ready(self, req):
        ready = getattr(self.ex, "ready", None)
        return True if ready is None else ready(req)

    def _admit(self, req, dt):
        self.ex.admit(req, dt)

    def run(self, requests, dt: float, horizon_s: float | None = None, log=None) -> ServingReport:
        """Serves `requests` (sorted by arrival).  Arrivals are replayed in
        real time from the start of the call; with `horizon_s`, the server
        stops at that time and reports the jobs completed by then."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        preparing: list[Request] = []
        active: list[Request] = []
        done: list[Request] = []
        in_use = 0
        rejected = 0
        peak_bytes = 0
        max_conc = 0
        launches = 0
        vs = 0.0
        t0 = self.clock()
        while pending or waiting or preparing or active:
            now = self.clock() - t0
            if horizon_s is not None and now > horizon_s:
                break
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # load shedding: jobs that waited too long for admission leave
            if self.max_queue_s is not None:
                while waiting and now - waiting[0].arrival > self.max_queue_s:
                    waiting.popleft().rejected = True
                    rejected += 1
            # admission: FCFS while the estimated footprint fits and the
            # set-up queue has room
            while waiting and len(active) + len(preparing) < self.max_conc and len(preparing) < self.max_prep:
                req = waiting[0]
                est = BYTES_PER_VERTEX_ESTIMATE * req.spec.n
                if (active or preparing) and in_use + est > self.budget:
                    break
                waiting.popleft()
                req.admitted = self.clock() - t0
                req.bytes = est
                in_use += est
                self._begin(req, dt)
                preparing.append(req)
            # join: every prepared job enters the batch at this boundary; with
            # nothing running, wait for the oldest set-up.
            joined = [r for r in preparing if self._ready(r)]
            if not joined and not active and preparing:
                joined = [preparing[0]]
            for req in joined:
                preparing.remove(req)
                est = req.bytes
                self._admit(req, dt)
                in_use += (req.bytes or est) - est
                req.joined = self.clock() - t0
                active.append(req)
            if not active:
                if pending and not preparing:
                    self.sleep(max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.01)))
                continue
            max_conc = max(max_conc, len(active))
            peak_bytes = max(peak_bytes, in_use)
            self.ex.step(dt)
            launches += 1
            self.ex.sync()
            t_end = self.clock() - t0
            still = []
            for req in active:
                req.done += 1
                vs += req.spec.n
                if req.done >= req.substeps:
                    req.finished = t_end
                    self.ex.retire(req)
                    in_use -= req.bytes
                    done.append(req)
                else:
                    still.append(req)
            active = still
            if log and launches % 200 == 0:
                log(f"serving t={t_end:.1f}s running={len(active)} preparing={len(preparing)} "
                    f"waiting={len(waiting)} done={len(done)}")
        wall = self.clock() - t0
        for req in active:  # stopped at the horizon: release unfinished jobs
            self.ex.retire(req)
        for req in preparing:
            self._admit(req, dt)
            self.ex.retire(req)
        lat = [r.finished - r.arrival for r in done]
        qd = [r.admitted - r.arrival for r in done]
        sd = [r.joined - r.admitted for r in done]
        # every job that arrived by the end and was not shed: finished ones
        # with their latency, the others censored at the end of the run
        arrived = [r for r in requests if r.arrival <= wall]
        cens = [(r.finished if r.finished is not None else wall) - r.arrival
                for r in arrived if not getattr(r, "rejected", False)]
        pct = lambda a, q: float(np.percentile(a, q)) if a else None  # noqa: E731
        return ServingReport(
            completed=len(done), submitted=len(requests), vertex_substeps=vs, wall_s=wall,
            throughput=vs / wall if wall > 0 else 0.0, p50_latency_s=pct(lat, 50), p99_latency_s=pct(lat, 99),
            p50_queue_s=pct(qd, 50), p50_setup_s=pct(sd, 50), max_concurrency=max_conc, peak_bytes=int(peak_bytes),
            substep_launches=launches, unfinished=len(requests) - len(done), arrived=len(arrived), rejected=rejected,
            p50_latency_censored_s=pct(cens, 50), p99_latency_censored_s=pct(cens, 99), latencies=lat)
