"""Continuous-batching server of cloth-drape jobs on one GPU.

Jobs (a cloth spec and a substep count) arrive over time.  The server keeps
every admitted job's state resident in HBM (its ClothSimulator arrays and
torch-owned pos/vel/inv_mass) and advances all admitted jobs together, one
fused substep per loop iteration (`ClothBatch.step`: one element pass, one
gather pass, one persistent PCG launch for the whole mix).  Between substeps
it retires finished jobs and admits waiting ones, first come first served,
while their estimated HBM footprint fits the budget — so requests that
arrive while others run join as soon as memory frees up and their work
overlaps on the device.  Each job's result is identical alone and in any
mix (the batch runs the same kernels on the job's own arrays with
per-instance reductions).

Reported: throughput (vertex-substeps/s), p50/p99 request latency (arrival
to the end of the substep that finishes the job), queueing delay, peak
concurrency and peak HBM in use.
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
    spec: object                   # chysx.scenes.ClothSpec
    substeps: int
    admitted: float | None = None
    finished: float | None = None
    done: int = 0
    handle: object = None
    bytes: int = 0
    result: object = None


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
    max_concurrency: int
    peak_bytes: int
    substep_launches: int
    latencies: list = field(default_factory=list)

    def as_dict(self):
        d = dict(self.__dict__)
        d.pop("latencies", None)
        return d


class GpuExecutor:
    """Runs admitted jobs on the GPU through one chysx.ClothBatch."""

    def __init__(self, lanes_per_row: int = 8, stream=None, keep_results: bool = False):
        import torch

        import chysx

        self.torch = torch
        self.batch = chysx.ClothBatch()
        self.batch.set_lanes_per_row(lanes_per_row)
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self.keep_results = keep_results

    def admit(self, req: Request):
        from . import scenes

        with self.torch.cuda.stream(self.stream):
            cloth = scenes.DeviceCloth(req.spec)
        req.handle = cloth
        req.bytes = int(cloth.sim.device_bytes() + 28 * req.spec.n)
        self.batch.add_instance(cloth.sim)

    def retire(self, req: Request):
        cloth = req.handle
        if self.keep_results:
            req.result = cloth.state()
        self.batch.remove_instance(cloth.sim)
        req.handle = None

    def step(self, dt: float):
        with self.torch.cuda.stream(self.stream):
            self.batch.step(dt, self.stream.cuda_stream)

    def sync(self):
        self.stream.synchronize()


class ServingEngine:
    def __init__(self, executor, hbm_budget_bytes: float = 200e9, max_concurrency: int = 1 << 30,
                 clock=time.perf_counter, sleep=time.sleep):
        self.ex = executor
        self.budget = float(hbm_budget_bytes)
        self.max_conc = max_concurrency
        self.clock = clock
        self.sleep = sleep

    def run(self, requests, dt: float, horizon_s: float | None = None, log=None) -> ServingReport:
        """Serves `requests` (sorted by arrival).  Arrivals are replayed in
        real time from the start of the call; with `horizon_s`, the server
        stops admitting at that time and reports the jobs completed."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        active: list[Request] = []
        done: list[Request] = []
        in_use = 0
        peak_bytes = 0
        max_conc = 0
        launches = 0
        vs = 0.0
        t0 = self.clock()
        while pending or waiting or active:
            now = self.clock() - t0
            if horizon_s is not None and now > horizon_s:
                break
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: FCFS while the estimated footprint fits
            while waiting and len(active) < self.max_conc:
                req = waiting[0]
                est = BYTES_PER_VERTEX_ESTIMATE * req.spec.n
                if active and in_use + est > self.budget:
                    break
                waiting.popleft()
                self.ex.admit(req)
                req.admitted = self.clock() - t0
                in_use += req.bytes or est
                active.append(req)
            if not active:
                if pending:
                    self.sleep(max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.01)))
                continue
            max_conc = max(max_conc, len(active))
            peak_bytes = max(peak_bytes, in_use)
            self.max_in_use = peak_bytes
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
                log(f"serving t={t_end:.1f}s active={len(active)} waiting={len(waiting)} done={len(done)}")
        wall = self.clock() - t0
        for req in active:  # stopped at the horizon: release unfinished jobs
            self.ex.retire(req)
        lat = [r.finished - r.arrival for r in done]
        qd = [r.admitted - r.arrival for r in done]
        pct = lambda a, q: float(np.percentile(a, q)) if a else None  # noqa: E731
        return ServingReport(
            completed=len(done), submitted=len(requests), vertex_substeps=vs, wall_s=wall,
            throughput=vs / wall if wall > 0 else 0.0, p50_latency_s=pct(lat, 50), p99_latency_s=pct(lat, 99),
            p50_queue_s=pct(qd, 50), max_concurrency=max_conc, peak_bytes=int(peak_bytes),
            substep_launches=launches, latencies=lat)
