"""Continuous-batching server on the GPU: jobs that arrive while others run
join the batch at a substep boundary (their set-up runs on worker threads),
and every job's result is the same bits as the job stepped alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone(gpu):
    from chysx import scenes, serving

    specs = scenes.batched(seed=31, count=6, lo=8, hi=40)
    dt = specs[0].dt
    # Job 0 runs 400 substeps (each one is at least a launch and a
    # synchronisation, > 50 us); the others arrive 5 ms apart while it runs.
    reqs = [serving.Request(job_id=i, arrival=0.005 * i, spec=s, substeps=400 if i == 0 else 10 + 4 * i)
            for i, s in enumerate(specs)]
    ex = serving.GpuExecutor(lanes_per_row=8, keep_results=True, prepare_workers=2)
    try:
        rep = serving.ServingEngine(ex).run(reqs, dt)
    finally:
        ex.close()
    print(rep.as_dict(), flush=True)
    assert rep.completed == len(reqs) and rep.unfinished == 0
    assert rep.max_concurrency >= 2  # the jobs overlapped on the device
    for r in reqs:
        assert r.done == r.substeps
        alone = scenes.DeviceCloth(r.spec, lanes_per_row=8)
        for _ in range(r.substeps):
            alone.step(dt)
        p, v = alone.state()
        assert np.array_equal(p, r.result[0]) and np.array_equal(v, r.result[1]), r.job_id


def test_budget_limits_concurrency(gpu):
    """An HBM budget of about two of these cloths keeps at most two in
    flight; all still complete, first come first served."""
    from chysx import scenes, serving

    specs = scenes.batched(seed=32, count=5, lo=20, hi=21)
    dt = specs[0].dt
    reqs = [serving.Request(job_id=i, arrival=0.0, spec=s, substeps=3) for i, s in enumerate(specs)]
    budget = 2.5 * serving.BYTES_PER_VERTEX_ESTIMATE * specs[0].n
    ex = serving.GpuExecutor(prepare_workers=1)
    try:
        rep = serving.ServingEngine(ex, hbm_budget_bytes=budget).run(reqs, dt)
    finally:
        ex.close()
    assert rep.completed == 5 and rep.max_concurrency <= 2
    order = sorted(reqs, key=lambda r: r.admitted)
    assert [r.job_id for r in order] == list(range(5))
