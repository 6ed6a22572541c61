"""GPU parity at the sizes of the benchmark workloads themselves (the small
stand-ins of the other test files cover the code paths; these cover the
shapes):

  256 x 256 single cloth, E seeded in [1e4, 1e6] Pa: 20 substeps;
  the full batch of 512 grids with sides log-uniform in [16, 128] (seed 2,
  2.08 M vertices), one fused substep for all: 2 substeps;
  a window of the serving stream (200 jobs/s, meshes log-uniform in
  [1e3, 2.5e5] vertices): served results equal the same jobs stepped alone.

Every substep starts from the FP32 oracle's state (positions, velocities
and the PCG warm start) and must land within 1e-5 relative max-norm of the
oracle's next state, with PCG iteration counts within +-1 (the bar of the
32 x 32 workload in test_gpu_step.py)."""
import time

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _load_state(dc, o):
    import torch

    po, vo = o.state()
    dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
    dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
    dc.sim.debug_set_warm_start(o.dx().astype(np.float32))


def test_plain_256x256_matches_oracle(gpu):
    from chysx import scenes

    spec = scenes.plain_large()
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    print(f"256x256: {spec.n} vertices, E = {spec.meta['E']:.4g} Pa", flush=True)
    worst, worst_it = 0.0, 0
    for s in range(20):
        _load_state(dc, o)
        dc.step()
        it_o = o.step(spec.dt)
        err = helpers.rel_maxnorm(dc.state()[0], o.state()[0])
        dit = abs(dc.sim.last_pcg_iterations() - it_o)
        worst, worst_it = max(worst, err), max(worst_it, dit)
        assert err <= 1e-5, (s, err)
        assert dit <= 1, (s, dc.sim.last_pcg_iterations(), it_o)
    print(f"256x256: worst one-substep rel err {worst:.3e}, worst |d it| {worst_it}", flush=True)


def test_full_batch_512_matches_oracle(gpu):
    """The batched workload in full: 512 instances in one ClothBatch (one
    element pass, one gather pass, one persistent PCG launch), two substeps,
    every instance against its own FP32 oracle."""
    import torch

    import chysx
    import pyoracle
    from chysx import scenes

    t0 = time.perf_counter()
    specs = scenes.batched(seed=2, count=512)
    cloths = [scenes.DeviceCloth(s) for s in specs]
    batch = chysx.ClothBatch()
    for c in cloths:
        batch.add_instance(c.sim)
    batch.finalize(specs[0].dt)
    oracles = [helpers.oracle_from_spec(s) for s in specs]
    n_all = sum(s.n for s in specs)
    print(f"512 instances, {n_all} vertices, set up in {time.perf_counter() - t0:.1f} s "
          f"({pyoracle.num_threads()} oracle threads)", flush=True)
    assert n_all == 2076631
    stream = torch.cuda.Stream()
    for s in range(2):
        for c, o in zip(cloths, oracles):
            _load_state(c, o)
        t1 = time.perf_counter()
        with torch.cuda.stream(stream):
            batch.step(specs[0].dt, stream.cuda_stream)
        its = batch.last_pcg_iterations(stream.cuda_stream)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        its_o = [o.step(sp.dt) for o, sp in zip(oracles, specs)]
        t3 = time.perf_counter()
        errs = np.array([helpers.rel_maxnorm(c.state()[0], o.state()[0]) for c, o in zip(cloths, oracles)])
        dits = np.abs(np.array(its) - np.array(its_o))
        print(f"substep {s}: GPU {1e3 * (t2 - t1):.1f} ms, oracle {t3 - t2:.1f} s; worst rel err {errs.max():.3e} "
              f"(instance {int(errs.argmax())}), worst |d it| {dits.max()}, capped {int(np.sum(np.array(its) == 100))}",
              flush=True)
        assert errs.max() <= 1e-5, (s, int(errs.argmax()), errs.max())
        assert dits.max() <= 1, (s, int(dits.argmax()), its[int(dits.argmax())], its_o[int(dits.argmax())])


def test_serving_stream_window_equals_alone(gpu):
    """Two seconds of the serving stream at its configured rate (200 jobs/s,
    meshes log-uniform in [1e3, 2.5e5] vertices, materials as the batched
    workload), substeps cut to 5-12 per job so that the window drains: every
    served job's state equals the same job stepped alone, bit for bit, for a
    seeded sample of 12 jobs (including the largest)."""
    from dataclasses import replace

    from chysx import scenes, serving

    jobs = scenes.serving_jobs(seed=4, rate=200.0, duration=2.0)
    rng = np.random.default_rng(4)
    reqs = [serving.Request(job_id=i, arrival=j.arrival, spec=j.spec, substeps=int(rng.integers(5, 13)))
            for i, j in enumerate(jobs)]
    big = max(range(len(reqs)), key=lambda i: reqs[i].spec.n)
    sample = sorted(set(rng.choice(len(reqs), size=11, replace=False).tolist()) | {big})
    keep = set(sample)
    ex = serving.GpuExecutor(lanes_per_row=4, keep_results=True, prepare_workers=8,
                             keep_only=lambda r: r.job_id in keep)
    dt = scenes.ClothSpec.dt
    try:
        rep = serving.ServingEngine(ex, hbm_budget_bytes=200e9).run(reqs, dt)
    finally:
        ex.close()
    print({k: v for k, v in rep.as_dict().items() if not isinstance(v, list)}, flush=True)
    assert rep.completed == len(reqs) and rep.unfinished == 0
    assert rep.max_concurrency >= 8
    for i in sample:
        r = reqs[i]
        spec = r.spec.build()
        alone = scenes.DeviceCloth(replace(spec), lanes_per_row=4)
        for _ in range(r.substeps):
            alone.step(dt)
        p, v = alone.state()
        assert np.array_equal(p, r.result[0]) and np.array_equal(v, r.result[1]), (i, spec.n)
