"""Cloth vertex-substeps/s of the batched implicit cloth step on one MI355X.

Workload (`batched`, seed 2): 512 independent cloth grids, sides drawn
log-uniform in [16, 128] vertices, E ~ U[1e4, 1e6] Pa (BW98 k = E*h,
h = 1 mm), density ~ U[0.1, 0.4] kg/m^2, two random corners pinned,
bending 1e-4, dt = 1/600 s, block-Jacobi PCG to relative residual 1e-6
(cap 100), FP32 with FP64 reductions.  One step = one substep of every
instance (one assembly kernel + one persistent PCG launch).

value        = sum(N_i) * K / wall-seconds of K timed substeps (inputs in HBM)
plain        = the same instances stepped one ClothSimulator at a time
roofline     = persistent PCG kernel: algorithmic bytes per launch
               sum_i [B_setup_i + k_i B_iter_i + 48 N_i] / mean launch time
               (HIP events on the stream the kernel runs on)
cpu_baseline = the CPU oracle (a port of the reference step) on a bounded
               sample of the same instances, all host threads
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "newton-chysx_amd"))

METRIC = "cloth vertex-substeps/s (batched vs plain) at 1e-6 PCG tol; p99 req latency"
HBM_PEAK_GBS = 8000.0


def b_spmv(n, nnz):
    return 40 * nnz + 64 * n


def b_iter(n, nnz):
    return b_spmv(n, nnz) + 156 * n


def b_setup(n, nnz):
    return b_spmv(n, nnz) + 168 * n


def kernel_breakdown(cloth, dt, stream, lanes, reps=5, pcg_form=0, tuned=None):
    """Per-kernel times and algorithmic rates of one instance's substep:
    the assembly kernel (element forces and Hessians: x_n 12N + triangle 52T
    + dihedral 20D + pin 16P bytes; assembly: predictor 56N + rhs 28N +
    diagonal 36N + blocks 36 nnz), and the persistent PCG launch, split by
    its solver trace: in the
    single-barrier form (0) into iteration phases (B_iter each); in the
    two-barrier form (1) into the SpMV phase (B_spmv + 36N per iteration: the
    folded p update) and the vector / reduction phase (120N per iteration);
    plus the whole launch."""
    import numpy as np
    import torch

    import chysx

    sh = stream.cuda_stream
    sim = cloth.sim
    N, nnz = sim.num_particles(), sim.num_off_diag_blocks()
    T, D, P = sim.num_fem_stretch_triangles(), sim.num_bending_dihedrals(), sim.num_pins()
    b = chysx.ClothBatch()
    b.set_lanes_per_row(lanes)
    if tuned is not None:
        pcg_form = tuned["pcg_form"]
        b.set_lds_staging(tuned["lds_staging"])
        b.set_tile_rows(tuned["tile_rows"])
        b.set_grid_blocks(tuned["grid_blocks"])
    b.set_pcg_form(pcg_form)
    b.add_instance(sim)
    b.set_trace(1024)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    te, ts, ta, tb, ks = [], [], [], [], []
    with torch.cuda.stream(stream):
        b.step(dt, sh)
        for _ in range(reps):
            ev[0].record(stream)
            b.assemble(dt, sh)
            ev[1].record(stream)
            b.solve(sh)
            ev[2].record(stream)
            ev[2].synchronize()
            te.append(ev[0].elapsed_time(ev[1]))
            ts.append(ev[1].elapsed_time(ev[2]))
            tr = np.array(b.last_trace(sh)).astype(np.int64)
            d = np.diff(tr[:, 0]) / 100.0  # us (100 MHz stamps)
            if pcg_form == 0:
                ta.append(float(d.sum()))
                tb.append(0.0)
            else:
                ta.append(float(d[0::2].sum()))
                tb.append(float(d[1::2].sum()))
            ks.append(int(b.last_pcg_iterations(sh)[0]))
    b.set_trace(0)
    k = float(np.mean(ks))
    rate = lambda nbytes, ms: round(nbytes / (ms * 1e-3) / 1e9, 1)
    e_ms, s_ms = float(np.median(te)), float(np.median(ts))
    a_us, bv_us = float(np.median(ta)), float(np.median(tb))
    e_b = 12 * N + 52 * T + 20 * D + 16 * P + 120 * N + 36 * nnz
    a_b = k * (b_spmv(N, nnz) + 36 * N)
    v_b = k * 120 * N
    s_b = b_setup(N, nnz) + k * b_iter(N, nnz) + 48 * N
    out = {
        "vertices": N, "off_diag_blocks": nnz, "triangles": T, "dihedrals": D, "pins": P, "pcg_iterations": k,
        "assembly": {"ms": round(e_ms, 4), "bytes": e_b, "GB_per_s": rate(e_b, e_ms)},
        "pcg_launch": {"ms": round(s_ms, 4), "bytes": s_b, "GB_per_s": rate(s_b, s_ms)},
        "pcg_form": pcg_form,
        "note": "algorithmic bytes (reference counts); phase times from the solver's per-barrier trace, barrier "
                "waits included; a 256x256 working set fits the 256 MB Infinity Cache",
    }
    if pcg_form == 0:
        i_b = k * b_iter(N, nnz)
        out["pcg_iteration_phase"] = {"us_per_iteration": round(a_us / max(k, 1.0), 2), "bytes": i_b,
                                      "GB_per_s": rate(i_b, a_us * 1e-3)}
    else:
        out["pcg_spmv_phase"] = {"us_per_iteration": round(a_us / max(k, 1.0), 2), "bytes": a_b,
                                 "GB_per_s": rate(a_b, a_us * 1e-3)}
        out["pcg_vector_phase"] = {"us_per_iteration": round(bv_us / max(k, 1.0), 2), "bytes": v_b,
                                   "GB_per_s": rate(v_b, bv_us * 1e-3)}
    return out


def load_traffic():
    """HBM bytes per PCG launch from the committed rocprofv3 --pmc summary."""
    path = os.path.join(REPO, "profiles", "pcg_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(specs, budget_s=12.0, budget_1core_s=6.0, reps=5):
    """The CPU oracle (FP32 restatement of the reference step, OpenMP over
    rows and elements) on a bounded sample of the same instances: one
    warm-up substep per instance, then `reps` timed repetitions, each one
    substep of every sampled instance; value = the median repetition's
    vertex-substeps/s.  First on every host thread over as many instances
    as fit the budget, then on one thread over a smaller sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import helpers
    import pyoracle

    def sample(budget, threads):
        pyoracle.set_num_threads(threads)
        oracles, t_one = [], 0.0
        for spec in specs:  # warm-up substep per instance sizes the sample
            o = helpers.oracle_from_spec(spec, precision=32)
            s, _ = o.run(spec.dt, 1)
            oracles.append((spec, o))
            t_one += s
            if t_one * reps >= budget:
                break
        rates, iters = [], []
        for _ in range(reps):
            secs, vs = 0.0, 0
            for spec, o in oracles:
                s, _ = o.run(spec.dt, 1)
                iters.append(o.last_stats()[0])
                secs += s
                vs += spec.n
            rates.append(vs / secs)
        return float(np.median(rates)), len(oracles), t_one, float(np.mean(iters))

    import numpy as np

    threads = pyoracle.num_threads()
    v_all, used, secs, it_all = sample(budget_s, threads)
    v_one, used1, secs1, _ = sample(budget_1core_s, 1)
    pyoracle.set_num_threads(threads)
    return {
        "value": v_all,
        "unit": "vertex-substeps/s",
        "cores": threads,
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "value_1core": v_one,
        "pcg_iterations_mean": it_all,
        "sample": f"{used} of {len(specs)} batched instances (seeded order): 1 warm-up substep each, then the median "
                  f"of {reps} repetitions of one substep of every sampled instance ({secs:.1f} s per repetition), FP32 "
                  f"oracle on {threads} threads; 1-thread figure: the same over the first {used1} instances "
                  f"({secs1:.1f} s per repetition)",
    }


def fp64_error_sample(specs, lanes, stream, count=8, steps=20, seed=7):
    """Position error of the batched GPU step against the full-FP64 oracle:
    a seeded sample of the batched workload's instances, fresh from their
    initial state, advanced `steps` substeps in one ClothBatch and by the
    FP64 oracle (and the FP32 oracle, the reference algorithm's own FP32
    distance from FP64); relative max-norm per instance."""
    import numpy as np

    import chysx
    from chysx import scenes

    sys.path.insert(0, os.path.join(REPO, "oracle"))
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import helpers

    rng = np.random.default_rng(seed)
    idx = sorted(rng.choice(len(specs), size=count, replace=False).tolist())
    sub = [specs[i] for i in idx]
    cloths = [scenes.DeviceCloth(s, lanes_per_row=lanes) for s in sub]
    b = chysx.ClothBatch()
    for c in cloths:
        b.add_instance(c.sim)
    with __import__("torch").cuda.stream(stream):
        for _ in range(steps):
            b.step(sub[0].dt, stream.cuda_stream)
    stream.synchronize()
    err, err32, its = [], [], []
    for s, c in zip(sub, cloths):
        o64 = helpers.oracle_from_spec(s, precision=64)
        o32 = helpers.oracle_from_spec(s, precision=32)
        for _ in range(steps):
            its.append(o64.step(s.dt))
            o32.step(s.dt)
        p64 = o64.state()[0]
        err.append(helpers.rel_maxnorm(c.state()[0], p64))
        err32.append(helpers.rel_maxnorm(o32.state()[0], p64))
    del b, cloths
    return {"instances": idx, "vertices": [s.n for s in sub], "substeps": steps,
            "rel_maxnorm_error_vs_fp64": {"max": float(np.max(err)), "median": float(np.median(err)),
                                          "per_instance": [float(e) for e in err]},
            "fp32_oracle_error_vs_fp64": {"max": float(np.max(err32)), "median": float(np.median(err32))},
            "note": "fresh instances of the batched workload, PCG tol 1e-6 cap 100, free-running"}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_serving(rate, duration, horizon, lanes, grid, tile_rows, workers=4, seed=4, lds_staging=True,
                max_queue_s=None):
    """Continuous-batching server on a Poisson stream of cloth-drape jobs
    (mesh sizes log-uniform in [1e3, 2.5e5] vertices, 30-300 substeps,
    materials as in the batched workload), arrivals for `duration` s, the
    server stopped at `horizon` s of wall time; HBM budget 200 GB."""
    from chysx import scenes, serving

    jobs = scenes.serving_jobs(seed=seed, rate=rate, duration=duration)
    reqs = [serving.Request(job_id=i, arrival=j.arrival, spec=j.spec, substeps=j.substeps) for i, j in enumerate(jobs)]
    offered = sum(j.spec.n * j.substeps for j in jobs) / duration
    ex = serving.GpuExecutor(lanes_per_row=lanes, grid_blocks=grid or None, tile_rows=tile_rows,
                             prepare_workers=workers, lds_staging=lds_staging)
    dt = scenes.ClothSpec.dt
    try:
        rep = serving.ServingEngine(ex, hbm_budget_bytes=200e9, max_queue_s=max_queue_s).run(
            reqs, dt, horizon_s=horizon, log=log)
    finally:
        ex.close()
    out = rep.as_dict()
    out.update({"rate_per_s": rate, "arrival_window_s": duration, "horizon_s": horizon,
                "offered_vertex_substeps_per_s": offered, "prepare_workers": workers, "max_queue_s": max_queue_s,
                "latency_unit": "s (arrival to end of the job's last substep); *_censored: every arrived, not "
                                "rejected job, unfinished ones counted at the end of the run (lower bounds)"})
    log(f"serving rate={rate}/s: {rep.completed}/{rep.arrived} arrived jobs done, {rep.rejected} rejected, "
        f"{rep.throughput:.3e} vertex-substeps/s, p50/p99 completed {rep.p50_latency_s}/{rep.p99_latency_s} s, "
        f"censored {rep.p50_latency_censored_s}/{rep.p99_latency_censored_s} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the lanes-per-row x grid-size sweep and use --lanes / --grid")
    ap.add_argument("--grid", type=int, default=0, help="persistent PCG workgroups (0: all co-resident)")
    ap.add_argument("--tile-rows", type=int, default=256, help="rows per PCG tile (multiple of 256)")
    ap.add_argument("--no-lds-staging", action="store_true",
                    help="PCG SpMV gathers from global memory instead of LDS-staged stages (with --no-autotune)")
    ap.add_argument("--no-plain-large", action="store_true")
    ap.add_argument("--workload", choices=["batched", "serving"], default="batched",
                    help="batched (default; includes a short serving leg) or the full serving stream")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg of the batched workload")
    ap.add_argument("--no-bf16", action="store_true", help="skip the BF16-blocks leg of the batched workload")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager substeps instead of hipGraph replays of the captured substep")
    ap.add_argument("--serving-rate", type=float, default=None)
    ap.add_argument("--serving-duration", type=float, default=None)
    ap.add_argument("--pcg-form", type=int, default=0, choices=[0, 1],
                    help="FP32 persistent PCG: 0 one grid barrier per iteration, 1 two")
    ap.add_argument("--no-converged", action="store_true", help="skip the leg with a cap every instance converges under")
    ap.add_argument("--converged-cap", type=int, default=4000)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-oracle error sample")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run an elementwise copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    import chysx
    from chysx import scenes

    if args.workload == "serving":
        # The serving workload: 200 req/s for 60 s, the server stopped at 60 s.
        rate = args.serving_rate or 200.0
        dur = args.serving_duration or 60.0
        sv = run_serving(rate, dur, dur, args.lanes, args.grid, args.tile_rows, lds_staging=not args.no_lds_staging)
        print(json.dumps({
            "metric": "served cloth vertex-substeps/s and p99 request latency (continuous batching)",
            "value": sv["throughput"], "unit": "vertex-substeps/s", "n_gpus": 1, "higher_is_better": True,
            "p50_latency_s": sv["p50_latency_censored_s"], "p99_latency_s": sv["p99_latency_censored_s"],
            "p99_latency_completed_s": sv["p99_latency_s"], "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": f"serving: Poisson {rate:g} req/s for {dur:g} s, meshes logU[1e3,2.5e5] vertices, "
                                   "30-300 substeps, HBM budget 200 GB, PCG tol 1e-6 cap 100"},
            "serving": sv}))
        return

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    specs = scenes.batched(seed=2, count=args.instances)
    t0 = time.perf_counter()
    cloths = [scenes.DeviceCloth(s, lanes_per_row=args.lanes) for s in specs]
    batch = chysx.ClothBatch()
    batch.set_lanes_per_row(args.lanes)
    if args.grid > 0:
        batch.set_grid_blocks(args.grid)
    batch.set_tile_rows(args.tile_rows)
    batch.set_lds_staging(not args.no_lds_staging)
    batch.set_pcg_form(args.pcg_form)
    for c in cloths:
        batch.add_instance(c.sim)
    batch.finalize(specs[0].dt)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"set up {len(specs)} instances in {setup_s:.1f} s")
    sizes = np.array(batch.instance_sizes(), np.int64)
    n_rows = int(sizes[:, 0].sum())
    dt = specs[0].dt
    tuned = None
    if not args.no_autotune:
        # Lanes per block row x persistent grid size, picked on this batch
        # (untimed; a few dozen substeps).
        from chysx import autotune

        tuned = autotune.tune_batch(batch, dt, stream, cache_path=None, log=log)
        args.lanes = tuned["lanes_per_row"]
        args.grid = tuned["grid_blocks"]
        args.tile_rows = tuned["tile_rows"]
        args.no_lds_staging = not tuned["lds_staging"]
        log(f"autotune picked staging={int(tuned['lds_staging'])} lanes={tuned['lanes_per_row']} "
            f"grid={tuned['grid_blocks']} tile_rows={tuned['tile_rows']}")

    # The timed substeps replay one hipGraph (assembly kernel, persistent
    # PCG + finalize), captured during the warm-up.
    batch.set_graph(not args.no_graph)
    with torch.cuda.stream(stream):
        for _ in range(max(1, args.warmup)):
            batch.step(dt, sh)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.step(dt, sh)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    graph_captures = batch.graph_captures()
    batch.set_graph(False)
    log(f"timed {args.steps} batched substeps ({'hipGraph replays' if not args.no_graph else 'eager'}): "
        f"{wall / args.steps * 1e3:.2f} ms each")
    iters = np.array(batch.last_pcg_iterations(sh))
    res = np.array(batch.last_relative_residuals(sh))
    value = n_rows * args.steps / wall
    capped = iters >= 100
    convergence = {
        "fraction_at_tol": float(np.mean(res <= 1e-6)),
        "instances_at_cap": int(capped.sum()),
        "rows_at_cap": int(sizes[capped, 0].sum()),
        "capped_relative_residual": ({"median": float(np.median(res[capped])), "max": float(res[capped].max())}
                                     if capped.any() else None),
        "note": "last timed substep: instances whose relative residual met 1e-6, and the residuals of those "
                "stopped by the 100-iteration cap"}
    log(f"convergence: {convergence['fraction_at_tol']:.3f} of instances at 1e-6, {convergence['instances_at_cap']} "
        f"at the cap")

    # PCG kernel alone: assemble, then time the solve launch with events on `stream`.
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    pcg_ms, pcg_bytes = [], []
    with torch.cuda.stream(stream):
        for _ in range(max(3, min(args.steps, 10))):
            batch.assemble(dt, sh)
            ev0.record(stream)
            batch.solve(sh)
            ev1.record(stream)
            ev1.synchronize()
            pcg_ms.append(ev0.elapsed_time(ev1))
            k = np.array(batch.last_pcg_iterations(sh), np.int64)
            n, nnz = sizes[:, 0], sizes[:, 1]
            pcg_bytes.append(float(np.sum(b_setup(n, nnz) + k * b_iter(n, nnz) + 48 * n)))
    mean_ms = float(np.mean(pcg_ms))
    achieved = float(np.mean(pcg_bytes)) / (mean_ms * 1e-3) / 1e9
    traffic = load_traffic()
    entries16 = bool(batch.entries16())
    key = (f"{'pcg1_' if args.pcg_form == 0 else ''}{'global_' if args.no_lds_staging else ''}"
           f"{'' if entries16 else 'entries32_'}"
           f"lanes{args.lanes}_grid{args.grid}_tile{args.tile_rows}_inst{len(specs)}")
    traffic_bytes = None
    if traffic and key in traffic.get("configs", {}):
        traffic_bytes = traffic["configs"][key]["bytes_per_launch"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_bytes,
                "kernel": "%s<%s, %s sliced entries> (persistent block-Jacobi PCG + finalize, %s)"
                          % ("pcg1_kernel" if args.pcg_form == 0 else "pcg_kernel",
                             "global SpMV" if args.no_lds_staging else "LDS-staged SpMV",
                             "16-bit" if entries16 else "32-bit",
                             "one grid barrier per iteration" if args.pcg_form == 0 else
                             "two grid barriers per iteration"),
                "mean_launch_ms": round(mean_ms, 3), "algorithmic_bytes_per_launch": float(np.mean(pcg_bytes))}

    log(f"PCG launch {mean_ms:.3f} ms, {achieved:.0f} GB/s algorithmic")

    # The phases of iterations in which every instance is solving, from the
    # solver's per-barrier trace of one more solve.  Single-barrier form: a
    # whole iteration per phase, algorithmic bytes B_iter.  Two-barrier form:
    # the SpMV phase (q = A p with the folded p update), B_spmv + 36N.
    spmv_phase = iteration_phase = None
    batch.set_trace(1024)
    with torch.cuda.stream(stream):
        batch.assemble(dt, sh)
        batch.solve(sh)
        tr = np.array(batch.last_trace(sh)).astype(np.int64)
    batch.set_trace(0)
    n_all, nnz_all = int(sizes[:, 0].sum()), int(sizes[:, 1].sum())
    if args.pcg_form == 0:
        full = [k for k in range(1, len(tr) - 1) if tr[k + 1, 1] == len(specs)]
        if full:
            us = np.array([(tr[k + 1, 0] - tr[k, 0]) / 100.0 for k in full])
            ib = b_iter(n_all, nnz_all)
            gbs = ib / (float(us.mean()) * 1e-6) / 1e9
            iteration_phase = {"iterations": len(full), "us_per_iteration": round(float(us.mean()), 1),
                               "algorithmic_bytes": ib, "GB_per_s": round(gbs, 1),
                               "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
                               "note": "iterations with every instance solving (SpMV, vector updates and "
                                       "reductions of one iteration, one barrier); barrier wait included"}
            log(f"iteration phase at full activity: {iteration_phase['us_per_iteration']} us, "
                f"{iteration_phase['GB_per_s']} GB/s algorithmic ({iteration_phase['frac_of_peak']:.3f} of peak)")
        full = []
    else:
        full = [k for k in range((len(tr) - 1) // 2) if tr[2 * k + 1, 1] == len(specs)]
    if full:
        us = np.array([(tr[2 * k + 1, 0] - tr[2 * k, 0]) / 100.0 for k in full])
        phase_bytes = b_spmv(n_all, nnz_all) + 36 * n_all
        spmv_phase = {"iterations": len(full), "us_per_iteration": round(float(us.mean()), 1),
                      "algorithmic_bytes": phase_bytes,
                      "GB_per_s": round(phase_bytes / (float(us.mean()) * 1e-6) / 1e9, 1),
                      "frac_of_peak": round(phase_bytes / (float(us.mean()) * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                      "note": "iterations with every instance solving; barrier wait included"}
        log(f"SpMV phase at full activity: {spmv_phase['us_per_iteration']} us, {spmv_phase['GB_per_s']} GB/s "
            f"algorithmic ({spmv_phase['frac_of_peak']:.3f} of peak)")

    # The other PCG form on the same batch (same substeps, same results up to
    # rounding): hipGraph replays, a few substeps each.
    form_cmp = {}
    for form in (args.pcg_form, 1 - args.pcg_form):
        batch.set_pcg_form(form)
        batch.set_graph(not args.no_graph)
        with torch.cuda.stream(stream):
            batch.step(dt, sh)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                batch.step(dt, sh)
            torch.cuda.synchronize()
            form_cmp[f"form{form}_ms_per_step"] = round((time.perf_counter() - t0) / 5 * 1e3, 3)
        batch.set_graph(False)
    batch.set_pcg_form(args.pcg_form)
    log(f"PCG forms: {form_cmp}")

    bf16_leg = None
    if not args.no_bf16:
        # The same batch with BF16-stored off-diagonal blocks (compensated
        # FP32 diagonal, one outer refinement against the FP32 residual):
        # same substeps, same tolerance and cap.
        for c in cloths:
            c.sim.set_bf16_blocks(True)
        batch.set_bf16_blocks(True)
        batch.set_graph(not args.no_graph)
        with torch.cuda.stream(stream):
            for _ in range(max(1, args.warmup)):
                batch.step(dt, sh)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                batch.step(dt, sh)
            torch.cuda.synchronize()
            bw16 = time.perf_counter() - t0
        it16 = np.array(batch.last_pcg_iterations(sh))
        batch.set_graph(False)
        bf16_leg = {"value": n_rows * args.steps / bw16, "unit": "vertex-substeps/s",
                    "ms_per_step": bw16 / args.steps * 1e3,
                    "pcg_iterations": {"mean": float(it16.mean()), "max": int(it16.max())},
                    "vs_fp32": round((n_rows * args.steps / bw16) / value, 3)}
        log(f"BF16 blocks: {bf16_leg['ms_per_step']:.2f} ms per substep, {bf16_leg['vs_fp32']}x the FP32 rate")
        batch.set_bf16_blocks(False)
        for c in cloths:
            c.sim.set_bf16_blocks(False)
    converged_leg = None
    if not args.no_converged:
        # The same batch with a cap high enough for every instance to reach
        # the 1e-6 tolerance.
        for c in cloths:
            c.sim.set_pcg_iterations(args.converged_cap)
        batch.set_graph(not args.no_graph)
        with torch.cuda.stream(stream):
            for _ in range(max(1, args.warmup)):
                batch.step(dt, sh)
            torch.cuda.synchronize()
            ksteps = max(3, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(ksteps):
                batch.step(dt, sh)
            torch.cuda.synchronize()
            wc = time.perf_counter() - t0
        itc = np.array(batch.last_pcg_iterations(sh))
        resc = np.array(batch.last_relative_residuals(sh))
        batch.set_graph(False)
        converged_leg = {"value": n_rows * ksteps / wc, "unit": "vertex-substeps/s", "steps": ksteps,
                         "ms_per_step": wc / ksteps * 1e3, "pcg_cap": args.converged_cap,
                         "pcg_iterations": {"mean": float(itc.mean()), "max": int(itc.max())},
                         "fraction_at_tol": float(np.mean(resc <= 1e-6)),
                         "max_relative_residual": float(resc.max())}
        log(f"converged leg (cap {args.converged_cap}): {converged_leg['ms_per_step']:.1f} ms per substep, "
            f"{converged_leg['fraction_at_tol']:.3f} at 1e-6, iterations max {converged_leg['pcg_iterations']['max']}")
        for c in cloths:
            c.sim.set_pcg_iterations(100)
    fp64 = None
    if not args.no_fp64:
        fp64 = fp64_error_sample(specs, args.lanes, stream)
        log(f"FP64 error sample: max {fp64['rel_maxnorm_error_vs_fp64']['max']:.3e} after {fp64['substeps']} "
            f"substeps (FP32 oracle {fp64['fp32_oracle_error_vs_fp64']['max']:.3e})")
    copy_probe = None
    if args.copy_probe_mib > 0:
        # A plain streaming kernel of known size (read B, write B bytes):
        # torch's elementwise multiply by one, timed with events on `stream`.
        nbytes = args.copy_probe_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        with torch.cuda.stream(stream):
            torch.mul(src, 1.0, out=dst)
            ev0.record(stream)
            for _ in range(5):
                torch.mul(src, 1.0, out=dst)
            ev1.record(stream)
            ev1.synchronize()
        ms = ev0.elapsed_time(ev1) / 5
        copy_probe = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(ms, 4),
                      "GB_per_s": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}
        log(f"copy probe {args.copy_probe_mib} MiB: {copy_probe['GB_per_s']} GB/s")
        del src, dst
    plain = None
    if not args.no_plain:
        # The same instances, one ClothSimulator.step at a time (each packs
        # its own single-instance system on first use, untimed).
        with torch.cuda.stream(stream):
            for c in cloths:
                c.sim.step(dt, sh)
            torch.cuda.synchronize()
            reps = 2
            t0 = time.perf_counter()
            for _ in range(reps):
                for c in cloths:
                    c.sim.step(dt, sh)
            torch.cuda.synchronize()
            pw = time.perf_counter() - t0
        plain = {"value": n_rows * reps / pw, "unit": "vertex-substeps/s", "ms_per_substep_all": pw / reps * 1e3,
                 "method": "sequential ClothSimulator.step over the same instances, 2 timed substeps"}

    if plain:
        log(f"plain sequential: {plain['value']:.3e} vertex-substeps/s")
    plain_large = None
    if not args.no_plain_large:
        # One large instance (256x256 grid, E ~ U[1e4, 1e6], seed 1) through
        # ClothSimulator.step: the single-instance path at full occupancy.
        big = scenes.DeviceCloth(scenes.plain_large(), lanes_per_row=args.lanes)
        tuned_big = None
        if not args.no_autotune:
            from chysx import autotune

            tuned_big = autotune.tune_simulator(big.sim, dt, stream, log=log)
            log(f"single-cloth autotune picked staging={int(tuned_big['lds_staging'])} "
                f"tile_rows={tuned_big['tile_rows']} form={tuned_big['pcg_form']} grid={tuned_big['grid_blocks']}")
            big = scenes.DeviceCloth(scenes.plain_large(), lanes_per_row=args.lanes)
            big.sim.set_lds_staging(tuned_big["lds_staging"])
            big.sim.set_tile_rows(tuned_big["tile_rows"])
            big.sim.set_pcg_form(tuned_big["pcg_form"])
            big.sim.set_grid_blocks(tuned_big["grid_blocks"])
        with torch.cuda.stream(stream):
            for _ in range(3):
                big.sim.step(dt, sh)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = 10
            for _ in range(reps):
                big.sim.step(dt, sh)
            torch.cuda.synchronize()
            bw = time.perf_counter() - t0
        plain_large = {"value": big.spec.n * reps / bw, "unit": "vertex-substeps/s", "vertices": big.spec.n,
                       "ms_per_substep": bw / reps * 1e3, "pcg_iterations": big.sim.last_pcg_iterations(sh)}
        log(f"plain 256x256: {plain_large['ms_per_substep']:.2f} ms/substep")
        if tuned_big is not None:
            plain_large["autotune"] = tuned_big
        plain_large["kernels"] = kernel_breakdown(big, dt, stream, args.lanes, pcg_form=args.pcg_form,
                                                  tuned=tuned_big)
        kb = plain_large["kernels"]
        log("plain 256x256 kernels: " + ", ".join(
            f"{name} {kb[name]['GB_per_s']} GB/s" for name in
            ("assembly", "pcg_iteration_phase", "pcg_spmv_phase", "pcg_vector_phase",
             "pcg_launch") if name in kb))
    serving_leg = serving_sustainable = None
    if not args.no_serving:
        # A window of the serving workload at its configured rate (200 jobs/s):
        # arrivals for 10 s, the server stopped at 20 s; jobs not finished by
        # then count as censored at the horizon.
        serving_leg = run_serving(args.serving_rate or 200.0, args.serving_duration or 10.0, 20.0, args.lanes,
                                  args.grid, args.tile_rows, workers=8, lds_staging=not args.no_lds_staging)
        # A sustainable stream (about half the batched rate) from the same
        # distribution: arrivals for 15 s, drained within 60 s.
        serving_sustainable = run_serving(4.0, 15.0, 60.0, args.lanes, args.grid, args.tile_rows,
                                          lds_staging=not args.no_lds_staging)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(specs, args.cpu_budget)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "vertex-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched: 512 cloth grids, side logU[16,128] vertices (seed 2), BW98 + bending, "
                               "PCG tol 1e-6 cap 100, one fused substep per step"
                               + ("" if args.no_graph else ", replayed from one hipGraph"),
                   "hipgraph": not args.no_graph, "graph_captures": graph_captures,
                   "instances": len(specs), "vertices": n_rows, "off_diag_blocks": int(sizes[:, 1].sum()),
                   "pcg_tol": 1e-6, "pcg_cap": 100, "lanes_per_row": args.lanes, "grid_blocks": args.grid, "tile_rows": args.tile_rows,
                   "lds_staging": not args.no_lds_staging, "sliced_entries": "16-bit" if entries16 else "32-bit",
                   "parallelism": "single GPU"},
        "pcg_iterations": {"mean": float(iters.mean()), "max": int(iters.max()), "min": int(iters.min())},
        "max_relative_residual": float(res.max()),
        "convergence": convergence,
        "converged_leg": converged_leg,
        "fp64_error_sample": fp64,
        "plain": plain,
        "plain_large": plain_large,
        "bf16_blocks": bf16_leg,
        "spmv_phase": spmv_phase,
        "iteration_phase": iteration_phase,
        "pcg_form_comparison": form_cmp,
        "autotune": tuned,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "p99_request_latency_s": serving_leg["p99_latency_censored_s"] if serving_leg else None,
        "serving": serving_leg,
        "serving_sustainable": serving_sustainable,
        "copy_probe": copy_probe,
        "setup_s": round(setup_s, 2),
    }
    print(json.dumps(out), flush=True)
    # Release every device object and wait for the device before the
    # interpreter tears down the HIP runtime.
    del batch, cloths
    if plain_large is not None:
        del big
    import gc

    gc.collect()
    torch.cuda.synchronize()


if __name__ == "__main__":
    main()
