"""BF16 precision mode of the cloth step: Hessian blocks stored in BF16 for
the inner PCG iterations (FP32 arithmetic), iterative refinement against the
FP32 true residual.  Checked against the full-FP64 oracle."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _bf16_cloth(spec, ir_period=10):
    from chysx import scenes

    dc = scenes.DeviceCloth(spec)
    dc.sim.set_bf16_blocks(True, ir_period)
    return dc


@pytest.mark.parametrize("ir_period", [5, 10])
def test_bf16_substeps_match_fp64_oracle(gpu, ir_period):
    """60 substeps of the 32x32 workload, each started from the FP64
    oracle's state: relative max-norm position error <= 1e-5 per substep,
    and the solve stops on the true FP32 residual (<= tol) or at the cap."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small()
    dc = _bf16_cloth(spec, ir_period)
    o = helpers.oracle_from_spec(spec, precision=64)
    worst = 0.0
    for s in range(60):
        po, vo = o.state()
        dxo = o.system()[3] if s else np.zeros((spec.n, 3))
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        o.step(spec.dt)
        p, _ = dc.state()
        err = helpers.rel_maxnorm(p, o.state()[0])
        worst = max(worst, err)
        it = dc.sim.last_pcg_iterations()
        res = dc.sim.last_relative_residual()
        assert err <= 1e-5, (s, err)
        assert it == spec.max_iter or res <= spec.tol, (s, it, res)
        # Convergence is tested only on true residuals (restart points).
        assert it == spec.max_iter or it % ir_period == 0, (s, it)
    print(f"BF16 ir_period={ir_period}: worst one-substep rel err vs FP64 {worst:.3e}")


def test_bf16_batch_equals_alone_bitwise(gpu):
    import chysx
    from chysx import scenes

    specs = scenes.batched(seed=21, count=4, lo=8, hi=36)
    alone = [_bf16_cloth(s) for s in specs]
    mixed = [_bf16_cloth(s) for s in specs]
    batch = chysx.ClothBatch()
    batch.set_bf16_blocks(True, 10)
    for c in mixed:
        batch.add_instance(c.sim)
    for step in range(6):
        for c in alone:
            c.step()
        batch.step(specs[0].dt)
        assert batch.last_pcg_iterations() == [c.sim.last_pcg_iterations() for c in alone], step
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.state()[0], m.state()[0]) and np.array_equal(a.state()[1], m.state()[1])


def test_bf16_batch_needs_bf16_instances(gpu):
    import chysx
    from chysx import scenes

    spec = scenes.batched(seed=22, count=1, lo=8, hi=9)[0]
    plain = scenes.DeviceCloth(spec)
    batch = chysx.ClothBatch()
    batch.set_bf16_blocks(True, 10)
    batch.add_instance(plain.sim)
    with pytest.raises(ValueError):
        batch.step(spec.dt)


def test_precision_report(gpu):
    """The precision workload: one n x n grid (512 in the workload; default
    128 here, CHYSX_PRECISION_N to change), E = 1e6 Pa, seeded initial shear
    of up to 5 cm, 100 substeps (CHYSX_PRECISION_STEPS).  Two variants -- FP32
    blocks / SpMV with FP64 reductions, and BF16 blocks + iterative
    refinement -- against a full-FP64 oracle run of the same substeps.
    Reports the relative max-norm position error after the last substep, the
    GPU time per substep and the PCG iterations.  Both variants must stay
    within the drift a free-running FP32 trajectory shows."""
    import json
    import os
    import time

    import torch

    from chysx import scenes

    n = int(os.environ.get("CHYSX_PRECISION_N", "128"))
    steps = int(os.environ.get("CHYSX_PRECISION_STEPS", "100"))
    spec = scenes.precision(n=n)
    report = {"workload": f"precision grid {n}x{n}, E=1e6 Pa, shear start, {steps} substeps", "vertices": spec.n}
    finals = {}
    for name, bf16 in (("fp32", False), ("bf16_ir10", True)):
        dc = scenes.DeviceCloth(spec)
        if bf16:
            dc.sim.set_bf16_blocks(True, 10)
        its = []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ms = 0.0
        for s in range(steps):
            e0.record()
            dc.step()
            e1.record()
            e1.synchronize()
            ms += e0.elapsed_time(e1)
            its.append(dc.sim.last_pcg_iterations())
        finals[name] = dc.state()[0]
        report[name] = {"ms_per_substep": round(ms / steps, 3), "pcg_iterations_mean": float(np.mean(its)),
                        "pcg_iterations_max": int(np.max(its)),
                        "last_relative_residual": float(dc.sim.last_relative_residual())}
        print(f"{name}: {ms / steps:.3f} ms/substep, iterations mean {np.mean(its):.1f}", flush=True)
        del dc
    o = helpers.oracle_from_spec(spec, precision=64)
    t0 = time.perf_counter()
    its64 = []
    for s in range(steps):
        its64.append(o.step(spec.dt))
        if (s + 1) % 10 == 0:
            print(f"FP64 oracle substep {s + 1}/{steps} ({time.perf_counter() - t0:.0f} s)", flush=True)
    p64 = o.state()[0]
    report["fp64_oracle"] = {"pcg_iterations_mean": float(np.mean(its64)), "seconds": round(time.perf_counter() - t0, 1),
                             "threads": pyoracle_threads()}
    for name in finals:
        report[name]["rel_maxnorm_error_vs_fp64"] = helpers.rel_maxnorm(finals[name], p64)
        report[name]["max_abs_error_m"] = float(np.max(np.abs(finals[name] - p64)))
    print("PRECISION " + json.dumps(report), flush=True)
    for name in finals:
        assert report[name]["rel_maxnorm_error_vs_fp64"] <= 1e-3, (name, report[name])


def pyoracle_threads():
    import pyoracle

    return pyoracle.num_threads()
