"""BF16 precision mode of the cloth step: Hessian blocks stored in BF16 for
the inner PCG iterations (FP32 arithmetic) with a diagonal that absorbs
their rounding, outer iterative refinement against the FP32 true residual.
Checked against the full-FP64 oracle."""
import os

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _bf16_cloth(spec):
    from chysx import scenes

    dc = scenes.DeviceCloth(spec)
    dc.sim.set_bf16_blocks(True)
    return dc


def _bf16_round(v):
    """float32 -> BF16 bit patterns, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def test_bf16_inner_matrix_bits(gpu):
    """The inner matrix of the BF16 mode, bit for bit: off-diagonal blocks
    are the round-to-nearest-even BF16 of the FP32 blocks, and the diagonal
    is D + sum_j (V_ij - bf16(V_ij)) accumulated in FP32 in column order.
    Hence the BF16 matrix maps block-constant vectors like the FP32 one."""
    from chysx import scenes

    spec = scenes.plain_small()
    dc = _bf16_cloth(spec)
    for _ in range(3):
        dc.step()
    D, ro, ci, V, _, _ = dc.sim.debug_dump_last_solve()
    Dc, V16 = dc.sim.debug_bf16_matrix()
    assert np.array_equal(V16, _bf16_round(V))
    back = (V16.astype(np.uint32) << 16).view(np.float32)
    delta = (V - back).astype(np.float32)  # exact in FP32
    expect = D.astype(np.float32).copy()
    deg = np.diff(ro)
    for pos in range(int(deg.max())):
        rows = np.nonzero(deg > pos)[0]
        expect[rows] = (expect[rows] + delta[ro[rows] + pos]).astype(np.float32)
    assert np.array_equal(Dc, expect)
    rows = np.repeat(np.arange(len(D)), deg)
    c = np.array([0.3, -1.2, 0.7])
    a32 = D.astype(np.float64) @ c
    np.add.at(a32, rows, V.astype(np.float64) @ c)
    a16 = Dc.astype(np.float64) @ c
    np.add.at(a16, rows, back.astype(np.float64) @ c)
    assert np.abs(a16 - a32).max() <= 1e-6 * np.abs(D @ c).max()


def test_bf16_substeps_match_fp64_oracle(gpu):
    """60 substeps of the 32x32 workload, each started from the FP64
    oracle's state: relative max-norm position error <= 1e-5 per substep;
    the solve stops on the true FP32 residual (<= tol) or at the cap."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small()
    dc = _bf16_cloth(spec)
    o = helpers.oracle_from_spec(spec, precision=64)
    worst = 0.0
    its, its64 = [], []
    for s in range(60):
        po, vo = o.state()
        dxo = o.system()[3] if s else np.zeros((spec.n, 3))
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        its64.append(o.step(spec.dt))
        p, _ = dc.state()
        err = helpers.rel_maxnorm(p, o.state()[0])
        worst = max(worst, err)
        it = dc.sim.last_pcg_iterations()
        its.append(it)
        res = dc.sim.last_relative_residual()
        assert err <= 1e-5, (s, err)
        assert it == spec.max_iter or res <= spec.tol, (s, it, res)
    print(f"BF16: worst one-substep rel err vs FP64 {worst:.3e}; PCG iterations mean {np.mean(its):.1f} "
          f"(FP64 oracle {np.mean(its64):.1f}), max excess {max(a - b for a, b in zip(its, its64))}")


def test_bf16_scattered_numbering_matches_fp64_oracle(gpu):
    """BF16 blocks on a randomly renumbered 64x64 grid, whose columns lie up
    to 4095 rows from their rows (32-bit sliced entries, SpMV from global
    memory): one-substep error from the oracle's state <= 1e-5."""
    import torch

    from chysx import scenes

    base = scenes.plain_small(seed=6, n=64)
    rng = np.random.default_rng(6)
    perm = rng.permutation(base.n)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(base.n)
    spec = scenes.ClothSpec(pos=base.pos[perm].copy(), tris=inv[base.tris].astype(np.int32),
                            pins=np.sort(inv[base.pins]).astype(np.int32), k_stretch=base.k_stretch,
                            k_shear=base.k_shear, k_bend=base.k_bend, density=base.density, name="scattered64")
    dc = _bf16_cloth(spec)
    o = helpers.oracle_from_spec(spec, precision=64)
    for s in range(10):
        po, vo = o.state()
        dxo = o.system()[3] if s else np.zeros((spec.n, 3))
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        o.step(spec.dt)
        assert not dc.sim.debug_entries16()
        err = helpers.rel_maxnorm(dc.state()[0], o.state()[0])
        assert err <= 1e-5, (s, err)
        it, res = dc.sim.last_pcg_iterations(), dc.sim.last_relative_residual()
        assert it == spec.max_iter or res <= spec.tol, (s, it, res)


def test_bf16_batch_equals_alone_bitwise(gpu):
    import chysx
    from chysx import scenes

    specs = scenes.batched(seed=21, count=4, lo=8, hi=36)
    alone = [_bf16_cloth(s) for s in specs]
    mixed = [_bf16_cloth(s) for s in specs]
    batch = chysx.ClothBatch()
    batch.set_bf16_blocks(True)
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
    batch.set_bf16_blocks(True)
    batch.add_instance(plain.sim)
    with pytest.raises(ValueError):
        batch.step(spec.dt)


def _precision_run(n, steps, max_iter, per_substep=False, floor=False):
    """The precision workload: one n x n grid, E = 1e6 Pa, seeded initial
    shear of up to 5 cm, `steps` substeps, PCG tol 1e-6 and `max_iter`.  Two
    GPU forms -- FP32 blocks / SpMV with FP64 reductions, and BF16 blocks +
    outer iterative refinement -- run free against a full-FP64 oracle run of
    the same substeps, beside the FP32 oracle (the reference algorithm in
    FP32 on the CPU, the arithmetic baseline).  Reported: relative max-norm
    position error after the last substep, GPU time per substep, PCG
    iterations.  per_substep: two more GPU simulators restart every substep
    from the FP64 oracle's state (positions, velocities, warm start) and
    their one-substep errors are recorded.  floor: an FP64 oracle at
    tolerance 1e-10 gives the distance the 1e-6 tolerance alone leaves."""
    import dataclasses
    import json
    import time

    import torch

    from chysx import scenes

    spec = dataclasses.replace(scenes.precision(n=n), max_iter=max_iter)
    report = {"workload": f"precision grid {n}x{n}, E=1e6 Pa, shear start, {steps} substeps, "
                          f"tol {spec.tol:g}, max_iter {max_iter}", "vertices": spec.n}
    forms = (("fp32", False), ("bf16_ir", True))
    free = {name: scenes.DeviceCloth(spec) for name, _ in forms}
    reset = {name: scenes.DeviceCloth(spec) for name, _ in forms} if per_substep else {}
    for name, bf16 in forms:
        if bf16:
            free[name].sim.set_bf16_blocks(True)
            if per_substep:
                reset[name].sim.set_bf16_blocks(True)
    o64 = helpers.oracle_from_spec(spec, precision=64)
    o32 = helpers.oracle_from_spec(spec, precision=32)
    o32r = helpers.oracle_from_spec(spec, precision=32) if per_substep else None
    otight = helpers.oracle_from_spec(spec, precision=64, pcg=(max(max_iter, 40 * n), 1e-10)) if floor else None
    its = {name: [] for name, _ in forms}
    ms = {name: 0.0 for name, _ in forms}
    worst = {name: 0.0 for name, _ in forms}
    worst32 = 0.0
    its64, its32 = [], []
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    for s in range(steps):
        if per_substep:
            po, vo = o64.state()
            dxo = o64.dx()
            for dc in reset.values():
                dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
                dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
                dc.sim.debug_set_warm_start(dxo.astype(np.float32))
                dc.step()
            # the FP32 oracle from the same state: the reference algorithm's
            # own one-substep distance from FP64
            o32r.set_state(po, vo, o32r.inv_mass())
            o32r.set_dx(dxo)
            o32r.step(spec.dt)
        for name, _ in forms:
            e0.record()
            free[name].step()
            e1.record()
            e1.synchronize()
            ms[name] += e0.elapsed_time(e1)
            its[name].append(free[name].sim.last_pcg_iterations())
        its64.append(o64.step(spec.dt))
        its32.append(o32.step(spec.dt))
        if otight is not None:
            otight.step(spec.dt)
        if per_substep:
            p64 = o64.state()[0]
            for name, dc in reset.items():
                worst[name] = max(worst[name], helpers.rel_maxnorm(dc.state()[0], p64))
            worst32 = max(worst32, helpers.rel_maxnorm(o32r.state()[0], p64))
        if (s + 1) % 10 == 0:
            print(f"substep {s + 1}/{steps} ({time.perf_counter() - t0:.0f} s)", flush=True)
    p64 = o64.state()[0]
    report["fp64_oracle"] = {"pcg_iterations_mean": float(np.mean(its64)), "threads": pyoracle_threads()}
    report["fp32_oracle"] = {"pcg_iterations_mean": float(np.mean(its32)),
                             "rel_maxnorm_error_vs_fp64": helpers.rel_maxnorm(o32.state()[0], p64)}
    if per_substep:
        report["fp32_oracle"]["worst_one_substep_error_vs_fp64"] = worst32
    if otight is not None:
        report["tolerance_floor"] = {
            "fp64_tol1e-6_vs_fp64_tol1e-10": helpers.rel_maxnorm(p64, otight.state()[0]),
            "note": "distance of the FP64 oracle at the workload tolerance from a tightly converged FP64 run"}
    for name, _ in forms:
        pf = free[name].state()[0]
        report[name] = {"ms_per_substep": round(ms[name] / steps, 3), "pcg_iterations_mean": float(np.mean(its[name])),
                        "pcg_iterations_max": int(np.max(its[name])),
                        "last_relative_residual": float(free[name].sim.last_relative_residual()),
                        "rel_maxnorm_error_vs_fp64": helpers.rel_maxnorm(pf, p64),
                        "max_abs_error_m": float(np.max(np.abs(pf - p64)))}
        if per_substep:
            report[name]["worst_one_substep_error_vs_fp64"] = worst[name]
    print("PRECISION " + json.dumps(report), flush=True)
    return report


def test_precision_config3_capped(gpu):
    """The precision workload at its own size (512 x 512, 100 substeps,
    the reference's cap of 100 iterations).  At the cap no form converges
    on this stiff grid (relative residual ~0.2), so each follows its own
    truncated Krylov path and even one substep of the reference algorithm in
    FP32 lands ~1e-4 from FP64 (measured on the MI355X box: 1.1e-4 for the
    GPU's FP32 form).  The bounds are therefore set against the FP32 oracle,
    the reference algorithm on the CPU, fixed before any run:
      * every substep, restarted from the FP64 oracle's state: each form's
        one-substep distance from the FP64 oracle's next state at most twice
        the FP32 oracle's (restarted from the same state), plus 1e-6;
      * free-running for 100 substeps, the FP32 form no further from the
        FP64 run than twice the FP32 oracle's own distance, the BF16 form
        within three times it."""
    rep = _precision_run(512, 100, max_iter=100, per_substep=True)
    base = rep["fp32_oracle"]["rel_maxnorm_error_vs_fp64"]
    base1 = rep["fp32_oracle"]["worst_one_substep_error_vs_fp64"]
    for name in ("fp32", "bf16_ir"):
        assert rep[name]["worst_one_substep_error_vs_fp64"] <= 2 * base1 + 1e-6, rep
    assert rep["fp32"]["rel_maxnorm_error_vs_fp64"] <= 2 * base + 1e-5, rep
    assert rep["bf16_ir"]["rel_maxnorm_error_vs_fp64"] <= 3 * base + 1e-5, rep


def test_precision_report_converged(gpu):
    """Every substep solved to the tolerance (128 x 128, cap far above the
    needs).  Bounds fixed before the run: FP32 free-running error after 100
    substeps at most twice the FP32 oracle's own distance from the FP64 run,
    BF16 at most four times it; BF16 iterations at most 2 x FP32's + 20 --
    with one outer refinement the BF16 form runs one inner solve to the
    tolerance (about the FP32 count) and at most one more from the refined
    residual, whose norm does not exceed the right-hand side's.  The report
    also states the tolerance floor: the FP64 oracle at 1e-6 against FP64 at
    1e-10, the error no arithmetic can remove at this tolerance."""
    # CHYSX_PRECISION_N / CHYSX_PRECISION_STEPS select another size for a
    # one-off report (the 512 x 512 floor run: N=512, STEPS=30).
    n = int(os.environ.get("CHYSX_PRECISION_N", "128"))
    steps = int(os.environ.get("CHYSX_PRECISION_STEPS", "100"))
    rep = _precision_run(n, steps, max_iter=40 * n, floor=True)
    f32, b16 = rep["fp32"], rep["bf16_ir"]
    base = rep["fp32_oracle"]["rel_maxnorm_error_vs_fp64"]
    for name in ("fp32", "bf16_ir"):
        assert rep[name]["pcg_iterations_max"] < 40 * n, (name, rep[name])
    assert f32["rel_maxnorm_error_vs_fp64"] <= 2 * base + 1e-5, rep
    assert b16["rel_maxnorm_error_vs_fp64"] <= 4 * base + 1e-5, rep
    assert b16["pcg_iterations_mean"] <= 2 * f32["pcg_iterations_mean"] + 20, rep


def pyoracle_threads():
    import pyoracle

    return pyoracle.num_threads()
