"""GPU parity of the drop-in boundary's less-travelled branches against the
oracle: damping in the finalize (cloth_simulator.cu:222-244), mid-run
update_pin_targets on the caller's stream, eager and replayed from captured
graphs (cloth_simulator.cu:294-313), the standalone PCGSolver's asynchronous
solve and last_residual (pcg_solver.cu:282-323), move semantics
(cloth_simulator.h:72-75), and the solver's sticky abort report."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _reset_to_oracle(dc, o, s):
    import torch

    po, vo = o.state()
    dxo = o.system()[3] if s else np.zeros((dc.spec.n, 3))
    dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
    dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
    dc.sim.debug_set_warm_start(dxo.astype(np.float32))


def test_damping_matches_oracle(gpu):
    """damping > 0: v = exp(-c dt) dx / dt in the finalize; every substep
    from the oracle's state within 1e-5 (positions) and iterations within
    +-1.  From the same x_n, v - v_o = (p - p_o) exp(-c dt) / dt up to the
    FP32 rounding of x_n + dx, which bounds the velocity difference."""
    from dataclasses import replace

    from chysx import scenes

    spec = replace(scenes.plain_small(seed=21), damping=4.0)
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    for s in range(60):
        _reset_to_oracle(dc, o, s)
        dc.step()
        it_o = o.step(spec.dt)
        p, v = dc.state()
        po, vo = o.state()
        assert helpers.rel_maxnorm(p, po) <= 1e-5, s
        dp = np.max(np.abs(p - po)) + 2.0 ** -23 * np.max(np.abs(po))
        assert np.max(np.abs(v - vo)) <= dp / spec.dt * (1 + 1e-6) + 1e-6 * np.max(np.abs(vo)), s
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, s
    # the damped velocity differs from dx/dt by the decay factor
    dx = dc.sim.debug_dump_last_solve()[5].astype(np.float64)
    assert np.allclose(v, np.exp(-np.float32(4.0) * np.float32(spec.dt)) * dx / np.float32(spec.dt), rtol=1e-5,
                       atol=1e-7)


def _twist(spec, s):
    """Pin targets rotated about the cloth's vertical axis by 0.02 rad per
    substep (the reference's twist example drives update_pin_targets every
    frame, example_chysx_twist.py:439-447)."""
    t0 = spec.pos[spec.pins].astype(np.float64)
    c = t0.mean(0)
    a = 0.02 * (s + 1)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    return ((t0 - c) @ R.T + c).astype(np.float32)


def test_update_pin_targets_eager_matches_oracle(gpu):
    import torch

    from chysx import scenes

    spec = scenes.plain_small(seed=22)
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    stream = torch.cuda.Stream()
    for s in range(40):
        _reset_to_oracle(dc, o, s)
        t = _twist(spec, s)
        with torch.cuda.stream(stream):
            dc.sim.update_pin_targets(t, stream.cuda_stream)
            dc.sim.step(spec.dt, stream.cuda_stream)
        stream.synchronize()
        o.update_pin_targets(t.astype(np.float64))
        it_o = o.step(spec.dt)
        p, _ = dc.state()
        po, _ = o.state()
        assert helpers.rel_maxnorm(p, po) <= 1e-5, s
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, s
    # the pinned corners follow the moving targets
    assert np.max(np.abs(p[spec.pins] - t)) < 1e-3


def test_update_pin_targets_between_graph_replays(gpu):
    """A ClothBatch substep captured once as a hipGraph: targets uploaded on
    the stream between replays reach the captured element pass (no new
    capture); every replay matches the oracle stepped with the same
    targets, free-running within 1e-5 over 30 substeps."""
    import torch

    import chysx
    from chysx import scenes

    spec = scenes.plain_small(seed=23)
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    batch = chysx.ClothBatch()
    batch.add_instance(dc.sim)
    batch.set_graph(True)
    stream = torch.cuda.Stream()
    for s in range(30):
        t = _twist(spec, s)
        with torch.cuda.stream(stream):
            dc.sim.update_pin_targets(t, stream.cuda_stream)
            batch.step(spec.dt, stream.cuda_stream)
        stream.synchronize()
        o.update_pin_targets(t.astype(np.float64))
        o.step(spec.dt)
        p, _ = dc.state()
        po, _ = o.state()
        assert helpers.rel_maxnorm(p, po) <= 1e-5, s
    assert batch.graph_captures() == 1


def test_update_pin_targets_between_torch_graph_replays(gpu):
    """ClothSimulator.step captured in a torch CUDA graph (the way Newton
    captures the solver step, example_chysx_cloth_drop.py:182-214); targets
    change between replays, with eager work and synchronisations in between.
    Every replay equals an eager simulator given the same targets, bit for
    bit, and stays within 1e-5 of the oracle."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small(seed=24)
    cap = scenes.DeviceCloth(spec)
    ref = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        t = _twist(spec, -1)
        for c in (cap, ref):
            c.sim.update_pin_targets(t, stream.cuda_stream)
            c.sim.step(spec.dt, stream.cuda_stream)  # warm-up: builds the device arrays
        stream.synchronize()
        o.update_pin_targets(t.astype(np.float64))
        o.step(spec.dt)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            cap.sim.step(spec.dt, stream.cuda_stream)
    for s in range(20):
        t = _twist(spec, s)
        with torch.cuda.stream(stream):
            cap.sim.update_pin_targets(t, stream.cuda_stream)
            g.replay()
            ref.sim.update_pin_targets(t, stream.cuda_stream)
            ref.sim.step(spec.dt, stream.cuda_stream)
        stream.synchronize()
        o.update_pin_targets(t.astype(np.float64))
        o.step(spec.dt)
        pc, vc = cap.state()
        pr, vr = ref.state()
        assert np.array_equal(pc, pr) and np.array_equal(vc, vr), s
        assert helpers.rel_maxnorm(pc, o.state()[0]) <= 1e-5, s


def test_move_roundtrip_keeps_the_simulator(gpu):
    """Move construction and move assignment keep buffers, topology and the
    warm start: a moved simulator keeps stepping like the oracle."""
    from chysx import scenes

    spec = scenes.plain_small(seed=25)
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    for s in range(10):
        dc.step()
        o.step(spec.dt)
        dc.sim.debug_move_roundtrip()
    dc.step()
    o.step(spec.dt)
    assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5
    assert dc.sim.num_particles() == spec.n


def test_aborted_solve_is_reported_by_the_next_step(gpu):
    """A grid-barrier wait bound of one tick makes the solve give up at its
    first wait (every workgroup leaves the kernel normally).  The abort stays
    reported across substeps: the next step raises RuntimeError; after the
    report the simulator steps normally again and matches the oracle."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small(seed=26, n=64)  # 16 row tiles: several workgroups wait at each barrier
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    dc.step()
    o.step(spec.dt)
    dc.sim.set_barrier_timeout_ticks(1)
    dc.step()
    torch.cuda.synchronize()
    dc.sim.set_barrier_timeout_ticks(0)
    with pytest.raises(RuntimeError, match="grid barrier timed out"):
        dc.step()
    # the report is cleared; continue from the oracle's state
    _reset_to_oracle(dc, o, 1)
    for s in range(5):
        dc.step()
        it_o = o.step(spec.dt)
        assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5, s
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, s


def test_pcg_solver_async_and_last_residual(gpu):
    """PCGSolver.solve without a tolerance runs exactly max_iterations and
    returns that count without waiting; last_residual() is the float
    <r, z> of the final iteration (the oracle's, within FP32 rounding), and
    the additions last_iterations() / last_relative_residual() read the
    rest.  With a tolerance solve() waits and returns the iterations used."""
    import torch

    import chysx
    import pyoracle
    from chysx import scenes

    spec = scenes.plain_small(seed=27)
    dc = scenes.DeviceCloth(spec)
    dc.step()
    diag, ro, ci, vals, rhs, _ = dc.sim.debug_dump_last_solve()
    n = len(ro) - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(ro))
    A = chysx.BlockCSR3()
    A.build_topology(n, rows, ci)
    assert np.array_equal(A.row_offsets(), ro) and np.array_equal(A.col_indices(), ci)
    A.upload(diag, vals)
    b = torch.from_numpy(rhs).cuda()
    for iters in (3, 17):
        x = torch.zeros_like(b)
        S = chysx.PCGSolver()
        assert S.solve(A, b.data_ptr(), x.data_ptr(), max_iterations=iters) == iters
        rho = S.last_residual()
        assert isinstance(rho, float)
        xo, it_o, rr_o, bb_o, rho_o = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, np.zeros_like(rhs), iters, 0.0,
                                                         precision=32, return_rho=True)
        assert S.last_iterations() == iters == it_o
        assert abs(rho - rho_o) <= 1e-3 * abs(rho_o), (rho, rho_o)
        assert abs(S.last_relative_residual() - np.sqrt(rr_o / bb_o)) <= 1e-3 * np.sqrt(rr_o / bb_o)
        assert helpers.rel_maxnorm(x.cpu().numpy(), xo) <= 1e-5
    x = torch.zeros_like(b)
    it = S.solve(A, b.data_ptr(), x.data_ptr(), max_iterations=500, tolerance=1e-6)
    _, it_o, _, _ = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, np.zeros_like(rhs), 500, 1e-6, precision=32)
    assert abs(it - it_o) <= 1 and it < 500
    assert S.last_relative_residual() <= 1e-6
