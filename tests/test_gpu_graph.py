"""hipGraph-captured substep (C ABI chx_cloth_step_graph / chx_graph_launch):
replays give the same bits as eager substeps of an identical batch, with the
program's own eager work and synchronisations between replays; a change of
the batch re-captures."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs(count=5, seed=21):
    from chysx import scenes

    return scenes.batched(seed=seed, count=count, lo=8, hi=40)


def _states(cloths):
    return [c.state() for c in cloths]


def test_graph_replays_equal_eager_bitwise(gpu):
    import torch

    import chysx
    from chysx import scenes

    specs = _specs()
    eager = [scenes.DeviceCloth(s) for s in specs]
    graph = [scenes.DeviceCloth(s) for s in specs]
    be, bg = chysx.ClothBatch(), chysx.ClothBatch()
    for c in eager:
        be.add_instance(c.sim)
    for c in graph:
        bg.add_instance(c.sim)
    bg.set_graph(True)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    dt = specs[0].dt
    print(f"graph batch of {len(specs)}: sizes {[s.n for s in specs]}", flush=True)
    scratch = torch.ones(1 << 20, device="cuda")
    for step in range(6):
        with torch.cuda.stream(stream):
            be.step(dt, sh)
            bg.step(dt, sh)
            # eager work of the program between replays
            scratch.mul_(1.0001)
            _ = float(scratch.sum())
        stream.synchronize()
        assert bg.last_pcg_iterations(sh) == be.last_pcg_iterations(sh), step
        for (pe, ve), (pg, vg) in zip(_states(eager), _states(graph)):
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), step
    assert bg.graph_captures() == 1
    # one instance fewer: the descriptor changes, so the substep is captured again
    be.remove_instance(eager[-1].sim)
    bg.remove_instance(graph[-1].sim)
    for step in range(3):
        with torch.cuda.stream(stream):
            be.step(dt, sh)
            bg.step(dt, sh)
        stream.synchronize()
        for (pe, ve), (pg, vg) in zip(_states(eager[:-1]), _states(graph[:-1])):
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), step
    assert bg.graph_captures() == 2


def test_graph_needs_a_stream(gpu):
    import chysx
    from chysx import scenes

    spec = _specs(count=1)[0]
    c = scenes.DeviceCloth(spec)
    b = chysx.ClothBatch()
    b.add_instance(c.sim)
    b.set_graph(True)
    with pytest.raises(ValueError, match="non-null stream"):
        b.step(spec.dt, 0)


def test_graph_with_static_contact_and_mixed_instances(gpu):
    """A mix of drape (static plane / box contact with friction) and plain
    instances: graph replays equal eager substeps bitwise; the contact
    parameters travel in the instance table, so 40 replays need one
    capture."""
    import torch

    import chysx
    from chysx import scenes

    specs = [scenes.drape(n=24, friction=0.3), scenes.drape(n=20), *_specs(count=2, seed=22)]
    eager = [scenes.DeviceCloth(s) for s in specs]
    graph = [scenes.DeviceCloth(s) for s in specs]
    be, bg = chysx.ClothBatch(), chysx.ClothBatch()
    for c in eager:
        be.add_instance(c.sim)
    for c in graph:
        bg.add_instance(c.sim)
    bg.set_graph(True)
    stream = torch.cuda.Stream()
    sh = stream.cuda_stream
    dt = specs[0].dt
    print(f"graph contact mix: sizes {[s.n for s in specs]}", flush=True)
    for step in range(40):
        with torch.cuda.stream(stream):
            be.step(dt, sh)
            bg.step(dt, sh)
        if step % 10 == 9:
            stream.synchronize()
            for (pe, ve), (pg, vg) in zip(_states(eager), _states(graph)):
                assert np.array_equal(pe, pg) and np.array_equal(ve, vg), step
    assert bg.graph_captures() == 1
