"""Batched form: uneven cloth instances advanced by one fused substep give
the same results as each instance stepped alone (bitwise: every reduction
is per instance, over the same tiles, in the same order)."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _specs(count=6, seed=11):
    from chysx import scenes

    specs = scenes.batched(seed=seed, count=count, lo=8, hi=40)
    return specs


def test_device_info_reports_grid_cap(gpu):
    from chysx import abi

    cus, cap, occ = abi.device_info()
    print(f"compute units {cus}, PCG grid cap {cap}, runtime occupancy {occ} workgroups/CU", flush=True)
    assert cus > 0 and 0 < cap <= 4 * cus


def test_batch_equals_alone_bitwise(gpu):
    import chysx
    from chysx import scenes

    specs = _specs()
    alone = [scenes.DeviceCloth(s) for s in specs]
    mixed = [scenes.DeviceCloth(s) for s in specs]
    batch = chysx.ClothBatch()
    for c in mixed:
        batch.add_instance(c.sim)
    print(f"batch of {len(specs)}: sizes {[s.n for s in specs]}", flush=True)
    for step in range(8):
        for c in alone:
            c.step()
        batch.step(specs[0].dt)
        its_b = batch.last_pcg_iterations()
        its_a = [c.sim.last_pcg_iterations() for c in alone]
        assert its_b == its_a, (step, its_b, its_a)
    assert batch.entries16()  # row-numbered grids: 16-bit sliced entries
    for a, m in zip(alone, mixed):
        pa, va = a.state()
        pm, vm = m.state()
        assert np.array_equal(pa, pm) and np.array_equal(va, vm)


def test_one_instance_batch_at_512_row_tiles_bitwise(gpu):
    """A one-instance ClothBatch runs as a single-instance batch (descriptor
    in the kernel arguments, tile list implied by the tile size); at 512-row
    tiles it matches the same cloth stepped alone and the same cloth inside
    a two-instance batch (instance table, explicit tile list) bit for bit."""
    import chysx
    from chysx import scenes

    spec, other = scenes.batched(seed=5, count=2, lo=40, hi=48)
    alone, one, two = (scenes.DeviceCloth(spec) for _ in range(3))
    partner = scenes.DeviceCloth(other)
    alone.sim.set_tile_rows(512)
    b1, b2 = chysx.ClothBatch(), chysx.ClothBatch()
    b1.set_tile_rows(512)
    b2.set_tile_rows(512)
    b1.add_instance(one.sim)
    b2.add_instance(two.sim)
    b2.add_instance(partner.sim)
    print(f"one-instance batch: {spec.n} rows, 512-row tiles", flush=True)
    for step in range(6):
        alone.step()
        b1.step(spec.dt)
        b2.step(spec.dt)
        its = (alone.sim.last_pcg_iterations(), b1.last_pcg_iterations()[0], b2.last_pcg_iterations()[0])
        assert its[0] == its[1] == its[2], (step, its)
    pa, va = alone.state()
    for c in (one, two):
        pc, vc = c.state()
        assert np.array_equal(pa, pc) and np.array_equal(va, vc)


def test_32bit_entries_equal_16bit_bitwise(gpu):
    """A row-numbered grid solved alone reads 16-bit sliced entries; batched
    with a randomly renumbered grid (columns up to 4095 rows away) the whole
    batch reads 32-bit entries.  Same blocks, same sums in the same order:
    the grid's states and iteration counts agree bit for bit."""
    import chysx
    from chysx import scenes

    grid = scenes.plain_small(seed=15, n=48)
    base = scenes.plain_small(seed=16, n=64)
    perm = np.random.default_rng(16).permutation(base.n)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(base.n)
    scattered = scenes.ClothSpec(pos=base.pos[perm].copy(), tris=inv[base.tris].astype(np.int32),
                                 pins=np.sort(inv[base.pins]).astype(np.int32), k_stretch=base.k_stretch,
                                 k_shear=base.k_shear, k_bend=base.k_bend, density=base.density, name="scattered64")
    alone = scenes.DeviceCloth(grid)
    mixed = [scenes.DeviceCloth(grid), scenes.DeviceCloth(scattered)]
    batch = chysx.ClothBatch()
    for c in mixed:
        batch.add_instance(c.sim)
    for step in range(8):
        alone.step()
        batch.step(grid.dt)
        assert alone.sim.debug_entries16() and not batch.entries16(), step
        assert batch.last_pcg_iterations()[0] == alone.sim.last_pcg_iterations(), step
    pa, va = alone.state()
    pm, vm = mixed[0].state()
    assert np.array_equal(pa, pm) and np.array_equal(va, vm)


def test_batch_matches_oracle(gpu):
    import chysx
    from chysx import scenes

    specs = _specs(count=4, seed=12)
    cloths = [scenes.DeviceCloth(s) for s in specs]
    oracles = [helpers.oracle_from_spec(s) for s in specs]
    batch = chysx.ClothBatch()
    for c in cloths:
        batch.add_instance(c.sim)
    for _ in range(5):
        batch.step(specs[0].dt)
        its_o = [o.step(s.dt) for o, s in zip(oracles, specs)]
        assert all(abs(a - b) <= 1 for a, b in zip(batch.last_pcg_iterations(), its_o))
    for c, o in zip(cloths, oracles):
        assert helpers.rel_maxnorm(c.state()[0], o.state()[0]) <= 1e-5
    res = batch.last_relative_residuals()
    assert max(res) <= 1e-6


def test_batch_grid_smaller_than_tiles(gpu):
    """Fewer persistent workgroups than tiles: each workgroup walks several
    tiles of several instances; results unchanged."""
    import chysx
    from chysx import scenes

    specs = _specs(count=5, seed=13)
    ref = [scenes.DeviceCloth(s) for s in specs]
    test = [scenes.DeviceCloth(s) for s in specs]
    b1, b2 = chysx.ClothBatch(), chysx.ClothBatch()
    for c in ref:
        b1.add_instance(c.sim)
    for c in test:
        b2.add_instance(c.sim)
    b2.set_grid_blocks(2)
    for _ in range(4):
        b1.step(specs[0].dt)
        b2.step(specs[0].dt)
    for a, m in zip(ref, test):
        assert np.array_equal(a.state()[0], m.state()[0])


@pytest.mark.parametrize("grid", [0, 16, 13])
def test_global_spmv_form_equals_staged_bitwise(gpu, grid):
    """The SpMV form without LDS staging (gathers through the caches, more
    workgroups per CU) computes the same sums in the same order: bitwise the
    staged form's results, with the full grid, a multiple of 8 (XCD
    segments of the compacted tile list) and an odd grid."""
    import chysx
    from chysx import scenes

    specs = _specs(count=6, seed=14)
    staged = [scenes.DeviceCloth(s) for s in specs]
    plain = [scenes.DeviceCloth(s) for s in specs]
    b1, b2 = chysx.ClothBatch(), chysx.ClothBatch()
    for c in staged:
        b1.add_instance(c.sim)
    for c in plain:
        b2.add_instance(c.sim)
    b2.set_lds_staging(False)
    assert b1.lds_staging() and not b2.lds_staging()
    if grid:
        b1.set_grid_blocks(grid)
        b2.set_grid_blocks(grid)
    for step in range(6):
        b1.step(specs[0].dt)
        b2.step(specs[0].dt)
        assert b1.last_pcg_iterations() == b2.last_pcg_iterations(), step
    for a, m in zip(staged, plain):
        assert np.array_equal(a.state()[0], m.state()[0]) and np.array_equal(a.state()[1], m.state()[1])


def test_many_instances_static_tile_list(gpu):
    """More instances than the compacted tile list holds (1024): the solve
    walks the static tile list; every instance still equals its solve alone."""
    import chysx
    from chysx import scenes

    specs = scenes.batched(seed=15, count=1030, lo=4, hi=7)
    alone = [scenes.DeviceCloth(s) for s in specs]
    mixed = [scenes.DeviceCloth(s) for s in specs]
    batch = chysx.ClothBatch()
    for c in mixed:
        batch.add_instance(c.sim)
    for _ in range(2):
        for c in alone:
            c.step()
        batch.step(specs[0].dt)
    assert batch.last_pcg_iterations() == [c.sim.last_pcg_iterations() for c in alone]
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.state()[0], m.state()[0])
