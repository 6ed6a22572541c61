"""GPU parity of the full implicit cloth substep (ClothSimulator.step through
the C ABI) against the CPU oracle on the 32x32 workload, plus the
reference's own example assertions (free fall, hanging cloth)."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _make(spec, lanes=None):
    from chysx import scenes

    return scenes.DeviceCloth(spec, lanes_per_row=lanes)


def test_topology_and_pattern_bit_exact(gpu):
    from chysx import scenes

    spec = scenes.plain_small()
    dc = _make(spec)
    o = helpers.oracle_from_spec(spec)
    dc.step()
    o.step(spec.dt)
    diag, ro, ci, vals, rhs, dx = dc.sim.debug_dump_last_solve()
    ro_o, ci_o = o.pattern()
    assert np.array_equal(ro, ro_o) and np.array_equal(ci, ci_o)
    assert dc.sim.num_bending_dihedrals() == o.sizes()["dihedrals"]
    d_g, _ = dc.sim.debug_dihedrals()
    d_o, _ = o.dihedrals()
    assert np.array_equal(d_g, d_o)  # same order (libstdc++ unordered_map)
    slots = dc.sim.debug_slot_tables()
    for kind, cnt in (("pin", 2), ("stretch", 9 * 1922), ("shear", 9 * 1922), ("bending", 16 * 2821)):
        key = "pins" if kind == "pin" else kind
        assert np.array_equal(slots[key], o.slots(kind, cnt)), kind


def test_first_step_system_matches_oracle(gpu):
    from chysx import scenes

    spec = scenes.plain_small()
    dc = _make(spec)
    o = helpers.oracle_from_spec(spec)
    o64 = helpers.oracle_from_spec(spec, precision=64)
    dc.step()
    o.step(spec.dt)
    o64.step(spec.dt)
    diag, ro, ci, vals, rhs, dx = dc.sim.debug_dump_last_solve()
    diag_o, vals_o, rhs_o, dx_o = o.system()
    _, _, rhs_64, dx_64 = o64.system()
    assert helpers.rel_maxnorm(diag, diag_o) <= 1e-5
    assert np.max(np.abs(vals - vals_o)) <= 1e-5 * np.abs(diag_o).max()
    # At the rest state the elastic forces are rounding noise; the GPU may
    # differ from the FP32 oracle by at most twice the FP32 oracle's own
    # distance from the FP64 oracle.
    assert np.max(np.abs(rhs - rhs_o)) <= 2.0 * np.max(np.abs(rhs_o - rhs_64))
    assert np.max(np.abs(dx - dx_o)) <= 2.0 * np.max(np.abs(dx_o - dx_64))
    p, v = dc.state()
    po, vo = o.state()
    assert helpers.rel_maxnorm(p, po) <= 1e-5
    assert abs(dc.sim.last_pcg_iterations() - o.last_stats()[0]) <= 1


def test_substep_parity_60_frames(gpu):
    """Every one of the 600 substeps (60 frames x 10), started from the
    oracle's state: relative max-norm position error <= 1e-5 and PCG
    iteration counts within +-1.  The free-running trajectories are also
    compared."""
    from chysx import scenes

    spec = scenes.plain_small()
    dc = _make(spec)
    o = helpers.oracle_from_spec(spec)
    free = _make(spec)
    worst, worst_it = 0.0, 0
    for s in range(600):
        po, vo = o.state()
        _, _, _, dxo = o.system() if s else (None, None, None, np.zeros((spec.n, 3)))
        import torch

        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        free.step()
        it_o = o.step(spec.dt)
        p, v = dc.state()
        pn, _ = o.state()
        err = helpers.rel_maxnorm(p, pn)
        worst = max(worst, err)
        it_g = dc.sim.last_pcg_iterations()
        worst_it = max(worst_it, abs(it_g - it_o))
        assert err <= 1e-5, (s, err)
        assert abs(it_g - it_o) <= 1, (s, it_g, it_o)
    pf, _ = free.state()
    po, _ = o.state()
    print(f"worst one-substep rel err {worst:.3e}, worst |dit| {worst_it}, "
          f"free-running rel err after 600 substeps {helpers.rel_maxnorm(pf, po):.3e}")


def test_deterministic_bitwise(gpu):
    from chysx import scenes

    spec = scenes.plain_small(seed=7)
    outs = []
    for _ in range(2):
        dc = _make(spec)
        for _ in range(20):
            dc.step()
        outs.append(dc.state())
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("lanes", [4, 16])
def test_lanes_per_row_variants(gpu, lanes):
    from chysx import scenes

    spec = scenes.plain_small(seed=3)
    dc = _make(spec, lanes)
    o = helpers.oracle_from_spec(spec)
    for _ in range(5):
        dc.step()
        o.step(spec.dt)
    assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5


@pytest.mark.parametrize("tile_rows", [512, 1024])
def test_tile_rows_variants(gpu, tile_rows):
    """Tiles of several 256-row stages (one workgroup and one FP64 partial
    per tile) give the oracle's substeps and iteration counts."""
    from chysx import scenes

    spec = scenes.plain_small(seed=4)
    dc = _make(spec)
    dc.sim.set_tile_rows(tile_rows)
    o = helpers.oracle_from_spec(spec)
    for _ in range(5):
        dc.step()
        it_o = o.step(spec.dt)
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1
    assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5


def test_scattered_numbering_uses_global_path(gpu):
    """A randomly renumbered 64x64 grid: each 256-row stage references
    columns across the whole mesh, beyond the LDS window, so the PCG's SpMV
    runs from global memory.  Same oracle bar, substeps started from the
    oracle's state."""
    import torch

    from chysx import scenes

    base = scenes.plain_small(seed=6, n=64)
    rng = np.random.default_rng(6)
    perm = rng.permutation(base.n)          # new index of old vertex i: inv[i]
    inv = np.empty_like(perm)
    inv[perm] = np.arange(base.n)
    spec = scenes.ClothSpec(pos=base.pos[perm].copy(), tris=inv[base.tris].astype(np.int32),
                            pins=np.sort(inv[base.pins]).astype(np.int32), k_stretch=base.k_stretch,
                            k_shear=base.k_shear, k_bend=base.k_bend, density=base.density, name="scattered64")
    dc = _make(spec)
    o = helpers.oracle_from_spec(spec)
    for s in range(20):
        po, vo = o.state()
        dxo = o.system()[3] if s else np.zeros((spec.n, 3))
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        it_o = o.step(spec.dt)
        if s == 0:
            windows = np.asarray(dc.sim.debug_stage_windows())
            assert np.all(windows[:, 1] - windows[:, 0] + 1 > 2816), windows[:4]
            assert not dc.sim.debug_entries16()  # columns up to 4095 rows away: 32-bit entries
        assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5, s
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, s


def test_grid_numbering_uses_16bit_entries(gpu):
    """A grid numbered row by row keeps every column within 2047 rows of its
    row, so the SpMV reads the 16-bit sliced entries; parity as above."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small(seed=7, n=64)
    dc = _make(spec)
    o = helpers.oracle_from_spec(spec)
    for s in range(10):
        po, vo = o.state()
        dxo = o.system()[3] if s else np.zeros((spec.n, 3))
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        it_o = o.step(spec.dt)
        assert dc.sim.debug_entries16()
        assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5, s
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, s


def test_freefall_closed_form(gpu):
    """The reference's free-fall example check (example_chysx_freefall.py:
    116-178): no elastic elements, particle 0 pinned; free particles follow
    x = x0 + dt^2 g n(n+1)/2, v = n dt g (implicit Euler with PCG), within
    5 %; the pin stays put."""
    import torch

    import chysx
    from chysx import scenes

    pos, tris = scenes.cloth_grid(11, side=1.0, height=2.0)
    n = len(pos)
    p = torch.from_numpy(pos).cuda()
    v = torch.zeros_like(p)
    w = torch.full((n,), 1.0 / 0.05, device="cuda")
    sim = chysx.ClothSimulator()
    m = chysx.ClothMaterial()
    m.gx, m.gy, m.gz = 0.0, 0.0, -9.81
    sim.set_material(m)
    sim.set_external_buffers(p.data_ptr(), v.data_ptr(), n, w.data_ptr())
    sim.set_pins(np.array([0], np.int32), pos[:1], 1.0e6)
    dt = 1.0 / 600.0
    steps = 600
    for _ in range(steps):
        sim.step(dt)
    torch.cuda.synchronize()
    q = p.cpu().numpy().astype(np.float64)
    qd = v.cpu().numpy().astype(np.float64)
    t = steps * dt
    z_exp = pos[1:, 2] - 0.5 * 9.81 * t * t
    vz_exp = -9.81 * t
    assert np.all(np.abs(q[1:, 2] - z_exp) < max(0.05 * 0.5 * 9.81 * t * t, 0.05))
    assert np.all(np.abs(qd[1:, 2] - vz_exp) < max(0.05 * 9.81 * t, 0.05))
    assert np.all(np.abs(qd[1:, :2]) < 1e-4)
    # exact implicit-Euler closed form of the discrete scheme
    z_disc = pos[1:, 2] - 9.81 * dt * dt * steps * (steps + 1) / 2
    assert np.max(np.abs(q[1:, 2] - z_disc)) < 2e-3
    assert np.max(np.abs(q[0] - pos[0])) < 1e-4 and np.max(np.abs(qd[0])) < 1e-2


def test_hanging_cloth_acceptance(gpu):
    """The reference's hanging-cloth example checks
    (example_chysx_hanging_cloth.py:182-227) on a 101x101 grid:
    BW98 stretch = shear = 1e2, bending 5e-4, pins k = 1e9 at two corners,
    50 PCG iterations, density 0.1, 2 s of simulation."""
    import torch

    import chysx
    from chysx import scenes

    nx = 101
    pos, tris = scenes.cloth_grid(nx, side=1.0, height=0.0)
    # lay in the XY plane at z = 2 (the example's z-up set-up)
    pos = np.stack([pos[:, 0] - 0.5, pos[:, 2] - 0.5, np.full(len(pos), 2.0, np.float32)], 1).astype(np.float32)
    n = len(pos)
    p = torch.from_numpy(pos).cuda()
    v = torch.zeros_like(p)
    w = torch.ones(n, device="cuda")
    sim = chysx.ClothSimulator()
    m = chysx.ClothMaterial()
    m.gx, m.gy, m.gz = 0.0, 0.0, -9.81
    sim.set_material(m)
    sim.set_pcg_iterations(50)
    sim.set_external_buffers(p.data_ptr(), v.data_ptr(), n, w.data_ptr())
    sim.set_mesh(tris)
    sim.redistribute_mass_area_weighted(0.1, w.data_ptr(), n)
    sim.build_fem_stretch_from_current_positions(1.0e2)
    sim.build_fem_shear_from_current_positions(1.0e2)
    sim.build_bending_from_current_positions(5.0e-4)
    pins = np.array([0, nx - 1], np.int32)
    sim.set_pins(pins, pos[pins], 1.0e9)
    fps, substeps, frames = 60, 10, 120
    far = (nx - 1) * nx + (nx - 1) // 2
    drop = []
    for _ in range(frames):  # the cloth swings: drape checked on the max drop
        for _ in range(substeps):
            sim.step(1.0 / fps / substeps)
        torch.cuda.synchronize()
        q = p.cpu().numpy()
        qd = v.cpu().numpy()
        assert np.isfinite(q).all() and np.isfinite(qd).all()
        for i in pins:
            assert np.linalg.norm(q[i] - pos[i]) <= 1e-3
        assert q[:, 2].mean() < 2.0
        assert np.linalg.norm(qd, axis=1).max() <= 20.0
        drop.append(2.0 - q[far, 2])
    assert max(drop) >= 0.5


def test_error_paths(gpu):
    import torch

    import chysx

    sim = chysx.ClothSimulator()
    sim.step(0.01)  # no particles: no-op
    p = torch.zeros((4, 3), device="cuda")
    sim.set_external_buffers(0, 0, 4)
    with pytest.raises(RuntimeError):
        sim.step(0.01)
    sim.set_external_buffers(p.data_ptr(), p.data_ptr(), 4)
    with pytest.raises(ValueError):
        sim.step(0.0)
    sim.set_solver_type(1)
    with pytest.raises(RuntimeError):
        sim.step(0.01)
    sim.set_solver_type(0)
    sim.set_pins(np.array([0], np.int32), np.zeros((1, 3), np.float32))
    with pytest.raises(ValueError):
        sim.update_pin_targets(np.zeros((2, 3), np.float32))
    sim.set_mesh(np.array([[0, 1, 9]], np.int32))
    with pytest.raises(ValueError):
        sim.redistribute_mass_area_weighted(0.1, 0, 4)
    with pytest.raises(ValueError):
        sim.redistribute_mass_area_weighted(-1.0, p.data_ptr(), 4)


def test_energy_matches_oracle(gpu):
    from chysx import scenes

    spec = scenes.plain_small(seed=9)
    dc = _make(spec)
    o = helpers.oracle_from_spec(spec)
    for _ in range(3):
        dc.step()
        o.step(spec.dt)
    e_g = dc.sim.compute_energy()
    e_o = o.energy()
    assert abs(e_g - e_o) <= 1e-3 * abs(e_o) + 1e-9


@pytest.mark.parametrize("form", [0, 1])
def test_pcg_forms_match_oracle(gpu, form):
    """Both persistent PCG forms of the FP32 solve -- one grid barrier per
    iteration (the default: sigma and the next rho's recurrence terms in one
    reduction) and two -- on 60 substeps of the 32x32 workload, each from the
    oracle's state: 1e-5 relative max-norm and iterations within +-1."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small(seed=8)
    dc = _make(spec)
    dc.sim.set_pcg_form(form)
    assert dc.sim.pcg_form() == form
    o = helpers.oracle_from_spec(spec)
    for s in range(60):
        po, vo = o.state()
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(o.dx().astype(np.float32))
        dc.step()
        it_o = o.step(spec.dt)
        assert helpers.rel_maxnorm(dc.state()[0], o.state()[0]) <= 1e-5, (form, s)
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, (form, s, dc.sim.last_pcg_iterations(), it_o)
