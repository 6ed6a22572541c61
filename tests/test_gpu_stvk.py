"""GPU parity of the StVK membrane step (no reference counterpart; the
oracle's StVK is pinned in test_oracle_stvk.py) on the 32x32 workload with
E = 1e5 Pa, nu = 0.3, h = 1 mm."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def test_stvk_pattern_and_slots_bit_exact(gpu):
    from chysx import scenes

    spec = scenes.plain_small_stvk()
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    dc.step()
    o.step(spec.dt)
    _, ro, ci, _, _, _ = dc.sim.debug_dump_last_solve()
    ro_o, ci_o = o.pattern()
    assert np.array_equal(ro, ro_o) and np.array_equal(ci, ci_o)
    assert np.array_equal(dc.sim.debug_slot_tables()["stvk"], o.slots("stvk", 9 * len(spec.tris)))


def test_stvk_substeps_match_oracle(gpu):
    """All 600 substeps of the 32x32 workload (60 frames x 10), each
    started from the FP32 oracle's state: relative max-norm position error
    <= 1e-5 and PCG iterations within +-1."""
    import torch

    from chysx import scenes

    spec = scenes.plain_small_stvk(seed=1)
    dc = scenes.DeviceCloth(spec)
    o = helpers.oracle_from_spec(spec)
    worst = 0.0
    for s in range(600):
        po, vo = o.state()
        dxo = o.dx()
        dc.pos.copy_(torch.from_numpy(po.astype(np.float32)).cuda())
        dc.vel.copy_(torch.from_numpy(vo.astype(np.float32)).cuda())
        dc.sim.debug_set_warm_start(dxo.astype(np.float32))
        dc.step()
        it_o = o.step(spec.dt)
        err = helpers.rel_maxnorm(dc.state()[0], o.state()[0])
        worst = max(worst, err)
        assert err <= 1e-5, (s, err)
        assert abs(dc.sim.last_pcg_iterations() - it_o) <= 1, (s, dc.sim.last_pcg_iterations(), it_o)
    print(f"StVK worst one-substep rel err {worst:.3e}")


def test_stvk_energy_matches_oracle(gpu):
    """Total energy (pins, StVK, bending) at a sagged state against the
    FP64 oracle, within twice the FP32 oracle's own distance from it (the
    Green strain (|f|^2 - 1)/2 cancels in FP32 at small strains); the StVK
    part is positive and the kind mask splits the total."""
    from chysx import scenes

    spec = scenes.plain_small_stvk(seed=2)
    dc = scenes.DeviceCloth(spec)
    for _ in range(30):
        dc.step()
    p, _ = dc.state()
    energies = {}
    for prec in (32, 64):
        o = helpers.oracle_from_spec(spec, precision=prec)
        o.set_state(p, np.zeros_like(p), o.inv_mass())
        energies[prec] = o.energy()
    e_stvk = dc.sim.compute_energy(8)
    e_all = dc.sim.compute_energy()
    e_rest = dc.sim.compute_energy(1 | 4)
    assert e_stvk > 0.0
    assert abs((e_stvk + e_rest) - e_all) <= 1e-6 * abs(e_all)
    assert abs(e_all - energies[64]) <= 2.0 * abs(energies[32] - energies[64]) + 1e-7 * abs(energies[64])
