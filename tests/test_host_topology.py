"""The product's host topology (chysx.topology, the code ClothSimulator runs)
is bit-exact with the oracle: edges, dihedrals in the same order, BSR
pattern and every slot table; errors map to the reference's exception types."""
import numpy as np
import pytest

import chysx
import pyoracle
from chysx import scenes


def _meshes():
    yield "grid32", *scenes.cloth_grid(32)
    yield "grid7x5", *scenes.cloth_grid(7, 5)
    pos, tris = scenes.cloth_grid(10)
    rng = np.random.default_rng(0)
    keep = np.sort(rng.choice(len(tris), size=len(tris) - 30, replace=False))
    yield "ragged", pos, tris[keep]
    # non-manifold fan: three triangles share edge (0, 1)
    pos = np.zeros((5, 3), np.float32)
    yield "fan", pos, np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.int32)


@pytest.mark.parametrize("name,pos,tris", list(_meshes()))
def test_topology_bit_exact(name, pos, tris):
    n = len(pos)
    pins = np.array([0, n - 1], np.int32)
    d = chysx.topology.cloth_system(n, tris, pins)
    o = pyoracle.OracleSim(32)
    o.set_state(pos, np.zeros_like(pos), np.ones(n))
    o.set_mesh(tris)
    o.build_stretch(1.0)
    o.build_shear(1.0)
    o.build_bending(1.0)
    o.set_pins(pins, pos[pins], 1.0)
    s = o.sizes()
    ro, ci = o.pattern()
    assert np.array_equal(d["row_offsets"], ro) and np.array_equal(d["col_indices"], ci)
    assert np.array_equal(d["dihedrals"], o.dihedrals()[0])
    assert np.array_equal(chysx.topology.detect_dihedrals(tris), pyoracle.detect_dihedrals(tris))
    assert np.array_equal(chysx.topology.build_edges(tris), pyoracle.build_edges(tris))
    assert np.array_equal(d["pin_slots"], o.slots("pin", 2))
    assert np.array_equal(d["stretch_slots"], o.slots("stretch", 9 * s["tris"]))
    assert np.array_equal(d["shear_slots"], o.slots("shear", 9 * s["tris"]))
    assert np.array_equal(d["bending_slots"], o.slots("bending", 16 * s["dihedrals"]))


def test_pattern_errors():
    with pytest.raises(IndexError):
        chysx.topology.build_pattern(3, np.array([0, 5], np.int32), np.array([1, 0], np.int32))
    with pytest.raises(ValueError):
        chysx.topology.build_pattern(3, np.array([0], np.int32), np.array([1, 2], np.int32))
    with pytest.raises(IndexError):
        pos, tris = scenes.cloth_grid(3)
        chysx.topology.cloth_system(9, tris, np.array([42], np.int32))


def test_bad_shapes_raise_value_error():
    sim = chysx.ClothSimulator()
    with pytest.raises(ValueError):
        sim.set_mesh(np.zeros((4, 2), np.int32))
    with pytest.raises(ValueError):
        sim.set_pins(np.zeros((2, 2), np.int32), np.zeros((2, 3), np.float32))
    with pytest.raises(ValueError):
        sim.set_pins(np.zeros(2, np.int32), np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError):
        sim.set_mesh(np.full((1, 3), -1, np.int32)[:0].reshape(-1, 4))


def test_host_only_setup_without_device():
    """Material, pins, mesh and knobs are host state (no device needed)."""
    sim = chysx.ClothSimulator()
    m = chysx.ClothMaterial()
    m.gy, m.damping = -9.81, 0.5
    sim.set_material(m)
    assert sim.material.gy == pytest.approx(-9.81) and sim.material.damping == 0.5
    pos, tris = scenes.cloth_grid(4)
    sim.set_mesh(tris)
    sim.set_pins(np.array([0, 3], np.int32), pos[[0, 3]])
    assert sim.num_pins() == 2
    assert np.array_equal(sim.debug_edges(), pyoracle.build_edges(tris))
    sim.set_pcg_iterations(0)
    assert sim.pcg_iterations() == 1  # clamped as in cloth_simulator.h:597-600
    sim.set_pcg_iterations(77)
    assert sim.pcg_iterations() == 77
    sim.set_pcg_tolerance(1e-6)
    assert sim.pcg_tolerance() == 1e-6
    with pytest.raises(ValueError):
        sim.set_lanes_per_row(3)
    with pytest.raises(ValueError):
        sim.update_pin_targets(np.zeros((5, 3), np.float32))
    sim.update_pin_targets(np.ones((2, 3), np.float32))
    sim.clear_pins()
    assert sim.num_pins() == 0
    assert sim.num_particles() == 0 and sim.num_off_diag_blocks() == 0
    diag, ro, ci, vals, rhs, dx = sim.debug_dump_last_solve()
    assert diag.shape == (0, 3, 3) and ro.shape == (1,)
