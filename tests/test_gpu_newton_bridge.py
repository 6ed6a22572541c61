"""The Newton-shaped bridge (chysx.newton_bridge.SolverChysX over torch
buffers) against the assertions of the reference's own examples:
example_chysx_freefall.py:116-178 and example_chysx_hanging_cloth.py:193-227
(same scenes, step loops and checks)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(solver, model, frames, dt):
    s0, s1 = model.state(), model.state()
    for _ in range(frames):
        s0.clear_forces()
        solver.step(s0, s1, None, None, dt)
        s0, s1 = s1, s0
    import torch

    torch.cuda.synchronize()
    return s0.particle_q.cpu().numpy().astype(np.float64), s0.particle_qd.cpu().numpy().astype(np.float64)


def test_freefall_example(gpu):
    from chysx.newton_bridge import SolverChysX, cloth_grid_model

    z0 = 2.0
    model = cloth_grid_model(pos=(-0.5, -0.5, z0), dim_x=10, dim_y=10, cell_x=0.1, cell_y=0.1, mass=0.05)
    pinned = [0]
    solver = SolverChysX(model, pin_indices=pinned)
    targets = model.particle_q.cpu().numpy()[pinned].copy()
    fps, frames = 60, 120
    q, qd = _run(solver, model, frames, 1.0 / fps)
    t = frames / fps
    z_exp, vz_exp = z0 - 0.5 * 9.81 * t * t, -9.81 * t
    for k, idx in enumerate(pinned):
        assert np.abs(q[idx] - targets[k]).max() < 1e-5 and np.abs(qd[idx]).max() < 1e-5
    tol_z = max(0.05 * abs(z_exp - z0), 0.05)
    tol_v = max(0.05 * abs(vz_exp), 0.05)
    free = [i for i in range(model.particle_count) if i not in pinned]
    assert np.all(np.abs(q[free, 2] - z_exp) < tol_z)
    assert np.all(np.abs(qd[free, 2] - vz_exp) < tol_v)
    assert np.all(np.abs(q[free, :2]) < 2.0) and np.all(np.abs(qd[free, :2]) < 1e-4)


def test_hanging_cloth_example(gpu):
    from chysx.newton_bridge import SolverChysX, cloth_grid_model

    z0, dim = 2.0, 100
    model = cloth_grid_model(pos=(-0.5, -0.5, z0), dim_x=dim, dim_y=dim, cell_x=0.01, cell_y=0.01, mass=1.0)
    pinned = [0, dim]
    solver = SolverChysX(model, fem_stretch_stiffness=1.0e2, fem_shear_stiffness=1.0e2, bending_stiffness=5.0e-4,
                         damping=0.0, pin_indices=pinned, pin_stiffness=1.0e9, pcg_iterations=50, surface_density=0.1)
    targets = model.particle_q.cpu().numpy()[pinned].copy()
    far = dim * (dim + 1) + dim // 2
    q, qd = _run(solver, model, 240, 1.0 / 100)
    assert np.isfinite(q).all() and np.isfinite(qd).all()
    for k, idx in enumerate(pinned):
        assert np.linalg.norm(q[idx] - targets[k]) <= 1e-3
    assert q[:, 2].mean() < z0
    assert z0 - q[far, 2] >= 0.5
    assert np.linalg.norm(qd, axis=1).max() <= 20.0


def test_unsupported_features_raise(gpu):
    from chysx.newton_bridge import SolverChysX, cloth_grid_model

    model = cloth_grid_model(dim_x=2, dim_y=2)
    with pytest.raises(NotImplementedError):
        SolverChysX(model, self_collision_enabled=True)
    with pytest.raises(NotImplementedError):
        SolverChysX(model, solver_type="vbd")
    with pytest.raises(ValueError):
        SolverChysX(model, pin_indices=[100])


def test_table_drop_example(gpu):
    """example_chysx_cloth_drop.py's scene without self-collision (not
    implemented here): a 21x21, 1 m cloth dropped 0.5 m onto a 0.6 m table over
    a ground plane through static_contact_enabled; the example's test_final
    checks after 2 s, plus: the centre rests on the table top and the
    overhanging edges hang below it."""
    from chysx.newton_bridge import GEO_BOX, GEO_PLANE, SolverChysX, cloth_grid_model

    dim, size = 21, 1.0
    cell = size / (dim - 1)
    top = 0.5
    model = cloth_grid_model(pos=(-0.5 * size, -0.5 * size, top + 0.5), dim_x=dim - 1, dim_y=dim - 1, cell_x=cell,
                             cell_y=cell, mass=0.0)
    model.shape_type = np.array([GEO_BOX, GEO_PLANE])
    model.shape_body = np.array([-1, -1])
    model.shape_transform = np.array([[0.0, 0.0, top - 0.025, 0, 0, 0, 1], [0, 0, 0, 0, 0, 0, 1]], np.float64)
    model.shape_scale = np.array([[0.3, 0.3, 0.025], [1.0, 1.0, 1.0]])
    thickness = 0.5 * cell
    solver = SolverChysX(model, damping=0.05, fem_stretch_stiffness=5.0e2, fem_shear_stiffness=5.0e2,
                         bending_stiffness=1.0e-4, pcg_iterations=30, surface_density=0.3,
                         static_contact_enabled=True, static_contact_thickness=thickness,
                         static_contact_stiffness=1.0e4)
    assert solver.simulator.static_box_count() == 1 and solver.simulator.static_plane_count() == 1
    q, qd = _run(solver, model, 200, 1.0 / 100)
    assert np.isfinite(q).all() and np.isfinite(qd).all()
    assert np.abs(q).max() <= 5.0
    assert q[:, 2].min() >= -1.5 * thickness
    assert np.linalg.norm(qd, axis=1).max() <= 25.0
    centre = (dim // 2) * dim + dim // 2
    assert top - 1.5 * thickness <= q[centre, 2] <= top + 1.5 * thickness
    corners = [0, dim - 1, dim * (dim - 1), dim * dim - 1]
    assert np.all(q[corners, 2] < top - 0.05)
