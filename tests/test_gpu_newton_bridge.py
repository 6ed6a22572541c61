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
