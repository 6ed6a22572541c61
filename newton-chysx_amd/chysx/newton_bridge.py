"""Newton-shaped bridge: ``SolverChysX`` over torch-owned device buffers.

The reference's Newton solver (newton/_src/solvers/chysx/solver_chysx.py:
282-478, 963-1039) drives ``chysx.ClothSimulator`` with Warp arrays of a
Newton ``Model`` / ``State``.  Warp has no ROCm backend, so this module keeps
the same construction order and step contract over torch tensors:

* ``ClothModel`` / ``ClothState`` hold what the solver reads of Newton's
  Model / State: ``particle_q``, ``particle_qd``, ``particle_inv_mass``
  (torch float32 CUDA tensors, AoS (N, 3) like Warp ``vec3``),
  ``tri_indices`` ((M, 3) int32), ``gravity``, ``particle_count``,
  ``tri_count``;
* ``cloth_grid_model`` builds the model ``ModelBuilder.add_cloth_grid``
  builds for a flat grid (newton/_src/sim/builder.py:8063-8187: vertex
  index y*(dim_x+1)+x, triangles [v0, v1, v3], [v1, v2, v3], per-particle
  mass);
* ``SolverChysX(model, ...)`` installs material -> PCG iterations ->
  external buffers -> mesh -> area-weighted mass (writes the model's
  inv_mass) -> stretch -> shear -> bending -> pins (targets = initial q),
  and ``step(state_in, state_out, control, contacts, dt)`` seeds state_out
  from state_in when they differ, rebinds the pointers to state_out and
  steps on the caller's current torch stream.

``static_contact_enabled`` registers the model's world-static planes and
boxes (``shape_type`` / ``shape_body`` / ``shape_transform`` /
``shape_scale``, Newton's layout: a plane's normal is its local +Z, a box's
frame columns are its rotation's columns) with the simulator's static
contact set (solver_chysx.py:550-640).  Features of the reference solver
that are not on the implicit cloth step's path (self-collision, SDF /
mesh-body contact, untangling, tetrahedral FEM, the VBD solver) raise
``NotImplementedError`` when enabled.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence

import numpy as np


# newton.GeoType values of the primitives the static contact set takes
GEO_PLANE = 1
GEO_BOX = 7


def quat_to_matrix(q) -> np.ndarray:
    """Rotation matrix of a unit quaternion (x, y, z, w)."""
    x, y, z, w = (float(c) for c in q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


@dataclass
class ClothState:
    particle_q: object   # torch (N, 3) float32 on the device
    particle_qd: object  # torch (N, 3) float32 on the device

    def clear_forces(self):
        """External forces are not read by the PCG step (the reference
        ignores ClothBuffers::force there too)."""


@dataclass
class ClothModel:
    particle_q: object
    particle_qd: object
    particle_inv_mass: object | None
    tri_indices: np.ndarray          # (M, 3) int32
    gravity: tuple = (0.0, 0.0, -9.81)
    # Newton's collision shapes (numpy): type (GEO_PLANE / GEO_BOX ...), owning
    # body (-1: world-static), transform (px, py, pz, qx, qy, qz, qw), scale
    shape_type: np.ndarray | None = None
    shape_body: np.ndarray | None = None
    shape_transform: np.ndarray | None = None
    shape_scale: np.ndarray | None = None

    @property
    def shape_count(self) -> int:
        return 0 if self.shape_type is None else int(len(self.shape_type))

    @property
    def particle_count(self) -> int:
        return int(self.particle_q.shape[0])

    @property
    def tri_count(self) -> int:
        return int(self.tri_indices.shape[0]) if self.tri_indices is not None else 0

    @property
    def device(self):
        return self.particle_q.device

    def state(self) -> ClothState:
        return ClothState(self.particle_q.clone(), self.particle_qd.clone())


def cloth_grid_model(pos=(0.0, 0.0, 0.0), dim_x: int = 10, dim_y: int = 10, cell_x: float = 0.1,
                     cell_y: float = 0.1, mass: float = 1.0, vel=(0.0, 0.0, 0.0), gravity=(0.0, 0.0, -9.81),
                     device: str = "cuda") -> ClothModel:
    """A flat (dim_x+1) x (dim_y+1) particle grid in the model's XY plane at
    `pos` (identity rotation), the layout and winding of
    ModelBuilder.add_cloth_grid (builder.py:8118-8136), `mass` per particle."""
    import torch

    ys, xs = np.meshgrid(np.arange(dim_y + 1), np.arange(dim_x + 1), indexing="ij")
    q = np.zeros(((dim_y + 1) * (dim_x + 1), 3), np.float64)
    q[:, 0] = xs.ravel() * cell_x + pos[0]
    q[:, 1] = ys.ravel() * cell_y + pos[1]
    q[:, 2] = pos[2]
    tris = []
    w = dim_x + 1
    for y in range(1, dim_y + 1):
        for x in range(1, dim_x + 1):
            v0, v1 = (y - 1) * w + (x - 1), (y - 1) * w + x
            v2, v3 = y * w + x, y * w + (x - 1)
            tris.append((v0, v1, v3))
            tris.append((v1, v2, v3))
    n = len(q)
    qt = torch.from_numpy(q.astype(np.float32)).to(device)
    qd = torch.tensor(np.tile(np.asarray(vel, np.float32), (n, 1)), device=device)
    inv_m = torch.full((n,), 1.0 / mass if mass > 0 else 0.0, dtype=torch.float32, device=device)
    return ClothModel(qt, qd, inv_m, np.asarray(tris, np.int32).reshape(-1, 3), tuple(float(g) for g in gravity))


class SolverChysX:
    """The reference solver's constructor and step over torch buffers."""

    def __init__(self, model: ClothModel, gravity: tuple | None = None, damping: float = 0.0,
                 spring_stiffness: float = 0.0, fem_stretch_stiffness: float = 0.0, fem_shear_stiffness: float = 0.0,
                 bending_stiffness: float = 0.0, pin_indices: Sequence[int] | None = None, pin_stiffness: float = 1.0e6,
                 pcg_iterations: int = 50, surface_density: float | None = None,
                 self_collision_enabled: bool = False, static_contact_enabled: bool = False,
                 static_contact_thickness: float = 0.0, static_contact_stiffness: float = 1.0e4,
                 static_contact_friction: float = 0.0, static_contact_friction_epsilon: float = 1.0e-4,
                 untangle_enabled: bool = False, tet_fem_enabled: bool = False, solver_type: str = "pcg",
                 **_unused):
        import chysx

        for flag, name in ((self_collision_enabled, "self_collision_enabled"),
                           (untangle_enabled, "untangle_enabled"), (tet_fem_enabled, "tet_fem_enabled")):
            if flag:
                raise NotImplementedError(f"SolverChysX: {name} is not implemented by this library (implicit cloth step only)")
        if solver_type.lower() != "pcg":
            raise NotImplementedError("SolverChysX: only solver_type='pcg' is implemented by this library")
        if spring_stiffness != 0.0:
            import warnings

            warnings.warn("SolverChysX(spring_stiffness=...) is deprecated and ignored (as in the reference)",
                          DeprecationWarning, stacklevel=2)
        self.model = model
        if gravity is None:
            gravity = model.gravity
        material = chysx.ClothMaterial()
        material.gx, material.gy, material.gz = (float(g) for g in gravity)
        material.damping = float(damping)
        self._sim = chysx.ClothSimulator()
        self._sim.set_material(material)
        self._sim.set_pcg_iterations(int(pcg_iterations))
        n = model.particle_count
        inv_mass_ptr = model.particle_inv_mass.data_ptr() if model.particle_inv_mass is not None else 0
        if n > 0:
            self._sim.set_external_buffers(pos_ptr=model.particle_q.data_ptr(), vel_ptr=model.particle_qd.data_ptr(),
                                           particle_count=n, inv_mass_ptr=inv_mass_ptr)
        wants_mesh = (fem_stretch_stiffness > 0.0 or fem_shear_stiffness > 0.0 or bending_stiffness > 0.0
                      or surface_density is not None)
        if wants_mesh and model.tri_count > 0:
            self._sim.set_mesh(np.ascontiguousarray(model.tri_indices, np.int32))
            if surface_density is not None and surface_density > 0.0 and model.particle_inv_mass is not None:
                self._sim.redistribute_mass_area_weighted(surface_density=float(surface_density),
                                                          inv_mass_ptr=inv_mass_ptr, particle_count=n)
            if fem_stretch_stiffness > 0.0:
                self._sim.build_fem_stretch_from_current_positions(stiffness=float(fem_stretch_stiffness))
            if fem_shear_stiffness > 0.0:
                self._sim.build_fem_shear_from_current_positions(stiffness=float(fem_shear_stiffness))
            if bending_stiffness > 0.0:
                self._sim.build_bending_from_current_positions(stiffness=float(bending_stiffness))
        self._sim.set_solver_type(0)
        self._pin_indices_np = None
        if pin_indices is not None and len(pin_indices) > 0 and n > 0:
            idx = np.asarray(list(pin_indices), dtype=np.int32)
            if idx.ndim != 1:
                raise ValueError("pin_indices must be a 1-D iterable of ints")
            if (idx < 0).any() or (idx >= n).any():
                raise ValueError(f"pin_indices out of range [0, {n})")
            q0 = model.particle_q.detach().cpu().numpy().reshape(-1, 3)
            self._sim.set_pins(idx, np.ascontiguousarray(q0[idx], np.float32), float(pin_stiffness))
            self._pin_indices_np = idx
        if static_contact_enabled and n > 0:
            if static_contact_thickness <= 0.0:
                raise ValueError("SolverChysX(static_contact_enabled=True) requires static_contact_thickness > 0")
            if static_contact_friction_epsilon <= 0.0:
                raise ValueError("SolverChysX: static_contact_friction_epsilon must be positive")
            self._sim.set_static_contact_thickness(float(static_contact_thickness))
            self._sim.set_static_contact_stiffness(float(static_contact_stiffness))
            self._sim.set_static_contact_friction(float(static_contact_friction))
            self._sim.set_static_contact_friction_epsilon(float(static_contact_friction_epsilon))
            self._register_static_shapes()

    def _register_static_shapes(self) -> None:
        """World-static planes and boxes of the model into the contact set:
        a plane's normal is its transform's local +Z, d = -dot(n, position);
        a box's frame columns are its rotation's columns, half extents its
        scale."""
        m = self.model
        if m.shape_count == 0:
            return
        body = m.shape_body if m.shape_body is not None else np.full(m.shape_count, -1)
        xf = np.asarray(m.shape_transform, np.float64).reshape(m.shape_count, 7)
        scale = np.asarray(m.shape_scale, np.float64).reshape(m.shape_count, 3)
        for s in range(m.shape_count):
            if int(body[s]) != -1:
                continue  # moving shapes are not static primitives
            R = quat_to_matrix(xf[s, 3:7])
            pos = xf[s, :3]
            t = int(m.shape_type[s])
            if t == GEO_PLANE:
                nrm = R[:, 2]
                self._sim.add_static_plane(n=nrm.astype(np.float32), d=float(-np.dot(nrm, pos)))
            elif t == GEO_BOX:
                self._sim.add_static_box(center=pos.astype(np.float32), half_ext=scale[s].astype(np.float32),
                                         ex=R[:, 0].astype(np.float32), ey=R[:, 1].astype(np.float32),
                                         ez=R[:, 2].astype(np.float32))

    @property
    def simulator(self):
        return self._sim

    def update_pin_targets(self, targets) -> None:
        if self._pin_indices_np is None:
            raise RuntimeError("SolverChysX.update_pin_targets: no pins were configured at construction time")
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        if targets.shape != (self._pin_indices_np.shape[0], 3):
            raise ValueError(f"targets must have shape ({self._pin_indices_np.shape[0]}, 3), got {targets.shape}")
        self._sim.update_pin_targets(targets)

    def step(self, state_in: ClothState, state_out: ClothState, control, contacts, dt: float) -> None:
        import torch

        del control, contacts
        n = self.model.particle_count
        if n == 0:
            return
        if state_out.particle_q.device.type != "cuda":
            raise RuntimeError("SolverChysX requires a GPU device; the chysx kernels only run on the GPU")
        if state_out is not state_in:
            state_out.particle_q.copy_(state_in.particle_q)
            state_out.particle_qd.copy_(state_in.particle_qd)
        inv_mass_ptr = self.model.particle_inv_mass.data_ptr() if self.model.particle_inv_mass is not None else 0
        self._sim.set_external_buffers(pos_ptr=state_out.particle_q.data_ptr(),
                                       vel_ptr=state_out.particle_qd.data_ptr(), particle_count=n,
                                       inv_mass_ptr=inv_mass_ptr)
        self._sim.step(dt=float(dt), cuda_stream=torch.cuda.current_stream().cuda_stream)
