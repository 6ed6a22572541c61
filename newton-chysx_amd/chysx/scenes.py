"""Cloth scenes and the synthetic workloads of the cloth-step benchmarks.

``cloth_grid`` reproduces the vertex order and triangle winding of Newton's
``ModelBuilder.add_cloth_grid`` (``newton/_src/sim/builder.py:8118-8136``):
vertex (x, y) has index ``y*(dim_x+1) + x`` and each cell (x-1..x, y-1..y)
gives triangles ``[v0, v1, v3]`` and ``[v1, v2, v3]``.  The grid is laid in
the XZ plane at height ``height`` (grid y -> world z).

``setup_simulator`` installs a :class:`ClothSpec` on a ``ClothSimulator`` in
the order ``SolverChysX.__init__`` uses (``newton/_src/solvers/chysx/
solver_chysx.py:349-478``): material, PCG knobs, buffers, mesh, area mass,
stretch, shear, bending, then pins whose targets are the initial positions.

Workloads (all seeded, all synthetic):
  plain_small   32x32 grid, 1 m, jitter N(0, 1e-4 m), BW98 k = E*h = 100,
                bending 1e-4, two corners at max z pinned (k = 1e6),
                g = (0, -9.81, 0), dt = 1/600, density 0.2, PCG tol 1e-6, cap 100
  plain_large   256x256 grid, same, E ~ U[1e4, 1e6] (k = E*h)
  batched       512 grids, side ~ logU[16, 128] vertices, E ~ U[1e4, 1e6],
                density ~ U[0.1, 0.4], two random distinct corners pinned
  precision     512x512 grid, E = 1e6, rest undeformed, then a shear
                x += s (z - z0) with max |s dz| = 5 cm
  serving       Poisson arrivals at 200/s for 60 s; N ~ logU[1e3, 2.5e5]
                (square grids), 30..300 substeps, materials as `batched`
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

THICKNESS = 1.0e-3  # membrane thickness [m]: BW98 stiffness k = E * h


@dataclass
class ClothSpec:
    pos: np.ndarray                 # (N, 3) float32 initial = rest positions
    tris: np.ndarray                # (T, 3) int32
    pins: np.ndarray                # (P,) int32
    k_stretch: float = 100.0
    k_shear: float = 100.0
    k_bend: float = 1.0e-4
    k_pin: float = 1.0e6
    density: float = 0.2
    damping: float = 0.0
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1.0 / 600.0
    max_iter: int = 100
    tol: float = 1.0e-6
    name: str = "cloth"
    pos_start: np.ndarray | None = None  # positions after set-up (None: = pos)
    stvk: tuple | None = None       # StVK membrane (mu [Pa], lambda [Pa], thickness [m]), or None
    planes: tuple = ()              # static planes ((nx, ny, nz), d)
    boxes: tuple = ()               # static boxes (center, half_ext, ex, ey, ez)
    contact: tuple | None = None    # (thickness, stiffness, friction, friction_eps), or None
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def cloth_grid(nx: int, ny: int | None = None, side: float = 1.0, height: float = 1.0):
    """(pos float32 (N,3), tris int32 (T,3)) of an nx-by-ny vertex grid."""
    ny = nx if ny is None else ny
    if nx < 2 or ny < 2:
        raise ValueError("cloth_grid: need at least 2 vertices per side")
    dim_x, dim_y = nx - 1, ny - 1
    cell = side / dim_x
    ys, xs = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.zeros((ny * nx, 3), np.float64)
    pos[:, 0] = xs.ravel() * cell
    pos[:, 1] = height
    pos[:, 2] = ys.ravel() * cell
    cy, cx = np.meshgrid(np.arange(1, ny), np.arange(1, nx), indexing="ij")
    cx, cy = cx.ravel(), cy.ravel()
    v0 = (cy - 1) * nx + (cx - 1)
    v1 = (cy - 1) * nx + cx
    v2 = cy * nx + cx
    v3 = cy * nx + (cx - 1)
    tris = np.empty((2 * len(cx), 3), np.int64)
    tris[0::2] = np.stack([v0, v1, v3], 1)
    tris[1::2] = np.stack([v1, v2, v3], 1)
    return pos.astype(np.float32), tris.astype(np.int32)


def top_corners(nx: int, ny: int | None = None) -> np.ndarray:
    """The two corners at max z (grid row y = ny-1)."""
    ny = nx if ny is None else ny
    return np.array([(ny - 1) * nx, ny * nx - 1], np.int32)


def grid_counts(n: int) -> dict:
    """Closed-form element counts of an n-by-n grid (vertices, triangles,
    edges, interior dihedrals, off-diagonal BSR blocks)."""
    return {
        "N": n * n,
        "T": 2 * (n - 1) ** 2,
        "E": 3 * (n - 1) ** 2 + 2 * (n - 1),
        "D": 3 * (n - 1) ** 2 - 2 * (n - 1),
        "nnz": 12 * n * n - 24 * (n - 1) - 12,
    }


def plain_small(seed: int = 0, n: int = 32) -> ClothSpec:
    rng = np.random.default_rng(seed)
    pos, tris = cloth_grid(n)
    pos = (pos.astype(np.float64) + rng.normal(0.0, 1.0e-4, pos.shape)).astype(np.float32)
    k = 1.0e5 * THICKNESS
    return ClothSpec(pos=pos, tris=tris, pins=top_corners(n), k_stretch=k, k_shear=k, name=f"grid{n}")


def stvk_lame(E: float, nu: float) -> tuple:
    """(mu, lambda) of a membrane: mu = E / (2 (1 + nu)) and the
    plane-stress lambda = E nu / (1 - nu^2) (= 2 mu lambda3 / (2 mu + lambda3)
    of the 3-D Lame lambda3)."""
    return E / (2.0 * (1.0 + nu)), E * nu / (1.0 - nu * nu)


def plain_small_stvk(seed: int = 0, n: int = 32, E: float = 1.0e5, nu: float = 0.3) -> ClothSpec:
    """The 32x32 workload with the StVK membrane (E = 1e5 Pa, nu = 0.3,
    h = 1 mm) in place of the BW98 stretch / shear pair."""
    spec = plain_small(seed=seed, n=n)
    mu, lam = stvk_lame(E, nu)
    spec.k_stretch = 0.0
    spec.k_shear = 0.0
    spec.stvk = (mu, lam, THICKNESS)
    spec.name = f"grid{n}_stvk"
    return spec


def drape(seed: int = 5, n: int = 32, friction: float = 0.0) -> ClothSpec:
    """The 32x32 cloth (no pins) dropped 2 cm onto a static table: an
    oriented box (top at y = 0.98, 0.5 x 0.5 m, turned 0.3 rad about y)
    over a ground plane y = 0.5; contact thickness 5 mm, stiffness 1e4 N/m,
    Coulomb friction `friction` (eps 1e-4 m)."""
    spec = plain_small(seed=seed, n=n)
    c, s_ = math.cos(0.3), math.sin(0.3)
    box = ((0.5, 0.95, 0.5), (0.25, 0.03, 0.25), (c, 0.0, -s_), (0.0, 1.0, 0.0), (s_, 0.0, c))
    return replace(spec, pins=np.zeros(0, np.int32), planes=(((0.0, 1.0, 0.0), -0.5),), boxes=(box,),
                   contact=(5.0e-3, 1.0e4, friction, 1.0e-4), name=f"grid{n}_drape")


def plain_large(seed: int = 1, n: int = 256) -> ClothSpec:
    rng = np.random.default_rng(seed)
    E = float(rng.uniform(1.0e4, 1.0e6))
    spec = plain_small(seed=seed, n=n)
    k = E * THICKNESS
    return replace(spec, k_stretch=k, k_shear=k, meta={"E": E}, name=f"grid{n}")


def batched(seed: int = 2, count: int = 512, lo: int = 16, hi: int = 128) -> list[ClothSpec]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        E = float(rng.uniform(1.0e4, 1.0e6))
        rho = float(rng.uniform(0.1, 0.4))
        corners = np.array([0, n - 1, (n - 1) * n, n * n - 1], np.int32)
        pins = np.sort(rng.choice(corners, size=2, replace=False)).astype(np.int32)
        pos, tris = cloth_grid(n)
        pos = (pos.astype(np.float64) + rng.normal(0.0, 1.0e-4, pos.shape)).astype(np.float32)
        k = E * THICKNESS
        specs.append(ClothSpec(pos=pos, tris=tris, pins=pins, k_stretch=k, k_shear=k, density=rho,
                               name=f"inst{i}_grid{n}", meta={"E": E, "n": n}))
    return specs


def precision(seed: int = 3, n: int = 512, max_shift: float = 0.05) -> ClothSpec:
    rng = np.random.default_rng(seed)
    pos, tris = cloth_grid(n)
    k = 1.0e6 * THICKNESS
    z0 = float(pos[:, 2].min())
    dz = float(pos[:, 2].max()) - z0
    s = float(rng.uniform(0.5, 1.0)) * max_shift / dz
    start = pos.astype(np.float64).copy()
    start[:, 0] += s * (start[:, 2] - z0)
    return ClothSpec(pos=pos, tris=tris, pins=top_corners(n), k_stretch=k, k_shear=k, name=f"grid{n}_shear",
                     pos_start=start.astype(np.float32), meta={"E": 1.0e6, "shear": s})


@dataclass
class JobSpec:
    """A serving job's cloth, built on demand (`build()`): an n_side x
    n_side grid with BW98 stiffness E*h, density rho and two pinned corners."""
    n_side: int
    E: float
    rho: float
    pins: np.ndarray

    @property
    def n(self) -> int:
        return self.n_side * self.n_side

    @property
    def dt(self) -> float:
        return ClothSpec.dt

    def build(self) -> ClothSpec:
        pos, tris = cloth_grid(self.n_side)
        k = self.E * THICKNESS
        return ClothSpec(pos=pos, tris=tris, pins=self.pins, k_stretch=k, k_shear=k, density=self.rho,
                         name=f"job_grid{self.n_side}", meta={"E": self.E})


@dataclass
class Job:
    arrival: float
    n_side: int
    substeps: int
    spec: JobSpec


def serving_jobs(seed: int = 4, rate: float = 200.0, duration: float = 60.0, lo: float = 1.0e3,
                 hi: float = 2.5e5) -> list[Job]:
    """Poisson arrivals at `rate` per second for `duration` seconds; mesh
    sizes log-uniform in [lo, hi] vertices (square grids), 30-300 substeps,
    E ~ U[1e4, 1e6] Pa, density ~ U[0.1, 0.4] kg/m^2, two random corners
    pinned."""
    rng = np.random.default_rng(seed)
    t, jobs = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        N = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        n = max(2, int(round(math.sqrt(N))))
        steps = int(rng.integers(30, 301))
        E = float(rng.uniform(1.0e4, 1.0e6))
        rho = float(rng.uniform(0.1, 0.4))
        corners = np.array([0, n - 1, (n - 1) * n, n * n - 1], np.int32)
        pins = np.sort(rng.choice(corners, size=2, replace=False)).astype(np.int32)
        jobs.append(Job(arrival=t, n_side=n, substeps=steps, spec=JobSpec(n_side=n, E=E, rho=rho, pins=pins)))
    return jobs


def setup_simulator(sim, spec: ClothSpec, pos_ptr: int, vel_ptr: int, inv_mass_ptr: int, stream: int = 0,
                    lanes_per_row: int | None = None):
    """Installs `spec` on a chysx.ClothSimulator bound to device buffers
    holding spec.pos (positions), zeros (velocities) and inverse masses."""
    import chysx

    mat = chysx.ClothMaterial()
    mat.gx, mat.gy, mat.gz = (float(g) for g in spec.gravity)
    mat.damping = float(spec.damping)
    mat.density = float(spec.density)
    sim.set_material(mat)
    sim.set_pcg_iterations(int(spec.max_iter))
    sim.set_pcg_tolerance(float(spec.tol))
    if lanes_per_row is not None:
        sim.set_lanes_per_row(int(lanes_per_row))
    sim.set_external_buffers(pos_ptr=pos_ptr, vel_ptr=vel_ptr, particle_count=spec.n, inv_mass_ptr=inv_mass_ptr)
    sim.set_mesh(spec.tris)
    sim.redistribute_mass_area_weighted(surface_density=float(spec.density), inv_mass_ptr=inv_mass_ptr,
                                        particle_count=spec.n, cuda_stream=stream)
    if spec.k_stretch > 0:
        sim.build_fem_stretch_from_current_positions(stiffness=float(spec.k_stretch), cuda_stream=stream)
    if spec.k_shear > 0:
        sim.build_fem_shear_from_current_positions(stiffness=float(spec.k_shear), cuda_stream=stream)
    if spec.stvk is not None:
        mu, lam, h = spec.stvk
        sim.build_fem_stvk_from_current_positions(mu=float(mu), lam=float(lam), thickness=float(h),
                                                  cuda_stream=stream)
    if spec.k_bend > 0:
        sim.build_bending_from_current_positions(stiffness=float(spec.k_bend), cuda_stream=stream)
    if len(spec.pins):
        sim.set_pins(spec.pins, spec.pos[spec.pins], float(spec.k_pin))
    for n, d in spec.planes:
        sim.add_static_plane(np.asarray(n, np.float32), float(d))
    for c, h, ex, ey, ez in spec.boxes:
        sim.add_static_box(*(np.asarray(a, np.float32) for a in (c, h, ex, ey, ez)))
    if spec.contact is not None:
        th, k, mu, eps = spec.contact
        sim.set_static_contact_thickness(float(th))
        sim.set_static_contact_stiffness(float(k))
        sim.set_static_contact_friction(float(mu))
        sim.set_static_contact_friction_epsilon(float(eps))
    return sim


class DeviceCloth:
    """A ClothSimulator with its own torch-allocated device buffers."""

    def __init__(self, spec: ClothSpec, device: str = "cuda", lanes_per_row: int | None = None, stream=None):
        """With `stream` (a torch.cuda.Stream) every set-up operation runs on
        that stream and waits only for it, so a cloth can be set up on one
        host thread while another steps a batch."""
        import contextlib

        import torch

        import chysx

        self.spec = spec
        ctx = torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()
        sync = stream.synchronize if stream is not None else torch.cuda.synchronize
        with ctx:
            self.pos = torch.from_numpy(np.ascontiguousarray(spec.pos, np.float32)).to(device)
            self.vel = torch.zeros_like(self.pos)
            self.inv_mass = torch.ones(spec.n, dtype=torch.float32, device=device)
            self.sim = chysx.ClothSimulator()
            sync()
            setup_simulator(self.sim, spec, self.pos.data_ptr(), self.vel.data_ptr(), self.inv_mass.data_ptr(),
                            stream=stream.cuda_stream if stream is not None else 0, lanes_per_row=lanes_per_row)
            if spec.pos_start is not None:
                self.pos.copy_(torch.from_numpy(np.ascontiguousarray(spec.pos_start, np.float32)).to(device))
            sync()

    def step(self, dt: float | None = None, stream: int = 0):
        self.sim.step(self.spec.dt if dt is None else dt, stream)

    def state(self):
        import torch

        torch.cuda.synchronize()
        return self.pos.cpu().numpy().astype(np.float64), self.vel.cpu().numpy().astype(np.float64)
