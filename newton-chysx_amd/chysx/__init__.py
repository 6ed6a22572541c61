"""chysx: ChysX's implicit cloth time-step on AMD Instinct MI355X.

Same public surface as the reference's ``chysx`` package for the cloth step
(``ChysX/chysx/__init__.py:36-39``): :class:`ClothMaterial` and
:class:`ClothSimulator`, whose ``step`` runs hand-written gfx950 HIP kernels
through the C ABI in ``include/chysx_hip.h``.  Additions:
:class:`ClothBatch` (many cloths per fused substep), the standalone
:class:`BlockCSR3` / :class:`PCGSolver` pair of the C++ API, and the host
``topology`` functions.

Particle buffers are borrowed device pointers (AoS float3), e.g.
``torch.Tensor.data_ptr()`` of a ``(n, 3)`` float32 tensor on ``cuda:0``::

    import chysx, torch
    pos = torch.zeros(n, 3, device="cuda"); vel = torch.zeros_like(pos)
    sim = chysx.ClothSimulator()
    mat = chysx.ClothMaterial(); mat.gz = -9.81
    sim.set_material(mat)
    sim.set_external_buffers(pos_ptr=pos.data_ptr(), vel_ptr=vel.data_ptr(), particle_count=n)
    sim.step(dt=1.0 / 60.0)
"""
from __future__ import annotations

from ._runtime import share_torch_hip_runtime

share_torch_hip_runtime()

from . import _chysx_native  # noqa: E402

ClothMaterial = _chysx_native.ClothMaterial
ClothSimulator = _chysx_native.ClothSimulator
ClothBatch = _chysx_native.ClothBatch
BlockCSR3 = _chysx_native.BlockCSR3
PCGSolver = _chysx_native.PCGSolver
topology = _chysx_native.topology
io = _chysx_native.io

__version__ = "0.1.0"

__all__ = ["BlockCSR3", "ClothBatch", "ClothMaterial", "ClothSimulator", "PCGSolver", "io", "topology",
           "__version__"]
