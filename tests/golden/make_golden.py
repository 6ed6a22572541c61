"""Writes tests/golden/grid32_*.npz: regression vectors of the 32x32 cloth
workload (seed 0) produced by the CPU oracle.

grid32_topology.npz    edges, dihedrals (reference order), BSR pattern, slot
                       tables of pins / stretch / shear / bending
grid32_trajectory.npz  FP64-oracle positions after 10, 100 and 600 substeps
                       and the PCG iteration count of every substep (tol 1e-6,
                       cap 100), plus the FP32 oracle's iteration counts

The reference publishes no numbers for this path; these files pin the
oracle against later changes and give the GPU tests FP64 answers without
re-running 600 FP64 substeps.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "tests"), os.path.join(REPO, "newton-chysx_amd")]

import helpers  # noqa: E402
from chysx import scenes  # noqa: E402


def main():
    spec = scenes.plain_small(seed=0)
    o = helpers.oracle_from_spec(spec, precision=32)
    s = o.sizes()
    ro, ci = o.pattern()
    dih, rest = o.dihedrals()
    np.savez_compressed(
        os.path.join(HERE, "grid32_topology.npz"),
        tris=spec.tris, pins=spec.pins, edges=o.edges(), dihedrals=dih, row_offsets=ro, col_indices=ci,
        pin_slots=o.slots("pin", len(spec.pins)), stretch_slots=o.slots("stretch", 9 * s["tris"]),
        shear_slots=o.slots("shear", 9 * s["tris"]), bending_slots=o.slots("bending", 16 * s["dihedrals"]))
    o64 = helpers.oracle_from_spec(spec, precision=64)
    it64, it32, ckpt = [], [], {}
    for k in range(1, 601):
        it64.append(o64.step(spec.dt))
        it32.append(o.step(spec.dt))
        if k in (10, 100, 600):
            ckpt[f"pos64_{k}"] = o64.state()[0]
            ckpt[f"pos32_{k}"] = o.state()[0]
    np.savez_compressed(os.path.join(HERE, "grid32_trajectory.npz"), pos0=spec.pos, iters64=np.array(it64),
                        iters32=np.array(it32), **ckpt)
    print("iterations fp64: mean %.1f max %d; fp32 mean %.1f max %d" % (np.mean(it64), max(it64), np.mean(it32),
                                                                          max(it32)))


if __name__ == "__main__":
    main()
