"""Shared test helpers: the oracle set up from a ClothSpec exactly the way
chysx.scenes.setup_simulator sets up the product."""
from __future__ import annotations

import numpy as np

import pyoracle


def oracle_from_spec(spec, precision=32, pcg=None):
    o = pyoracle.OracleSim(precision)
    n = spec.n
    o.set_state(spec.pos.astype(np.float64), np.zeros((n, 3)), np.ones(n))
    o.set_material(spec.damping, spec.gravity)
    o.set_pcg(*(pcg if pcg is not None else (spec.max_iter, spec.tol)))
    o.set_mesh(spec.tris)
    o.redistribute_mass(spec.density)
    if spec.k_stretch > 0:
        o.build_stretch(spec.k_stretch)
    if spec.k_shear > 0:
        o.build_shear(spec.k_shear)
    if getattr(spec, "stvk", None) is not None:
        o.build_stvk(*spec.stvk)
    if spec.k_bend > 0:
        o.build_bending(spec.k_bend)
    if len(spec.pins):
        o.set_pins(spec.pins, spec.pos[spec.pins].astype(np.float64), spec.k_pin)
    for n_, d_ in getattr(spec, "planes", ()) or ():
        o.add_static_plane(n_, d_)
    for box in getattr(spec, "boxes", ()) or ():
        o.add_static_box(*box)
    if getattr(spec, "contact", None) is not None:
        o.set_static_contact(*spec.contact)
    if spec.pos_start is not None:
        _, v = o.state()
        o.set_state(spec.pos_start.astype(np.float64), v, o.inv_mass())
    return o


def rel_maxnorm(a, b):
    """max|a-b| / max|b| (relative max-norm)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
