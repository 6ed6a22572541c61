"""ctypes wrapper of the CPU oracle — TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of ChysX's implicit cloth step
(``cloth_oracle.hpp``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module; the product package
never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_cpu.so")
_lib = None

_I = ctypes.c_int
_D = ctypes.c_double
_P = ctypes.c_void_p


def build(force: bool = False) -> str:
    """Compile liboracle_cpu.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_sim_create.restype = _P
        L.orc_sim_create.argtypes = [_I]
        L.orc_sim_destroy.argtypes = [_P]
        for name in (
            "orc_sim_set_state", "orc_sim_get_state", "orc_sim_get_inv_mass", "orc_sim_set_material",
            "orc_sim_set_mesh", "orc_sim_redistribute_mass", "orc_sim_build_stretch", "orc_sim_build_shear",
            "orc_sim_build_bending", "orc_sim_set_pins", "orc_sim_set_pcg", "orc_sim_step", "orc_sim_run",
            "orc_sim_last_stats", "orc_sim_sizes", "orc_sim_get_pattern", "orc_sim_get_slots",
            "orc_sim_get_dihedrals", "orc_sim_get_edges", "orc_sim_get_rest_shape", "orc_sim_get_system",
            "orc_sim_set_dx", "orc_sim_energy", "orc_sim_gradient", "orc_sim_hessian", "orc_sim_build_stvk",
            "orc_stvk_element", "orc_box_sdf", "orc_sim_add_plane", "orc_sim_add_box", "orc_sim_set_contact",
            "orc_contact_bake", "orc_contact_apply", "orc_pcg_solve_contact", "orc_sim_update_pin_targets",
            "orc_pcg_solve_rho", "orc_sim_get_dx",
        ):
            getattr(L, name).restype = _I
        L.orc_sim_step.argtypes = [_P, _D]
        L.orc_sim_build_stretch.argtypes = [_P, _D]
        L.orc_sim_build_shear.argtypes = [_P, _D]
        L.orc_sim_build_bending.argtypes = [_P, _D]
        L.orc_sim_build_stvk.argtypes = [_P, _D, _D, _D]
        L.orc_stvk_element.argtypes = [_I, _P, _P, _D, _D, _I, _P, _P, _P]
        L.orc_sim_redistribute_mass.argtypes = [_P, _D]
        L.orc_sim_set_material.argtypes = [_P, _D, _D, _D, _D]
        L.orc_sim_set_pcg.argtypes = [_P, _I, _D]
        L.orc_sim_run.argtypes = [_P, _D, _I, _P, _P]
        L.orc_box_sdf.argtypes = [_P, _P, _P, _P]
        L.orc_ipc_f1_over_x.restype = _D
        L.orc_ipc_f1_over_x.argtypes = [_D, _D]
        L.orc_sim_add_plane.argtypes = [_P, _P, _D]
        L.orc_sim_add_box.argtypes = [_P, _P]
        L.orc_sim_set_contact.argtypes = [_P, _D, _D, _D, _D]
        L.orc_contact_bake.argtypes = [_I, _P, _P, _P, _D, _D, _D, _D, _P]
        L.orc_contact_apply.argtypes = [_I, _P, _P, _P, _D, _D, _D, _D, _P, _P]
        L.orc_pcg_solve_contact.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _D, _D, _D,
                                            _P, _P, _P]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def _check(rc: int):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def build_edges(tris) -> np.ndarray:
    t = _i32(tris).reshape(-1, 3)
    out = np.zeros((max(1, 3 * len(t)), 2), np.int32)
    cnt = _I()
    _check(lib().orc_build_edges(_ptr(t), _I(len(t)), _ptr(out), ctypes.byref(cnt)))
    return out[: cnt.value].copy()


def detect_dihedrals(tris) -> np.ndarray:
    t = _i32(tris).reshape(-1, 3)
    out = np.zeros((max(1, 3 * len(t)), 4), np.int32)
    cnt = _I()
    _check(lib().orc_detect_dihedrals(_ptr(t), _I(len(t)), _ptr(out), ctypes.byref(cnt)))
    return out[: cnt.value].copy()


class OracleSim:
    """Stateful CPU restatement of ``ClothSimulator`` (precision 32 or 64)."""

    def __init__(self, precision: int = 32):
        self.L = lib()
        self.h = ctypes.c_void_p(self.L.orc_sim_create(precision))
        self.precision = precision
        self.n = 0

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.orc_sim_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # setup -----------------------------------------------------------------
    def set_state(self, pos, vel, inv_mass=None):
        p = _f64(pos).reshape(-1, 3)
        v = _f64(vel).reshape(-1, 3)
        self.n = len(p)
        w = None if inv_mass is None else _f64(inv_mass)
        _check(self.L.orc_sim_set_state(self.h, _I(self.n), _ptr(p), _ptr(v), None if w is None else _ptr(w)))

    def set_material(self, damping=0.0, gravity=(0.0, 0.0, -9.81)):
        _check(self.L.orc_sim_set_material(self.h, damping, *gravity))

    def set_mesh(self, tris):
        t = _i32(tris).reshape(-1, 3)
        _check(self.L.orc_sim_set_mesh(self.h, _ptr(t), _I(len(t))))

    def redistribute_mass(self, rho):
        _check(self.L.orc_sim_redistribute_mass(self.h, rho))

    def build_stretch(self, k):
        _check(self.L.orc_sim_build_stretch(self.h, k))

    def build_shear(self, k):
        _check(self.L.orc_sim_build_shear(self.h, k))

    def build_stvk(self, mu, lam, thickness):
        _check(self.L.orc_sim_build_stvk(self.h, mu, lam, thickness))

    def build_bending(self, k):
        _check(self.L.orc_sim_build_bending(self.h, k))

    def update_pin_targets(self, targets):
        t = _f64(targets).reshape(-1, 3)
        _check(self.L.orc_sim_update_pin_targets(self.h, _ptr(t), _I(len(t))))

    def set_pins(self, idx, targets, k=1.0e6):
        i = _i32(idx)
        t = _f64(targets).reshape(-1, 3)
        _check(self.L.orc_sim_set_pins(self.h, _ptr(i), _ptr(t), _I(len(i)), _D(k)))

    # static plane / box contact (ChysX/src/collision/static_contact.h)
    def add_static_plane(self, n, d):
        nn = _f64(n).reshape(3)
        _check(self.L.orc_sim_add_plane(self.h, _ptr(nn), _D(d)))

    def add_static_box(self, center, half_ext, ex, ey, ez):
        b = np.concatenate([_f64(a).reshape(3) for a in (center, half_ext, ex, ey, ez)])
        _check(self.L.orc_sim_add_box(self.h, _ptr(b)))

    def set_static_contact(self, thickness, stiffness, friction=0.0, friction_epsilon=1.0e-4):
        _check(self.L.orc_sim_set_contact(self.h, thickness, stiffness, friction, friction_epsilon))

    def set_pcg(self, max_iter=50, tol=0.0):
        _check(self.L.orc_sim_set_pcg(self.h, max_iter, tol))

    # stepping ----------------------------------------------------------------
    def step(self, dt) -> int:
        it = self.L.orc_sim_step(self.h, dt)
        if it < 0:
            _check(-1)
        return it

    def run(self, dt, steps):
        sec = _D()
        tot = ctypes.c_longlong()
        _check(self.L.orc_sim_run(self.h, dt, steps, ctypes.byref(sec), ctypes.byref(tot)))
        return sec.value, tot.value

    def last_stats(self):
        it, rr, bb = _I(), _D(), _D()
        _check(self.L.orc_sim_last_stats(self.h, ctypes.byref(it), ctypes.byref(rr), ctypes.byref(bb)))
        return it.value, rr.value, bb.value

    # readback ----------------------------------------------------------------
    def state(self):
        p = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        _check(self.L.orc_sim_get_state(self.h, _ptr(p), _ptr(v)))
        return p, v

    def inv_mass(self):
        w = np.zeros(self.n)
        _check(self.L.orc_sim_get_inv_mass(self.h, _ptr(w)))
        return w

    def sizes(self):
        vals = [_I() for _ in range(5)]
        _check(self.L.orc_sim_sizes(self.h, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("n", "nnz", "edges", "dihedrals", "tris"), (v.value for v in vals)))

    def pattern(self):
        s = self.sizes()
        ro = np.zeros(s["n"] + 1, np.int32)
        ci = np.zeros(max(1, s["nnz"]), np.int32)
        _check(self.L.orc_sim_get_pattern(self.h, _ptr(ro), _ptr(ci)))
        return ro, ci[: s["nnz"]].copy()

    def slots(self, kind: str, count: int) -> np.ndarray:
        k = {"pin": 0, "stretch": 1, "shear": 2, "bending": 3, "stvk": 4}[kind]
        out = np.zeros(max(1, count), np.int32)
        _check(self.L.orc_sim_get_slots(self.h, k, _ptr(out)))
        return out[:count].copy()

    def dihedrals(self):
        d = self.sizes()["dihedrals"]
        out = np.zeros((max(1, d), 4), np.int32)
        rest = np.zeros(max(1, d))
        _check(self.L.orc_sim_get_dihedrals(self.h, _ptr(out), _ptr(rest)))
        return out[:d].copy(), rest[:d].copy()

    def edges(self):
        e = self.sizes()["edges"]
        out = np.zeros((max(1, e), 2), np.int32)
        _check(self.L.orc_sim_get_edges(self.h, _ptr(out)))
        return out[:e].copy()

    def rest_shape(self, kind: str, n_tris: int):
        dm = np.zeros((max(1, n_tris), 2, 2))
        ar = np.zeros(max(1, n_tris))
        _check(self.L.orc_sim_get_rest_shape(self.h, 1 if kind == "stretch" else 2, _ptr(dm), _ptr(ar)))
        return dm[:n_tris].copy(), ar[:n_tris].copy()

    def system(self):
        s = self.sizes()
        n, nnz = s["n"], s["nnz"]
        diag = np.zeros((n, 3, 3))
        vals = np.zeros((max(1, nnz), 3, 3))
        rhs = np.zeros((n, 3))
        dx = np.zeros((n, 3))
        _check(self.L.orc_sim_get_system(self.h, _ptr(diag), _ptr(vals), _ptr(rhs), _ptr(dx)))
        return diag, vals[:nnz].copy(), rhs, dx

    def dx(self):
        """The warm start / last solution dx (zeros before the first step)."""
        d = np.zeros((self.n, 3))
        _check(self.L.orc_sim_get_dx(self.h, _ptr(d)))
        return d

    def set_dx(self, dx):
        d = _f64(dx).reshape(-1, 3)
        _check(self.L.orc_sim_set_dx(self.h, _ptr(d)))

    def energy(self) -> float:
        e = _D()
        _check(self.L.orc_sim_energy(self.h, ctypes.byref(e)))
        return e.value

    def gradient(self) -> np.ndarray:
        g = np.zeros((self.n, 3))
        _check(self.L.orc_sim_gradient(self.h, _ptr(g)))
        return g

    def hessian(self):
        s = self.sizes()
        diag = np.zeros((s["n"], 3, 3))
        vals = np.zeros((max(1, s["nnz"]), 3, 3))
        _check(self.L.orc_sim_hessian(self.h, _ptr(diag), _ptr(vals)))
        return diag, vals[: s["nnz"]].copy()


def pcg_solve(row_offsets, cols, diag, values, b, x0, max_iter, tol=0.0, precision=32, return_rho=False):
    """Standalone block-Jacobi PCG of the oracle; returns (x, iters, rr, bb),
    plus the final <r, z> with return_rho."""
    ro = _i32(row_offsets)
    ci = _i32(cols) if len(cols) else np.zeros(1, np.int32)
    dg = _f64(diag).reshape(-1, 9)
    vl = _f64(values).reshape(-1, 9) if len(values) else np.zeros((1, 9))
    bb_ = _f64(b).reshape(-1, 3)
    x = _f64(x0).reshape(-1, 3).copy()
    it, rr, bb, rho = _I(), _D(), _D(), _D()
    n = len(ro) - 1
    _check(lib().orc_pcg_solve_rho(_I(precision), _I(n), _ptr(ro), _ptr(ci), _ptr(dg), _ptr(vl), _ptr(bb_), _ptr(x),
                                   _I(max_iter), _D(tol), ctypes.byref(it), ctypes.byref(rr), ctypes.byref(bb),
                                   ctypes.byref(rho)))
    if return_rho:
        return x, it.value, rr.value, bb.value, rho.value
    return x, it.value, rr.value, bb.value


def contact_pcg_solve(row_offsets, cols, diag, values, b, x0, max_iter, tol, pairs, weights, k, slips=None, mu=0.0,
                      eps=1e-4, precision=32):
    """The standalone PCG on (A + C_off) (diag already holds the baked
    same-slot terms); returns (x, iters, rr, bb)."""
    ro = _i32(row_offsets)
    ci = _i32(cols) if len(cols) else np.zeros(1, np.int32)
    dg = _f64(diag).reshape(-1, 9)
    vl = _f64(values).reshape(-1, 9) if len(values) else np.zeros((1, 9))
    bb_ = _f64(b).reshape(-1, 3)
    x = _f64(x0).reshape(-1, 3).copy()
    pr = _i32(pairs).reshape(-1, 4)
    w = _f64(weights).reshape(-1, 8)
    sl = None if slips is None else _f64(slips).reshape(-1, 4)
    it, rr, bb = _I(), _D(), _D()
    n = len(ro) - 1
    _check(lib().orc_pcg_solve_contact(_I(precision), _I(n), _ptr(ro), _ptr(ci), _ptr(dg), _ptr(vl), _ptr(bb_),
                                       _ptr(x), _I(max_iter), _D(tol), _I(len(pr)), _ptr(pr), _ptr(w),
                                       None if sl is None else _ptr(sl), _D(k), _D(mu), _D(eps), ctypes.byref(it),
                                       ctypes.byref(rr), ctypes.byref(bb)))
    return x, it.value, rr.value, bb.value


def spmv(row_offsets, cols, diag, values, x, precision=32):
    ro = _i32(row_offsets)
    ci = _i32(cols) if len(cols) else np.zeros(1, np.int32)
    dg = _f64(diag).reshape(-1, 9)
    vl = _f64(values).reshape(-1, 9) if len(values) else np.zeros((1, 9))
    xv = _f64(x).reshape(-1, 3)
    n = len(ro) - 1
    y = np.zeros((n, 3))
    _check(lib().orc_spmv(_I(precision), _I(n), _ptr(ro), _ptr(ci), _ptr(dg), _ptr(vl), _ptr(xv), _ptr(y)))
    return y


def set_num_threads(n: int):
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def stvk_element(x, dinv, muA, lamA, exact=False, precision=64):
    """Energy, gradient (3, 3) and Hessian blocks (3, 3, 3, 3) [a, b, i, j] of
    one StVK triangle with vertices x (3, 3) and rest shape Dm^-1 (2, 2)."""
    x = _f64(x).reshape(9)
    dinv = _f64(dinv).reshape(4)
    E = ctypes.c_double()
    g = np.zeros(9)
    H = np.zeros(81)
    _check(lib().orc_stvk_element(int(precision), _ptr(x), _ptr(dinv), float(muA), float(lamA), int(bool(exact)),
                                  ctypes.byref(E), _ptr(g), _ptr(H)))
    return E.value, g.reshape(3, 3), H.reshape(3, 3, 3, 3)


def box_sdf(p, center, half_ext, ex=(1.0, 0.0, 0.0), ey=(0.0, 1.0, 0.0), ez=(0.0, 0.0, 1.0)):
    """(signed distance, outward normal) of point p to an oriented box (FP64)."""
    b = np.concatenate([_f64(a).reshape(3) for a in (center, half_ext, ex, ey, ez)])
    pp = _f64(p).reshape(3)
    d = np.zeros(1)
    n = np.zeros(3)
    _check(lib().orc_box_sdf(_ptr(pp), _ptr(b), _ptr(d), _ptr(n)))
    return float(d[0]), n


def ipc_f1_over_x(u, eps):
    return float(lib().orc_ipc_f1_over_x(float(u), float(eps)))


def contact_bake(diag, pairs, weights, k, slips=None, mu=0.0, eps=1e-4, alpha=1.0):
    """diag (n, 3, 3) += alpha * sum_c w_i^2 H_c (FP64, contact order)."""
    d = np.ascontiguousarray(diag, np.float64).copy()
    pr = _i32(pairs).reshape(-1, 4)
    w = _f64(weights).reshape(-1, 8)
    sl = None if slips is None else _f64(slips).reshape(-1, 4)
    _check(lib().orc_contact_bake(_I(len(pr)), _ptr(pr), _ptr(w), None if sl is None else _ptr(sl), _D(k), _D(mu),
                                  _D(eps), _D(alpha), _ptr(d)))
    return d


def contact_apply(x, pairs, weights, k, slips=None, mu=0.0, eps=1e-4, alpha=1.0, y=None):
    """y (n, 3) + alpha * C_off x (FP64, contact order)."""
    xx = _f64(x).reshape(-1, 3)
    out = np.zeros_like(xx) if y is None else _f64(y).reshape(-1, 3).copy()
    pr = _i32(pairs).reshape(-1, 4)
    w = _f64(weights).reshape(-1, 8)
    sl = None if slips is None else _f64(slips).reshape(-1, 4)
    _check(lib().orc_contact_apply(_I(len(pr)), _ptr(pr), _ptr(w), None if sl is None else _ptr(sl), _D(k), _D(mu),
                                   _D(eps), _D(alpha), _ptr(xx), _ptr(out)))
    return out
