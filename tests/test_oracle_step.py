"""Pins the oracle's whole step and its PCG: free-fall and pin closed forms,
PCG against a dense solve and against scipy's CG iterates, FP32 vs FP64."""
import numpy as np
import pytest

import helpers
import pyoracle
from chysx import scenes


@pytest.mark.parametrize("prec,tol_x,tol_v", [(64, 1e-11, 1e-9), (32, 1e-4, 2e-2)])
def test_freefall_closed_form(prec, tol_x, tol_v):
    """No elastic elements: the implicit-Euler step with M/dt^2 on the
    diagonal gives dx = dt v + dt^2 g, so after n substeps
    x = x0 + n dt v0 + dt^2 g n(n+1)/2 and v = v0 + n dt g."""
    rng = np.random.default_rng(0)
    n = 50
    x0 = rng.standard_normal((n, 3))
    v0 = rng.standard_normal((n, 3))
    o = pyoracle.OracleSim(prec)
    o.set_state(x0, v0, 1.0 / rng.uniform(0.1, 2.0, n))
    g = np.array([0.3, -9.81, 1.2], np.float32).astype(np.float64)
    o.set_material(0.0, tuple(g))
    o.set_pcg(5, 0.0)
    dt = float(np.float32(1.0 / 600.0))
    steps = 40
    for _ in range(steps):
        o.step(dt)
    p, v = o.state()
    xp = x0 + steps * dt * v0 + dt * dt * g * steps * (steps + 1) / 2
    vp = v0 + steps * dt * g
    # FP32 loses digits in the reference's right-hand side
    # m/dt^2 (x~ - x_n): x~ - x_n cancels to dt v + dt^2 g.
    assert np.max(np.abs(p - xp)) <= tol_x * np.abs(xp).max()
    assert np.max(np.abs(v - vp)) <= tol_v * np.abs(vp).max()


def test_pin_closed_form():
    """One pinned particle: (m/dt^2 + k) dx = m v/dt + m g - k (x - t)."""
    m, k, dt = 0.3, 50.0, 0.01
    x = np.array([[0.2, 1.0, -0.4]])
    v = np.array([[1.0, 0.5, 0.0]])
    t = np.array([[0.0, 1.5, 0.0]])
    g = np.array([0.0, -9.81, 0.0])
    o = pyoracle.OracleSim(64)
    o.set_state(t, v * 0, np.array([1 / m]))
    o.set_pins(np.array([0], np.int32), t, k)
    o.set_state(x, v, np.array([1 / m]))
    o.set_material(0.0, tuple(g))
    o.set_pcg(3, 0.0)
    o.step(dt)
    p, _ = o.state()
    # step(dt) and the gravity vector enter as float32 values (the API types)
    dt, g = float(np.float32(dt)), g.astype(np.float32).astype(np.float64)
    dx = (m * v / dt + m * g - k * (x - t)) / (m / dt ** 2 + k)
    assert np.allclose(p, x + dx, rtol=1e-12, atol=1e-12)


def _assembled(seed=0, n=12, prec=64):
    spec = scenes.plain_small(seed=seed, n=n)
    o = helpers.oracle_from_spec(spec, precision=prec)
    for _ in range(3):
        o.step(spec.dt)
    ro, ci = o.pattern()
    diag, vals, rhs, _ = o.system()
    N = len(ro) - 1
    A = np.zeros((3 * N, 3 * N))
    for i in range(N):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = diag[i]
        for kk in range(ro[i], ro[i + 1]):
            j = ci[kk]
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = vals[kk]
    return ro, ci, diag, vals, rhs, A


def test_pcg_converges_to_dense_solve():
    """The method of scripts/diagnose_pcg.py:202-214: compare with a dense
    solve of the assembled system."""
    ro, ci, diag, vals, rhs, A = _assembled()
    x, it, rr, bb = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, np.zeros_like(rhs), 2000, 1e-13, precision=64)
    xd = np.linalg.solve(A, rhs.ravel()).reshape(-1, 3)
    assert np.sqrt(rr / bb) <= 1e-13
    assert helpers.rel_maxnorm(x, xd) <= 1e-8
    assert np.allclose(A, A.T, atol=1e-9 * np.abs(A).max())


@pytest.mark.parametrize("iters", [1, 5, 17])
def test_pcg_iterates_match_scipy_cg(iters):
    """Fixed-iteration block-Jacobi PCG from a warm start == scipy's CG with
    the same preconditioner and iteration count."""
    import scipy.sparse.linalg as spla

    ro, ci, diag, vals, rhs, A = _assembled(seed=1)
    N = len(ro) - 1
    Minv = np.zeros_like(A)
    for i in range(N):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.linalg.inv(diag[i])
    x0 = 1e-3 * np.random.default_rng(2).standard_normal((N, 3))
    x, it, _, _ = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, x0, iters, 0.0, precision=64)
    M = spla.LinearOperator(A.shape, matvec=lambda v: Minv @ v)
    xs, _ = spla.cg(A, rhs.ravel(), x0=x0.ravel(), maxiter=iters, rtol=0.0, atol=0.0, M=M)
    assert it == iters
    assert np.max(np.abs(x.ravel() - xs)) <= 1e-9 * np.abs(xs).max()


def test_pcg_tolerance_counts_and_zero_rhs():
    ro, ci, diag, vals, rhs, A = _assembled(seed=2)
    x, it, rr, bb = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, np.zeros_like(rhs), 500, 1e-6, precision=64)
    assert 0 < it < 500 and np.sqrt(rr / bb) <= 1e-6
    # the iteration before stopping had not converged yet
    x2, it2, rr2, bb2 = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, np.zeros_like(rhs), it - 1, 0.0, precision=64)
    assert np.sqrt(rr2 / bb2) > 1e-6
    # already converged warm start: zero iterations
    x3, it3, _, _ = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, x, 500, 1e-6, precision=64)
    assert it3 == 0 and np.array_equal(x3, x)
    # zero right-hand side and zero start: the guard keeps x at zero
    x4, it4, _, _ = pyoracle.pcg_solve(ro, ci, diag, vals, 0 * rhs, 0 * rhs, 7, 0.0, precision=64)
    assert it4 == 7 and np.all(x4 == 0)


def test_fp32_tracks_fp64():
    spec = scenes.plain_small(seed=5, n=16)
    o32 = helpers.oracle_from_spec(spec, precision=32)
    o64 = helpers.oracle_from_spec(spec, precision=64)
    for _ in range(30):
        o32.step(spec.dt)
        o64.step(spec.dt)
    assert helpers.rel_maxnorm(o32.state()[0], o64.state()[0]) <= 1e-5


def test_hanging_cloth_qualitative():
    """The reference's hanging-cloth assertions (example_chysx_hanging_cloth.py:
    193-227) on a smaller 21x21 grid for 1 s: finite, pins hold, the cloth
    falls, speeds bounded."""
    nx = 21
    pos, tris = scenes.cloth_grid(nx, side=1.0, height=0.0)
    pos = np.stack([pos[:, 0] - 0.5, pos[:, 2] - 0.5, np.full(len(pos), 2.0)], 1)
    o = pyoracle.OracleSim(32)
    o.set_state(pos, np.zeros_like(pos), np.ones(len(pos)))
    o.set_material(0.0, (0.0, 0.0, -9.81))
    o.set_pcg(50, 0.0)
    o.set_mesh(tris)
    o.redistribute_mass(0.1)
    o.build_stretch(1e2)
    o.build_shear(1e2)
    o.build_bending(5e-4)
    pins = np.array([0, nx - 1], np.int32)
    o.set_pins(pins, pos[pins], 1e9)
    far = (nx - 1) * nx + (nx - 1) // 2
    drop = []
    for _ in range(60):  # frames of 10 substeps; the cloth swings, so the
        for _ in range(10):  # drape is checked on the largest far-point drop
            o.step(1.0 / 600.0)
        q, qd = o.state()
        assert np.isfinite(q).all() and np.isfinite(qd).all()
        assert np.max(np.linalg.norm(q[pins] - pos[pins], axis=1)) <= 1e-3
        assert q[:, 2].mean() < 2.0
        assert np.linalg.norm(qd, axis=1).max() <= 20.0
        drop.append(2.0 - q[far, 2])
    assert max(drop) >= 0.5
