"""Pins the oracle's damping branch of the finalize
(ChysX/src/cloth/cloth_simulator.cu:222-244) and its update_pin_targets
(cloth_simulator.cu:294-313) by closed forms: parts of the step that the
GPU parity tests of those branches compare against."""
import numpy as np
import pytest

import pyoracle


@pytest.mark.parametrize("prec,rtol", [(64, 1e-12), (32, 2e-4)])
def test_damping_closed_form(prec, rtol):
    """No elastic elements and no pins: dx = dt v + dt^2 g exactly (A =
    M/dt^2), so the finalize gives x1 = x0 + dx and
    v1 = exp(-c dt) dx / dt; a second substep starts from that damped v."""
    rng = np.random.default_rng(3)
    n = 40
    x0 = rng.standard_normal((n, 3))
    v0 = rng.standard_normal((n, 3))
    c = 2.5
    g = np.array([0.0, -9.81, 0.0])
    o = pyoracle.OracleSim(prec)
    o.set_state(x0, v0, 1.0 / rng.uniform(0.1, 2.0, n))
    o.set_material(c, tuple(g))
    o.set_pcg(5, 0.0)
    dt = float(np.float32(1.0 / 600.0))
    cf = float(np.float32(c))
    gf = g.astype(np.float32).astype(np.float64)
    x, v = x0, v0
    for _ in range(2):
        o.step(dt)
        dx = dt * v + dt * dt * gf
        x = x + dx
        v = np.exp(-cf * dt) * dx / dt
        p, q = o.state()
        assert np.max(np.abs(p - x)) <= rtol * np.abs(x).max()
        assert np.max(np.abs(q - v)) <= rtol * np.abs(v).max() * (1 if prec == 64 else 50)
    # the decay is what separates the damped from the undamped velocity
    assert np.max(np.abs(v - dx / dt)) > 1e-3 * np.abs(v).max()


def test_update_pin_targets_closed_form():
    """One pinned particle whose target moves between substeps:
    (m/dt^2 + k) dx = m v/dt + m g - k (x - t_new), for the new target and
    the state the previous substep left."""
    m, k, dt = 0.3, 50.0, float(np.float32(0.01))
    g = np.array([0.0, -9.81, 0.0])
    gf = g.astype(np.float32).astype(np.float64)
    x = np.array([[0.2, 1.0, -0.4]])
    v = np.array([[1.0, 0.5, 0.0]])
    targets = [np.array([[0.0, 1.5, 0.0]]), np.array([[0.5, 1.2, 0.1]]), np.array([[-0.3, 0.9, 0.4]])]
    o = pyoracle.OracleSim(64)
    o.set_state(targets[0], 0 * v, np.array([1 / m]))
    o.set_pins(np.array([0], np.int32), targets[0], k)
    o.set_state(x, v, np.array([1 / m]))
    o.set_material(0.0, tuple(g))
    o.set_pcg(3, 0.0)
    for t in targets:
        o.update_pin_targets(t)
        o.step(dt)
        dx = (m * v / dt + m * gf - k * (x - t)) / (m / dt ** 2 + k)
        x, v = x + dx, dx / dt
        p, q = o.state()
        assert np.allclose(p, x, rtol=1e-12, atol=1e-12)
        assert np.allclose(q, v, rtol=1e-10, atol=1e-10)
    with pytest.raises(RuntimeError):
        o.update_pin_targets(np.zeros((2, 3)))


def test_pcg_final_rho_is_preconditioned_residual():
    """The oracle's final <r, z> equals r^T M^-1 r of its final recursive
    residual; with an exact FP64 solve of a small SPD system it matches the
    true residual's value (b - A x) to rounding."""
    rng = np.random.default_rng(5)
    n = 30
    # a chain of 3x3 blocks: SPD, diagonally dominant
    ro = [0]
    cols = []
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        cols += nb
        ro.append(len(cols))
    ro = np.array(ro, np.int32)
    cols = np.array(cols, np.int32)
    off = -0.3 * np.eye(3)
    values = np.stack([off] * len(cols))
    diag = np.stack([np.eye(3) * 2.0 + 0.1 * np.diag(rng.uniform(0, 1, 3)) for _ in range(n)])
    b = rng.standard_normal((n, 3))
    x, it, rr, bb, rho = pyoracle.pcg_solve(ro, cols, diag, values, b, np.zeros((n, 3)), 4, 0.0, precision=64,
                                            return_rho=True)
    A = np.zeros((3 * n, 3 * n))
    for i in range(n):
        A[3 * i:3 * i + 3, 3 * i:3 * i + 3] = diag[i]
        for k in range(ro[i], ro[i + 1]):
            j = cols[k]
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = values[k]
    r = b.ravel() - A @ x.ravel()
    Minv = np.zeros_like(A)
    for i in range(n):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.linalg.inv(diag[i])
    assert it == 4
    assert abs(rho - r @ Minv @ r) <= 1e-10 * abs(rho)
    assert abs(rr - r @ r) <= 1e-10 * rr
