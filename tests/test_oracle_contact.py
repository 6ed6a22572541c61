"""Pinning of the oracle's static plane / box contact (reference:
ChysX/src/collision/static_contact.{h,cu}, friction.cuh).  Checked against
independent formulations (closest point by clamping for the box distance),
the closed form of a single particle's implicit step, continuity of the
friction ramp and the stick / slip invariants of the Coulomb projection."""
import numpy as np
import pytest

import pyoracle


def f32(v):
    """Scalars reach the oracle through float (the reference's parameter
    types); the closed forms use the same rounded values."""
    return float(np.float32(v))


def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@pytest.mark.parametrize("seed", range(6))
def test_box_sdf_matches_closest_point(seed):
    """Outside: distance to the clamped closest point, normal along it;
    inside: the smallest face gap (negative), normal to that face."""
    rng = np.random.default_rng(seed)
    R = _rot(rng)
    c = rng.normal(size=3)
    h = rng.uniform(0.2, 1.5, size=3)
    for _ in range(200):
        p = c + rng.normal(size=3) * 1.5
        d, n = pyoracle.box_sdf(p, c, h, R[:, 0], R[:, 1], R[:, 2])
        q = R.T @ (p - c)
        cl = np.clip(q, -h, h)
        if np.any(np.abs(q) > h):
            gap = q - cl
            assert d == pytest.approx(np.linalg.norm(gap), rel=1e-12, abs=1e-12)
            assert np.allclose(n, R @ gap / np.linalg.norm(gap), atol=1e-12)
        else:
            face = np.argmin(h - np.abs(q))
            assert d == pytest.approx(-(h - np.abs(q)).min(), abs=1e-12)
            assert np.allclose(n, R[:, face] * (1.0 if q[face] >= 0 else -1.0), atol=1e-12)
        assert np.linalg.norm(n) == pytest.approx(1.0, abs=1e-12)


def test_friction_ramp_continuity():
    """f(u) = f1(u)/u: 2/eps at rest, 1/u past eps; f1 = u f(u) reaches 1
    at eps with zero slope (C1), so the force saturates at mu f_n."""
    eps = 1e-4
    f = pyoracle.ipc_f1_over_x
    assert f(0.0, eps) == pytest.approx(2 / eps)
    assert f(eps, eps) == pytest.approx(1 / eps)
    assert f(eps * (1 + 1e-9), eps) == pytest.approx(1 / eps, rel=1e-8)
    f1 = lambda u: u * f(u, eps)  # noqa: E731
    # left slope at eps from two step sizes, extrapolated to a zero step
    s1 = (f1(eps) - f1(eps - 1e-8)) / 1e-8
    s2 = (f1(eps) - f1(eps - 2e-8)) / 2e-8
    assert abs(2 * s1 - s2) < 1e-3
    assert (f1(eps + 1e-8) - f1(eps)) / 1e-8 == pytest.approx(0.0, abs=1e-6)
    assert f1(3 * eps) == pytest.approx(1.0)


def _particle(x, v, dt, gravity=(0.0, 0.0, -9.81), precision=64):
    o = pyoracle.OracleSim(precision)
    o.set_state(np.asarray([x], np.float64), np.asarray([v], np.float64), np.ones(1))
    o.set_material(0.0, tuple(gravity))
    o.set_pcg(10, 0.0)
    return o


@pytest.mark.parametrize("z0", [0.004, 0.001, -0.002, 0.02])
def test_plane_contact_single_particle_closed_form(z0):
    """One unit-mass particle above a ground plane z = 0 (thickness h,
    stiffness k): A = I/dt^2 + k e_z e_z^T when depth = h - z > 0, b =
    (x~ - x)/dt^2 + k depth e_z; the normal and tangential parts solve in
    closed form."""
    dt, h, k = f32(1.0 / 600.0), f32(0.005), 1.0e4
    x = np.array([0.3, -0.2, z0])
    v = np.array([0.5, 0.1, -0.7])
    o = _particle(x, v, dt)
    o.add_static_plane([0.0, 0.0, 1.0], 0.0)
    o.set_static_contact(h, k)
    o.step(dt)
    p, vel = o.state()
    g = np.array([0.0, 0.0, f32(-9.81)])
    xt = x + dt * v + dt * dt * g
    depth = max(0.0, h - z0)
    m = 1.0 / dt ** 2
    dx_t = (xt - x)[:2]
    dx_n = (m * (xt - x)[2] + k * depth) / (m + (k if depth > 0 else 0.0))
    assert np.allclose(p[0, :2], x[:2] + dx_t, rtol=0, atol=1e-12)
    assert p[0, 2] == pytest.approx(z0 + dx_n, abs=1e-12)
    assert np.allclose(vel[0], (p[0] - x) / dt, atol=1e-9)


def test_box_contact_pushes_toward_closest_face():
    """A particle just inside the +x face of an axis-aligned unit box is
    pushed along +x with depth h - d (d < 0) and nothing else changes."""
    dt, h, k = f32(1.0 / 600.0), f32(0.01), 5.0e4
    x = np.array([0.495, 0.1, 0.2])          # 5 mm inside the face x = 0.5
    o = _particle(x, np.zeros(3), dt, gravity=(0.0, 0.0, 0.0))
    o.add_static_box([0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [1, 0, 0], [0, 1, 0], [0, 0, 1])
    o.set_static_contact(h, k)
    o.step(dt)
    p, _ = o.state()
    depth = h + (0.5 - x[0])  # signed distance inside the box: x - 0.5 < 0
    m = 1.0 / dt ** 2
    assert p[0, 0] == pytest.approx(x[0] + k * depth / (m + k), abs=1e-12)
    assert np.allclose(p[0, 1:], x[1:], atol=1e-15)


def test_deepest_primitive_wins():
    """Two planes: the contact uses the one with the larger depth."""
    dt, h, k = f32(1.0 / 600.0), f32(0.01), 1.0e4
    x = np.array([0.0, 0.0, 0.002])
    o = _particle(x, np.zeros(3), dt, gravity=(0.0, 0.0, 0.0))
    o.add_static_plane([0.0, 0.0, 1.0], 0.0)       # depth 0.008
    o.add_static_plane([1.0, 0.0, 0.0], 0.005)     # x + 0.005: d = 0.005, depth 0.005
    o.set_static_contact(h, k)
    o.step(dt)
    p, _ = o.state()
    m = 1.0 / dt ** 2
    assert p[0, 2] == pytest.approx(x[2] + k * (h - x[2]) / (m + k), abs=1e-12)
    assert p[0, 0] == pytest.approx(0.0, abs=1e-15)


def test_friction_sticks_below_the_cone():
    """Resting on the ground with a small tangential velocity and mu = 1:
    the tangential part F0_t of the assembled force (inertia, gravity and
    the reference's +alpha u_t term) is inside the cone mu f_n, so the
    projection cancels it (stick); the tangential rows of A are then
    (1/dt^2 + alpha) with a zero right-hand side: dx_t = 0 exactly."""
    dt, h, k, mu = f32(1.0 / 600.0), f32(0.005), 1.0e4, 1.0
    x = np.array([0.0, 0.0, 0.003])
    v = np.array([1e-3, 0.0, 0.0])
    o = _particle(x, v, dt)
    o.add_static_plane([0.0, 0.0, 1.0], 0.0)
    o.set_static_contact(h, k, mu, 1e-4)
    o.step(dt)
    p, _ = o.state()
    assert np.allclose(p[0, :2], x[:2], atol=1e-15)


def test_friction_slip_is_capped_at_the_cone():
    """A fast tangential velocity: F0_t exceeds mu f_n, the projection
    shrinks it to exactly mu f_n (slip), so the tangential rhs has norm
    |F0_t| - mu f_n and keeps F0_t's direction."""
    dt, h, k, mu, eps = f32(1.0 / 600.0), f32(0.005), 1.0e4, f32(0.3), f32(1e-4)
    x = np.array([0.0, 0.0, 0.003])
    v = np.array([2.0, 1.0, 0.0])
    o = _particle(x, v, dt)
    o.add_static_plane([0.0, 0.0, 1.0], 0.0)
    o.set_static_contact(h, k, mu, eps)
    o.step(dt)
    p, _ = o.state()
    m = 1.0 / dt ** 2
    depth = h - x[2]
    ut = v * dt
    ut[2] = 0.0
    alpha = mu * k * depth * pyoracle.ipc_f1_over_x(np.linalg.norm(ut), eps)
    g = np.array([0.0, 0.0, f32(-9.81)])
    xt = x + dt * v + dt * dt * g
    F0 = m * (xt - x) + alpha * ut                   # rhs without the normal push
    fn = max(0.0, -F0[2])
    F0t = F0.copy()
    F0t[2] = 0.0
    assert np.linalg.norm(F0t) > mu * fn             # slip regime
    bt = F0t * (1.0 - mu * fn / np.linalg.norm(F0t))
    dxt = bt[:2] / (m + alpha)                       # A_t = (1/dt^2 + alpha) I_t
    assert np.allclose(p[0, :2] - x[:2], dxt, rtol=1e-9, atol=1e-15)
