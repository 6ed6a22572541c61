"""Pins the oracle's element math (rest shape, lumped mass, BW98
stretch/shear, dihedral bending, pins) with closed forms, finite
differences of its FP64 energy, symmetry / PSD, and invariances."""
import numpy as np
import pytest

import pyoracle
from chysx import scenes


def _oracle(pos, tris, stretch=0.0, shear=0.0, bend=0.0, pins=None, kpin=0.0, prec=64, rest=None):
    o = pyoracle.OracleSim(prec)
    rest = pos if rest is None else rest
    o.set_state(rest, np.zeros_like(rest), np.ones(len(rest)))
    o.set_mesh(tris)
    if stretch:
        o.build_stretch(stretch)
    if shear:
        o.build_shear(shear)
    if bend:
        o.build_bending(bend)
    if pins is not None:
        o.set_pins(pins, rest[pins], kpin)
    o.set_state(pos, np.zeros_like(pos), np.ones(len(pos)))
    return o


def test_rest_shape_closed_form():
    """Unjittered Newton grid of spacing h: tri (v0,v1,v3) has Dm^-1 = I/h,
    tri (v1,v2,v3) has Dm^-1 = [[1/h, -1/h], [0, 1/h]], area h^2/2; shear
    multiplies by R(pi/4)^-1 = [[c, s], [-s, c]]."""
    n = 6
    pos, tris = scenes.cloth_grid(n, side=1.0)
    h = 1.0 / (n - 1)
    o = _oracle(pos.astype(np.float64), tris, stretch=1.0, shear=1.0)
    dm, area = o.rest_shape("stretch", len(tris))
    A = np.eye(2) / h
    B = np.array([[1 / h, -1 / h], [0, 1 / h]])
    assert np.allclose(dm[0::2], A, rtol=1e-6, atol=1e-6)
    assert np.allclose(dm[1::2], B, rtol=1e-6, atol=1e-6)
    assert np.allclose(area, h * h / 2, rtol=1e-6)
    c = s = np.cos(np.float32(0.7853981633974483))
    R = np.array([[c, s], [-s, c]])
    dm_s, _ = o.rest_shape("shear", len(tris))
    assert np.allclose(dm_s[0::2], A @ R, rtol=1e-6, atol=1e-6)
    assert np.allclose(dm_s[1::2], B @ R, rtol=1e-6, atol=1e-6)


def test_degenerate_rest_shape_identity():
    pos = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float64)  # collinear
    o = _oracle(pos, np.array([[0, 1, 2]], np.int32), stretch=1.0)
    dm, area = o.rest_shape("stretch", 1)
    assert np.allclose(dm[0], np.eye(2)) and area[0] == 0.0


def test_lumped_mass_closed_form():
    """m_v = rho * h^2/6 * (number of incident triangles)."""
    n, rho = 9, 0.3
    pos, tris = scenes.cloth_grid(n, side=2.0)
    h = 2.0 / (n - 1)
    o = _oracle(pos.astype(np.float64), tris)
    o.redistribute_mass(rho)
    w = o.inv_mass()
    count = np.bincount(tris.ravel(), minlength=n * n)
    assert np.allclose(1.0 / w, rho * h * h / 6 * count, rtol=1e-12)
    # Newton winding: corner (0,0) and (n-1,n-1) touch 1 triangle, the other two 2
    assert count[0] == 1 and count[n * n - 1] == 1 and count[n - 1] == 2 and count[(n - 1) * n] == 2


def _perturbed(seed, n=4, amp=0.08):
    rng = np.random.default_rng(seed)
    rest, tris = scenes.cloth_grid(n, side=1.0)
    rest = rest.astype(np.float64)
    pos = rest + amp * rng.standard_normal(rest.shape)
    return rest, pos, tris


def _fd_gradient(o, pos, eps=1e-6):
    g = np.zeros_like(pos)
    for i in range(pos.shape[0]):
        for d in range(3):
            p = pos.copy()
            p[i, d] += eps
            o.set_state(p, np.zeros_like(p), np.ones(len(p)))
            ep = o.energy()
            p[i, d] -= 2 * eps
            o.set_state(p, np.zeros_like(p), np.ones(len(p)))
            em = o.energy()
            g[i, d] = (ep - em) / (2 * eps)
    o.set_state(pos, np.zeros_like(pos), np.ones(len(pos)))
    return g


@pytest.mark.parametrize("kind", ["stretch", "shear", "bending", "pins"])
@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_fd(kind, seed):
    rest, pos, tris = _perturbed(seed)
    kw = {"stretch": dict(stretch=3.0), "shear": dict(shear=2.0), "bending": dict(bend=0.7),
          "pins": dict(pins=np.array([0, 5], np.int32), kpin=11.0)}[kind]
    o = _oracle(pos, tris, rest=rest, **kw)
    g = o.gradient()
    fd = _fd_gradient(o, pos)
    assert np.max(np.abs(g - fd)) <= 1e-6 * max(1.0, np.abs(fd).max())
    assert np.abs(g).max() > 1e-3  # non-trivial configuration
    # translation invariance of the elastic terms: the forces sum to zero
    if kind != "pins":
        assert np.allclose(g.sum(0), 0.0, atol=1e-10)


def _dense(o, n):
    diag, vals = o.hessian()
    ro, ci = o.pattern()
    H = np.zeros((3 * n, 3 * n))
    for i in range(n):
        H[3 * i:3 * i + 3, 3 * i:3 * i + 3] += diag[i]
        for k in range(ro[i], ro[i + 1]):
            j = ci[k]
            H[3 * i:3 * i + 3, 3 * j:3 * j + 3] += vals[k]
    return H


@pytest.mark.parametrize("kind", ["stretch", "shear"])
def test_triangle_hessian_exact_when_stretched(kind):
    """With every triangle stretched (c_u, c_v > 0) the PSD filter keeps
    the full Hessian, which must equal the FD Jacobian of the gradient."""
    rest, tris = scenes.cloth_grid(3, side=1.0)
    rest = rest.astype(np.float64)
    rng = np.random.default_rng(4)
    pos = rest * 1.3 + 0.01 * rng.standard_normal(rest.shape)
    o = _oracle(pos, tris, rest=rest, **({"stretch": 2.0} if kind == "stretch" else {"shear": 2.0}))
    H = _dense(o, len(pos))
    eps = 1e-6
    J = np.zeros_like(H)
    for c in range(3 * len(pos)):
        p = pos.copy().ravel()
        p[c] += eps
        o.set_state(p.reshape(-1, 3), np.zeros_like(pos), np.ones(len(pos)))
        gp = o.gradient().ravel()
        p[c] -= 2 * eps
        o.set_state(p.reshape(-1, 3), np.zeros_like(pos), np.ones(len(pos)))
        gm = o.gradient().ravel()
        J[:, c] = (gp - gm) / (2 * eps)
    assert np.max(np.abs(H - J)) <= 1e-5 * np.abs(J).max()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hessian_symmetric_psd(seed):
    rest, pos, tris = _perturbed(seed, amp=0.15)
    o = _oracle(pos, tris, rest=rest, stretch=5.0, shear=3.0, bend=0.5, pins=np.array([0], np.int32), kpin=7.0)
    H = _dense(o, len(pos))
    assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()


def test_bending_hessian_is_gauss_newton():
    """One dihedral: E = k/2 (theta - theta0)^2 and g_i = k (theta - theta0) a_i,
    so the Gauss-Newton blocks k a_i a_j^T equal g_i g_j^T / (2 E)."""
    rest = np.array([[0, 0, 0], [1, 0, 0], [0.4, 1, 0], [0.6, -1, 0]], np.float64)
    tris = np.array([[0, 1, 2], [1, 0, 3]], np.int32)
    pos = rest.copy()
    pos[2, 2] += 0.3
    pos[3, 2] += 0.1
    k = 0.9
    o = _oracle(pos, tris, rest=rest, bend=k)
    assert o.sizes()["dihedrals"] == 1
    E = o.energy()
    g = o.gradient().ravel()
    H = _dense(o, 4)
    assert E > 1e-4
    assert np.allclose(H, np.outer(g, g) / (2 * E), atol=1e-10 * np.abs(H).max())


def test_bending_rest_angle_flat_zero():
    pos, tris = scenes.cloth_grid(5)
    o = _oracle(pos.astype(np.float64), tris, bend=1.0)
    _, rest = o.dihedrals()
    assert np.allclose(rest, 0.0, atol=1e-12)


def test_energy_rotation_invariant():
    rest, pos, tris = _perturbed(3)
    o = _oracle(pos, tris, rest=rest, stretch=2.0, shear=1.0, bend=0.3)
    e0 = o.energy()
    a = 0.7
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    p2 = pos @ R.T + np.array([0.3, -2.0, 5.0])
    o.set_state(p2, np.zeros_like(p2), np.ones(len(p2)))
    assert abs(o.energy() - e0) <= 1e-10 * abs(e0)
