"""Pinning of the oracle's StVK membrane (no reference counterpart: the
reference's ClothMaterial keeps lame_mu / lame_lambda as unused
placeholders, ChysX/src/cloth/cloth_material.h:93-116).  Checked against
closed forms, finite differences, invariants and numpy's 2x2 eigensolver."""
import numpy as np
import pytest

import pyoracle

H = 0.1
DINV = np.eye(2) / H                # rest triangle (0,0,0), (h,0,0), (0,0,h)
REST = np.array([[0.0, 0.0, 0.0], [H, 0.0, 0.0], [0.0, 0.0, H]])
MU_A, LAM_A = 3.0, 2.0


def _stress(x, dinv=DINV, muA=MU_A, lamA=LAM_A):
    a, b = x[1] - x[0], x[2] - x[0]
    fu = a * dinv[0, 0] + b * dinv[1, 0]
    fv = a * dinv[0, 1] + b * dinv[1, 1]
    G = 0.5 * (np.array([[fu @ fu, fu @ fv], [fu @ fv, fv @ fv]]) - np.eye(2))
    return fu, fv, 2 * muA * G + lamA * np.trace(G) * np.eye(2), G


def _full(Hb):
    return Hb.transpose(0, 2, 1, 3).reshape(9, 9)  # [a, b, i, j] -> (3a+i, 3b+j)


def test_rest_state_has_no_energy_or_force():
    E, g, _ = pyoracle.stvk_element(REST, DINV, MU_A, LAM_A)
    assert abs(E) < 1e-15 and np.abs(g).max() < 1e-12


@pytest.mark.parametrize("s", [0.7, 1.0, 1.3])
def test_uniaxial_stretch_closed_form(s):
    x = REST.copy()
    x[1, 0] = s * H
    E, g, _ = pyoracle.stvk_element(x, DINV, MU_A, LAM_A)
    e = 0.5 * (s * s - 1.0)
    assert E == pytest.approx(MU_A * e * e + 0.5 * LAM_A * e * e, rel=1e-12, abs=1e-15)
    # S = diag(2 muA e + lamA e, lamA e): dE/dx1 = (1/h) S_uu s e_x, and the
    # Poisson term pulls vertex 2 across: dE/dx2 = (1/h) S_vv e_z.
    assert g[1] == pytest.approx([(2 * MU_A * e + LAM_A * e) * s / H, 0.0, 0.0], abs=1e-10)
    assert g[2] == pytest.approx([0.0, 0.0, LAM_A * e / H], abs=1e-10)


def _random_state(rng, stretch):
    x = REST + rng.normal(0.0, 0.02, REST.shape)
    x[1] = x[0] + (x[1] - x[0]) * stretch
    x[2] = x[0] + (x[2] - x[0]) * stretch
    return x


@pytest.mark.parametrize("seed,stretch", [(0, 1.2), (1, 0.8), (2, 1.0), (3, 0.5)])
def test_gradient_is_energy_derivative(seed, stretch):
    rng = np.random.default_rng(seed)
    x = _random_state(rng, stretch)
    _, g, _ = pyoracle.stvk_element(x, DINV, MU_A, LAM_A)
    eps = 1e-6
    fd = np.zeros((3, 3))
    for a in range(3):
        for i in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, i] += eps
            xm[a, i] -= eps
            fd[a, i] = (pyoracle.stvk_element(xp, DINV, MU_A, LAM_A)[0]
                        - pyoracle.stvk_element(xm, DINV, MU_A, LAM_A)[0]) / (2 * eps)
    assert np.abs(fd - g).max() <= 1e-6 * max(1.0, np.abs(g).max())
    assert np.abs(g.sum(axis=0)).max() < 1e-9  # translation invariance


@pytest.mark.parametrize("seed,stretch", [(4, 1.2), (5, 0.6), (6, 1.0)])
def test_exact_hessian_is_gradient_derivative(seed, stretch):
    rng = np.random.default_rng(seed)
    x = _random_state(rng, stretch)
    _, _, Hb = pyoracle.stvk_element(x, DINV, MU_A, LAM_A, exact=True)
    Hm = _full(Hb)
    eps = 1e-6
    fd = np.zeros((9, 9))
    for a in range(3):
        for i in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, i] += eps
            xm[a, i] -= eps
            gp = pyoracle.stvk_element(xp, DINV, MU_A, LAM_A)[1].reshape(9)
            gm = pyoracle.stvk_element(xm, DINV, MU_A, LAM_A)[1].reshape(9)
            fd[:, 3 * a + i] = (gp - gm) / (2 * eps)
    assert np.abs(fd - Hm).max() <= 1e-5 * max(1.0, np.abs(Hm).max())
    assert np.allclose(Hm, Hm.T, atol=1e-9)


@pytest.mark.parametrize("seed,stretch", [(7, 0.6), (8, 0.9), (9, 1.0), (10, 1.3)])
def test_projection_clamps_only_the_stress_term(seed, stretch):
    """Projected - exact = (S+ - S) (x) I3 mapped by the shape derivatives,
    with S+ from numpy's eigendecomposition; the projected Hessian is PSD."""
    rng = np.random.default_rng(seed)
    x = _random_state(rng, stretch)
    _, _, Hp = pyoracle.stvk_element(x, DINV, MU_A, LAM_A)
    _, _, He = pyoracle.stvk_element(x, DINV, MU_A, LAM_A, exact=True)
    _, _, S, _ = _stress(x)
    w, V = np.linalg.eigh(S)
    Sp = (V * np.maximum(w, 0.0)) @ V.T
    dS = Sp - S
    dwu = np.array([-DINV[0, 0] - DINV[1, 0], DINV[0, 0], DINV[1, 0]])
    dwv = np.array([-DINV[0, 1] - DINV[1, 1], DINV[0, 1], DINV[1, 1]])
    expect = np.zeros((3, 3, 3, 3))
    for a in range(3):
        for b in range(3):
            c = (dwu[a] * dwu[b] * dS[0, 0] + (dwu[a] * dwv[b] + dwv[a] * dwu[b]) * dS[0, 1]
                 + dwv[a] * dwv[b] * dS[1, 1])
            expect[a, b] = c * np.eye(3)
    assert np.abs((Hp - He) - expect).max() <= 1e-9 * max(1.0, np.abs(He).max())
    M = _full(Hp)
    assert np.allclose(M, M.T, atol=1e-9)
    assert np.linalg.eigvalsh(M).min() >= -1e-9 * np.abs(M).max()


def test_rigid_motion_invariance():
    rng = np.random.default_rng(11)
    x = _random_state(rng, 1.1)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    y = x @ q.T + np.array([0.3, -1.0, 2.0])
    E1, g1, _ = pyoracle.stvk_element(x, DINV, MU_A, LAM_A)
    E2, g2, _ = pyoracle.stvk_element(y, DINV, MU_A, LAM_A)
    assert E2 == pytest.approx(E1, rel=1e-12)
    assert np.abs(g2 - g1 @ q.T).max() < 1e-9


def test_system_with_stvk_membrane():
    """A 6x6 grid with the StVK membrane only: the pattern is the stretch
    pattern, system gradient = FD of system energy, and the slots resolve."""
    from chysx import scenes

    pos, tris = scenes.cloth_grid(6)
    rng = np.random.default_rng(12)
    o = pyoracle.OracleSim(64)
    n = len(pos)
    o.set_state(pos.astype(np.float64), np.zeros((n, 3)), np.ones(n))
    o.set_mesh(tris)
    o.build_stvk(38461.5, 32967.0, 1e-3)
    ref = pyoracle.OracleSim(64)
    ref.set_state(pos.astype(np.float64), np.zeros((n, 3)), np.ones(n))
    ref.set_mesh(tris)
    ref.build_stretch(100.0)
    assert all(np.array_equal(a, b) for a, b in zip(o.pattern(), ref.pattern()))
    assert np.array_equal(o.slots("stvk", 9 * len(tris)), ref.slots("stretch", 9 * len(tris)))
    x = pos.astype(np.float64) + rng.normal(0.0, 0.01, pos.shape)

    def at(y):
        o.set_state(y, np.zeros((n, 3)), np.ones(n))
        return o

    g = at(x).gradient()
    eps = 1e-7
    for (i, c) in [(7, 0), (14, 2), (20, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[i, c] += eps
        xm[i, c] -= eps
        fd = (at(xp).energy() - at(xm).energy()) / (2 * eps)
        assert fd == pytest.approx(g[i, c], rel=1e-5, abs=1e-7)
