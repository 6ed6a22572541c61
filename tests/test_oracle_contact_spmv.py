"""Pinning of the oracle's contact SpMV sidecar (reference:
ChysX/src/collision/contact_spmv.{h,cu}).  The bake (same-slot terms into the
diagonal) plus the SpMV (cross-slot terms) must equal the dense contact
Hessian C = sum_c (w w^T) kron H_c built here from its definition; H_c is a
projector sum (normal block k n n^T, tangent block a (I - n n^T)); weights
that sum to zero make C blind to rigid translations; alpha scales linearly;
the operator is symmetric positive semi-definite when 0 <= a <= k."""
import numpy as np
import pytest

import pyoracle


def _contacts(rng, n, m, signed=True):
    pairs = np.stack([rng.choice(n, 4, replace=False) for _ in range(m)]).astype(np.int32)
    w = np.zeros((m, 8))
    if signed:  # point-triangle style: 1, -b0, -b1, -b2 with b summing to 1
        b = rng.dirichlet(np.ones(3), size=m)
        w[:, 0] = 1.0
        w[:, 1:4] = -b
    else:
        w[:, 0:4] = rng.uniform(-1, 1, size=(m, 4))
    nrm = rng.normal(size=(m, 3))
    w[:, 4:7] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    w[:, 7] = rng.uniform(1e-4, 1e-2, size=m)
    slips = np.zeros((m, 4))
    slips[:, :3] = rng.normal(size=(m, 3)) * 1e-3
    slips[:, 3] = np.linalg.norm(slips[:, :3], axis=1)
    return pairs, w, slips


def _f1(u, eps):
    return 1.0 / u if u >= eps else (-u / eps + 2.0) / eps


def _dense(n, pairs, w, k, slips=None, mu=0.0, eps=1e-4):
    """C[3i:3i+3, 3j:3j+3] = sum_c w_i w_j H_c."""
    C = np.zeros((3 * n, 3 * n))
    for c in range(len(pairs)):
        nn = w[c, 4:7]
        a = mu * k * w[c, 7] * _f1(slips[c, 3], eps) if (mu > 0 and slips is not None) else 0.0
        H = k * np.outer(nn, nn) + a * (np.eye(3) - np.outer(nn, nn))
        for s in range(4):
            for t in range(4):
                i, j = pairs[c, s], pairs[c, t]
                C[3 * i:3 * i + 3, 3 * j:3 * j + 3] += w[c, s] * w[c, t] * H
    return C


def _operator(n, pairs, w, k, slips=None, mu=0.0, eps=1e-4, alpha=1.0):
    d = pyoracle.contact_bake(np.zeros((n, 3, 3)), pairs, w, k, slips, mu, eps, alpha)

    def apply(x):
        y = pyoracle.contact_apply(x, pairs, w, k, slips, mu, eps, alpha)
        return y + np.einsum("nij,nj->ni", d, x.reshape(n, 3))

    return d, apply


@pytest.mark.parametrize("mu", [0.0, 0.4])
@pytest.mark.parametrize("signed", [True, False])
def test_bake_plus_apply_is_the_dense_contact_hessian(mu, signed):
    rng = np.random.default_rng(11 + int(signed))
    n, m, k = 40, 60, 2.5e3
    pairs, w, slips = _contacts(rng, n, m, signed)
    C = _dense(n, pairs, w, k, slips, mu)
    d, apply = _operator(n, pairs, w, k, slips, mu)
    for i in range(n):  # the bake holds exactly the diagonal blocks
        assert np.allclose(d[i], C[3 * i:3 * i + 3, 3 * i:3 * i + 3], rtol=1e-12, atol=1e-9)
    for _ in range(3):
        x = rng.normal(size=(n, 3))
        assert np.allclose(apply(x).ravel(), C @ x.ravel(), rtol=1e-12, atol=1e-8)


def test_translation_invariance_and_spd():
    rng = np.random.default_rng(3)
    n, m, k = 30, 45, 1e4
    pairs, w, slips = _contacts(rng, n, m, signed=True)
    _, apply = _operator(n, pairs, w, k, slips, mu=0.3)
    t = np.tile(rng.normal(size=3), (n, 1))  # rigid translation, sum of weights 0
    assert np.abs(apply(t)).max() < 1e-9 * k
    C = np.stack([apply(e.reshape(n, 3)).ravel() for e in np.eye(3 * n)], axis=1)
    assert np.allclose(C, C.T, rtol=0, atol=1e-9 * k)
    ev = np.linalg.eigvalsh(0.5 * (C + C.T))
    assert ev.min() > -1e-9 * k


def test_single_contact_closed_form():
    """One contact between two particles, weights (1, -1), normal z:
    C = k [[n n^T, -n n^T], [-n n^T, n n^T]] — a spring along the normal."""
    n, k = 4, 7.0
    pairs = np.array([[0, 1, 2, 3]], np.int32)
    w = np.array([[1.0, -1.0, 0.0, 0.0, 0.0, 0.0, 1.0, 1e-3]])
    x = np.array([[0.1, 0.2, 0.3], [-0.5, 0.4, 0.9], [9.0, 9.0, 9.0], [9.0, 9.0, 9.0]])
    d, apply = _operator(n, pairs, w, k)
    assert np.allclose(d[0], np.diag([0, 0, k])) and np.allclose(d[1], np.diag([0, 0, k]))
    assert np.all(d[2:] == 0)
    y = apply(x)
    dz = x[0, 2] - x[1, 2]
    assert np.allclose(y, [[0, 0, k * dz], [0, 0, -k * dz], [0, 0, 0], [0, 0, 0]])


def test_friction_block_and_alpha_scaling():
    """With slip past eps, H = k n n^T + a (I - n n^T), a = mu k depth / |u|;
    alpha scales the whole operator (the friction stiffness included)."""
    n, k, mu, eps = 4, 50.0, 0.5, 1e-4
    pairs = np.array([[0, 1, 2, 3]], np.int32)
    w = np.array([[1.0, 0.0, 0.0, 0.0, 0.6, 0.0, 0.8, 2e-3]])
    slips = np.array([[3e-3, 0.0, 0.0, 3e-3]])
    d, _ = _operator(n, pairs, w, k, slips, mu, eps)
    nn = np.array([0.6, 0.0, 0.8])
    a = mu * k * 2e-3 / 3e-3
    assert np.allclose(d[0], k * np.outer(nn, nn) + a * (np.eye(3) - np.outer(nn, nn)), rtol=1e-12)
    d2, ap2 = _operator(n, pairs, w, k, slips, mu, eps, alpha=-2.5)
    assert np.allclose(d2, -2.5 * d, rtol=1e-12)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(n, 3))
    _, ap1 = _operator(n, pairs, w, k, slips, mu, eps)
    assert np.allclose(ap2(x), -2.5 * ap1(x), rtol=1e-12)


def test_contact_pcg_solves_the_sum():
    """The oracle's PCG on A with the contact bake in its diagonal and the
    SpMV in every product solves (A + C) x = b (dense check)."""
    rng = np.random.default_rng(5)
    n, m, k = 24, 20, 300.0
    pairs, w, slips = _contacts(rng, n, m, signed=True)
    diag = np.tile(np.eye(3) * 10.0, (n, 1, 1))
    ro = np.arange(n + 1, dtype=np.int32) * 0
    ci = np.zeros(0, np.int32)
    d_c = pyoracle.contact_bake(diag, pairs, w, k, slips, 0.2)
    b = rng.normal(size=(n, 3))
    x = pyoracle.contact_pcg_solve(ro, ci, d_c, np.zeros((0, 3, 3)), b, np.zeros_like(b), 400, 1e-12, pairs, w, k,
                                   slips, 0.2, precision=64)[0]
    A = _dense(n, pairs, w, k, slips, 0.2) + np.kron(np.eye(n), np.eye(3) * 10.0)
    assert np.allclose(x.ravel(), np.linalg.solve(A, b.ravel()), rtol=1e-8, atol=1e-10)
