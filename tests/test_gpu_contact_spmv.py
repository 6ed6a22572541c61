"""GPU parity of the contact SpMV sidecar (C ABI chx_contact_bake_diag /
chx_contact_spmv / chx_pcg_solve_bsr3_contact; reference:
ChysX/src/collision/contact_spmv.{h,cu} and the contact argument of
PCGSolver::solve, pcg_solver.cu:134-215) against the CPU oracle."""
import numpy as np
import pytest

import helpers
import pyoracle

pytestmark = pytest.mark.gpu


def _dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


def _contacts(rng, n, m, k_slip=1e-3):
    """Point-triangle style contacts: weights (1, -b0, -b1, -b2), unit
    normal, depth, slip (u_t, |u_t|)."""
    pairs = np.stack([rng.choice(n, 4, replace=False) for _ in range(m)]).astype(np.int32)
    w = np.zeros((m, 8), np.float32)
    w[:, 0] = 1.0
    w[:, 1:4] = -rng.dirichlet(np.ones(3), size=m)
    nrm = rng.normal(size=(m, 3))
    w[:, 4:7] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    w[:, 7] = rng.uniform(1e-4, 1e-2, size=m)
    slips = np.zeros((m, 4), np.float32)
    slips[:, :3] = rng.normal(size=(m, 3)) * k_slip
    slips[:, 3] = np.linalg.norm(slips[:, :3].astype(np.float64), axis=1)
    return pairs, w, slips


class _Op:
    """Device copies of a contact set and its chx_contact_op."""

    def __init__(self, torch, pairs, w, slips, count, k, mu, eps=1e-4, n=None):
        from chysx import abi

        self.pairs, self.w, self.slips = _dev(torch, pairs, np.int32), _dev(torch, w, np.float32), (
            _dev(torch, slips, np.float32) if slips is not None else None)
        self.count = _dev(torch, np.array([count], np.int32), np.int32)
        self.op = abi.ContactOp(self.pairs.data_ptr(), self.w.data_ptr(), self.count.data_ptr(),
                                self.slips.data_ptr() if self.slips is not None else None, len(pairs), k, mu, eps)
        self.ws_bytes = abi.contact_workspace_size(n, len(pairs))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device="cuda")


def _f32(v):
    return float(np.float32(v))


@pytest.mark.parametrize("mu", [0.0, 0.4])
def test_bake_and_spmv_match_oracle(gpu, mu):
    import torch

    from chysx import abi

    rng = np.random.default_rng(7)
    n, m, k, alpha = 3000, 5000, _f32(2.5e3), _f32(0.75)
    pairs, w, slips = _contacts(rng, n, m)
    count = m - 123  # the device count, below max_contacts
    o = _Op(torch, pairs, w, slips, count, k, mu, n=n)
    diag0 = rng.normal(size=(n, 3, 3)).astype(np.float32)
    x = rng.normal(size=(n, 3)).astype(np.float32)
    y0 = rng.normal(size=(n, 3)).astype(np.float32)
    d_diag, d_x, d_y = _dev(torch, diag0, np.float32), _dev(torch, x, np.float32), _dev(torch, y0, np.float32)
    print(f"contact bake/spmv n={n} m={m} count={count} mu={mu}", flush=True)
    abi.contact_bake_diag(n, d_diag.data_ptr(), o.op, alpha, o.ws.data_ptr(), o.ws_bytes)
    abi.contact_spmv(n, o.op, d_x.data_ptr(), d_y.data_ptr(), alpha, o.ws.data_ptr(), o.ws_bytes)
    torch.cuda.synchronize()
    sl = slips.astype(np.float64)
    ref_d = pyoracle.contact_bake(diag0, pairs[:count], w[:count], k, sl[:count], mu, 1e-4, alpha)
    ref_y = pyoracle.contact_apply(x, pairs[:count], w[:count], k, sl[:count], mu, 1e-4, alpha, y=y0)
    got_d = d_diag.cpu().numpy().astype(np.float64)
    got_y = d_y.cpu().numpy().astype(np.float64)
    # FP32 sums of a few contact terms per row against FP64, relative to the
    # largest contact term
    dscale = np.abs(ref_d - diag0).max() + np.abs(diag0).max()
    assert np.max(np.abs(got_d - ref_d)) <= 2e-6 * dscale
    yscale = np.abs(ref_y - y0).max() + np.abs(y0).max()
    assert np.max(np.abs(got_y - ref_y)) <= 4e-6 * yscale


def test_spmv_bitwise_repeatable_and_skips(gpu):
    """Same bits on every call; zero-weight slots and particle indices
    outside [0, n) add nothing; count 0 and count > max_contacts."""
    import torch

    from chysx import abi

    rng = np.random.default_rng(9)
    n, m, k = 500, 900, _f32(100.0)
    pairs, w, slips = _contacts(rng, n, m)
    pairs_ok = pairs.copy()
    for c in range(0, m, 7):  # edge-style contacts with two unused slots
        w[c, 2:4] = 0.0
        pairs[c, 2:4] = (-1, n + 5)
        pairs_ok[c, 2:4] = 0
    o = _Op(torch, pairs, w, slips, m, k, 0.3, n=n)
    x = rng.normal(size=(n, 3)).astype(np.float32)
    d_x = _dev(torch, x, np.float32)
    outs = []
    for _ in range(3):
        d_y = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        abi.contact_spmv(n, o.op, d_x.data_ptr(), d_y.data_ptr(), 1.0, o.ws.data_ptr(), o.ws_bytes)
        outs.append(d_y.cpu().numpy())
    assert all(np.array_equal(outs[0], v) for v in outs[1:])
    ref = pyoracle.contact_apply(x, pairs_ok, w, k, slips.astype(np.float64), 0.3, 1e-4, 1.0)
    assert np.max(np.abs(outs[0] - ref)) <= 1e-5 * np.abs(ref).max()
    for cnt, expect_zero in ((0, True), (10 * m, False)):
        o.count.fill_(cnt)
        d_y = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        abi.contact_spmv(n, o.op, d_x.data_ptr(), d_y.data_ptr(), 1.0, o.ws.data_ptr(), o.ws_bytes)
        got = d_y.cpu().numpy()
        if expect_zero:
            assert not np.any(got)
        else:  # clamped at max_contacts
            assert np.array_equal(got, outs[0])


def _system(seed=2, n=40):
    from chysx import scenes

    spec = scenes.plain_small(seed=seed, n=n)
    o = helpers.oracle_from_spec(spec)
    o.step(spec.dt)
    ro, ci = o.pattern()
    diag, vals, rhs, _ = o.system()
    return ro, ci, diag.astype(np.float32), vals.astype(np.float32), rhs.astype(np.float32)


@pytest.mark.parametrize("mu", [0.0, 0.3])
def test_contact_pcg_matches_oracle(gpu, mu):
    """Bake on the GPU, then the contact-aware PCG against the oracle's PCG
    on (A + C_off) with the same baked diagonal."""
    import torch

    from chysx import abi

    ro, ci, diag, vals, rhs = _system()
    n = len(ro) - 1
    rng = np.random.default_rng(4)
    m = n // 2
    pairs, w, slips = _contacts(rng, n, m)
    k = _f32(0.2 * np.abs(diag[:, [0, 1, 2], [0, 1, 2]]).mean())
    o = _Op(torch, pairs, w, slips, m, k, mu, n=n)
    d_ro, d_ci = _dev(torch, ro, np.int32), _dev(torch, ci, np.int32)
    d_diag, d_vals = _dev(torch, diag, np.float32), _dev(torch, vals, np.float32)
    abi.contact_bake_diag(n, d_diag.data_ptr(), o.op, 1.0, o.ws.data_ptr(), o.ws_bytes)
    baked = d_diag.cpu().numpy()
    ws_bytes = abi.pcg_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    sl = slips.astype(np.float64)
    for tol, max_iter in ((0.0, 40), (1e-6, 400)):
        d_b, d_x = _dev(torch, rhs, np.float32), torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        print(f"contact pcg n={n} m={m} mu={mu} tol={tol}", flush=True)
        it, rr, bb = abi.pcg_solve_bsr3_contact(n, d_diag.data_ptr(), d_ro.data_ptr(), d_ci.data_ptr(),
                                                d_vals.data_ptr(), d_b.data_ptr(), d_x.data_ptr(), max_iter, tol,
                                                ws.data_ptr(), ws_bytes, o.op, o.ws.data_ptr(), o.ws_bytes)
        x = d_x.cpu().numpy().astype(np.float64)
        xo, it_o, _, _ = pyoracle.contact_pcg_solve(ro, ci, baked, vals, rhs, np.zeros_like(rhs), max_iter, tol,
                                                    pairs, w, k, sl, mu, precision=32)
        x64, _, _, _ = pyoracle.contact_pcg_solve(ro, ci, baked, vals, rhs, np.zeros_like(rhs), max_iter, tol,
                                                  pairs, w, k, sl, mu, precision=64)
        if tol > 0:
            assert abs(it - it_o) <= 1, (it, it_o)
            assert np.sqrt(rr / bb) <= tol
        else:
            assert it == max_iter
        d32 = helpers.rel_maxnorm(xo, x64)
        assert helpers.rel_maxnorm(x, x64) <= 2.0 * d32 + 1e-5, (helpers.rel_maxnorm(x, x64), d32)


def test_contact_pcg_without_contacts_is_the_plain_solve(gpu):
    """count 0: the same bits as chx_pcg_solve_bsr3."""
    import torch

    from chysx import abi

    ro, ci, diag, vals, rhs = _system(seed=3, n=32)
    n = len(ro) - 1
    pairs, w, slips = _contacts(np.random.default_rng(0), n, 16)
    o = _Op(torch, pairs, w, slips, 0, 50.0, 0.2, n=n)
    d = [_dev(torch, a, t) for a, t in ((ro, np.int32), (ci, np.int32), (diag, np.float32), (vals, np.float32),
                                        (rhs, np.float32))]
    ws_bytes = abi.pcg_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    xa = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    xb = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    abi.pcg_solve_bsr3(n, d[2].data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[3].data_ptr(), d[4].data_ptr(),
                       xa.data_ptr(), 60, 1e-6, ws.data_ptr(), ws_bytes)
    abi.pcg_solve_bsr3_contact(n, d[2].data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[3].data_ptr(),
                               d[4].data_ptr(), xb.data_ptr(), 60, 1e-6, ws.data_ptr(), ws_bytes, o.op,
                               o.ws.data_ptr(), o.ws_bytes)
    assert torch.equal(xa, xb)


def test_contact_argument_errors(gpu):
    from chysx import abi

    # host-side checks only: each call is refused before any launch
    op = abi.ContactOp(None, None, None, None, 10, 1.0, 0.0, 1e-4)
    abi.contact_spmv(100, op, 1, 1, 1.0, None, 0)  # no count: nothing to do
    op.count = 1
    with pytest.raises(ValueError, match="workspace too small"):
        abi.contact_spmv(100, op, 1, 1, 1.0, None, 0)
    with pytest.raises(ValueError, match="null"):
        abi.contact_spmv(100, op, None, None, 1.0, None, 0)
    need = abi.contact_workspace_size(100, 10)
    assert need >= 4 * (101 + 100 + 40)
    with pytest.raises(ValueError):
        abi.contact_workspace_size(-1, 10)
