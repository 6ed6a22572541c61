"""GPU parity of the standalone BSR3 SpMV and block-Jacobi PCG entry points
(C ABI chx_spmv_bsr3 / chx_pcg_solve_bsr3) against the CPU oracle."""
import numpy as np
import pytest

import helpers
import pyoracle

pytestmark = pytest.mark.gpu


def _system(seed=0, n=24):
    """The assembled system of a small cloth after one oracle step (SPD)."""
    from chysx import scenes

    spec = scenes.plain_small(seed=seed, n=n)
    o = helpers.oracle_from_spec(spec)
    o.step(spec.dt)
    ro, ci = o.pattern()
    diag, vals, rhs, dx = o.system()
    return ro, ci, diag.astype(np.float32), vals.astype(np.float32), rhs.astype(np.float32)


def _dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a, dtype)).cuda()


@pytest.mark.parametrize("lanes", [4, 8, 16])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_spmv_matches_oracle(gpu, lanes, beta):
    import torch

    from chysx import abi

    ro, ci, diag, vals, _ = _system()
    n = len(ro) - 1
    rng = np.random.default_rng(1)
    x = rng.standard_normal((n, 3)).astype(np.float32)
    y0 = rng.standard_normal((n, 3)).astype(np.float32)
    d_ro, d_ci = _dev(torch, ro, np.int32), _dev(torch, ci, np.int32)
    d_diag, d_vals = _dev(torch, diag, np.float32), _dev(torch, vals, np.float32)
    d_x, d_y = _dev(torch, x, np.float32), _dev(torch, y0, np.float32)
    print(f"spmv n={n} nnz={len(ci)} lanes={lanes} beta={beta}", flush=True)
    abi.spmv_bsr3(n, d_diag.data_ptr(), d_ro.data_ptr(), d_ci.data_ptr(), d_vals.data_ptr(), d_x.data_ptr(),
                  d_y.data_ptr(), 1.0, beta, lanes)
    torch.cuda.synchronize()
    y = d_y.cpu().numpy().astype(np.float64)
    ref = pyoracle.spmv(ro, ci, diag, vals, x, precision=64) + beta * y0.astype(np.float64)
    scale = np.abs(diag).max() * np.abs(x).max() * 13
    assert np.max(np.abs(y - ref)) <= 1e-6 * scale


def test_spmv_empty_rows(gpu):
    """Rows without off-diagonal blocks and a ragged pattern."""
    import torch

    from chysx import abi

    n = 37
    rng = np.random.default_rng(5)
    rows, cols = [], []
    for i in range(n):
        for j in rng.choice(n, size=int(rng.integers(0, 20)), replace=False):
            if j != i:
                rows.append(i)
                cols.append(int(j))
    import chysx

    ro, ci = chysx.topology.build_pattern(n, np.array(rows, np.int32), np.array(cols, np.int32))
    diag = rng.standard_normal((n, 3, 3)).astype(np.float32)
    vals = rng.standard_normal((len(ci), 3, 3)).astype(np.float32)
    x = rng.standard_normal((n, 3)).astype(np.float32)
    d = [_dev(torch, a, t) for a, t in ((ro, np.int32), (ci, np.int32), (diag, np.float32), (vals, np.float32),
                                        (x, np.float32))]
    y = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    for lanes in (4, 8, 16):
        abi.spmv_bsr3(n, d[2].data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[3].data_ptr(), d[4].data_ptr(),
                      y.data_ptr(), 2.0, 0.0, lanes)
        torch.cuda.synchronize()
        ref = 2.0 * pyoracle.spmv(ro, ci, diag, vals, x, precision=64)
        assert np.max(np.abs(y.cpu().numpy() - ref)) <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("tol,max_iter", [(0.0, 30), (1e-6, 200)])
def test_pcg_matches_oracle(gpu, tol, max_iter):
    import torch

    from chysx import abi

    ro, ci, diag, vals, rhs = _system(seed=2, n=40)
    n = len(ro) - 1
    rng = np.random.default_rng(3)
    x0 = (1e-3 * rng.standard_normal((n, 3))).astype(np.float32)
    xo, it_o, rr_o, bb_o = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, x0, max_iter, tol, precision=32)
    ws_bytes = abi.pcg_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    d_ro, d_ci = _dev(torch, ro, np.int32), _dev(torch, ci, np.int32)
    d_diag, d_vals = _dev(torch, diag, np.float32), _dev(torch, vals, np.float32)
    d_b, d_x = _dev(torch, rhs, np.float32), _dev(torch, x0, np.float32)
    print(f"pcg n={n} tol={tol} max_iter={max_iter}", flush=True)
    it, rr, bb = abi.pcg_solve_bsr3(n, d_diag.data_ptr(), d_ro.data_ptr(), d_ci.data_ptr(), d_vals.data_ptr(),
                                    d_b.data_ptr(), d_x.data_ptr(), max_iter, tol, ws.data_ptr(), ws_bytes)
    x = d_x.cpu().numpy().astype(np.float64)
    if tol > 0:
        assert abs(it - it_o) <= 1, (it, it_o)
        assert np.sqrt(rr / bb) <= tol
    else:
        assert it == max_iter == it_o
    assert abs(bb - bb_o) <= 1e-12 * bb_o
    # x0 is ~40x larger than the solution, so r0 = b - A x0 cancels and its
    # FP32 rounding (which depends on the SpMV summation order) stays in x.
    # The FP32 oracle sits ~7e-5 (relative max-norm) from the FP64 one on this
    # system; the GPU solve must be as close to the FP64 result as that.
    x64, it64, _, _ = pyoracle.pcg_solve(ro, ci, diag, vals, rhs, x0, max_iter, tol, precision=64)
    d32 = helpers.rel_maxnorm(xo, x64)
    assert helpers.rel_maxnorm(x, x64) <= 2.0 * d32 + 1e-6, (helpers.rel_maxnorm(x, x64), d32)
    assert helpers.rel_maxnorm(x, xo) <= 3.0 * d32 + 1e-6
    # True residual ||b - A x|| / ||b|| (FP64): an FP32 solve cannot go
    # below ~kappa * eps_32 on this system (pins at k = 1e6), so it is
    # compared with the FP32 oracle's own true residual.
    def true_res(xx):
        return np.linalg.norm(pyoracle.spmv(ro, ci, diag, vals, xx, precision=64) - rhs) / np.linalg.norm(rhs)

    if tol > 0:
        assert true_res(x) <= max(2.0 * true_res(xo), 5e-6), (true_res(x), true_res(xo))


def test_pcg_deterministic(gpu):
    import torch

    from chysx import abi

    ro, ci, diag, vals, rhs = _system(seed=4, n=40)
    n = len(ro) - 1
    ws_bytes = abi.pcg_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    d = [_dev(torch, a, t) for a, t in ((ro, np.int32), (ci, np.int32), (diag, np.float32), (vals, np.float32),
                                        (rhs, np.float32))]
    outs = []
    for _ in range(2):
        x = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        abi.pcg_solve_bsr3(n, d[2].data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[3].data_ptr(), d[4].data_ptr(),
                           x.data_ptr(), 50, 0.0, ws.data_ptr(), ws_bytes)
        outs.append(x.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
