"""chx_pcg_solve_bsr3_scratch (no counterpart in the reference): the
standalone PCG without a workspace argument, for callers that cannot
allocate.  Through the C ABI, its results equal chx_pcg_solve_bsr3's with a
caller workspace bit for bit, on two calls in a row (the second on a smaller
system, so stale contents of the shared scratch must not leak in); systems
above CHX_PCG_SCRATCH_MAX_ROWS are rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(n_side, seed):
    """The assembled system of one substep of an n_side x n_side cloth."""
    from chysx import scenes

    dc = scenes.DeviceCloth(scenes.plain_small(seed=seed, n=n_side))
    dc.step()
    diag, ro, ci, vals, rhs, _ = dc.sim.debug_dump_last_solve()
    return diag, ro, ci, vals, rhs


def _solve(entry, sysm, iters, tol, x0):
    import torch

    from chysx import abi

    diag, ro, ci, vals, rhs = sysm
    n = len(ro) - 1
    d = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (diag, ro, ci, vals, rhs)]
    x = torch.from_numpy(x0.copy()).cuda()
    it, rr, bb = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    outs = (ctypes.byref(it), ctypes.byref(rr), ctypes.byref(bb))
    L = abi.lib()
    if entry == "workspace":
        ws = torch.zeros(abi.pcg_workspace_size(n), dtype=torch.uint8, device="cuda")
        rc = L.chx_pcg_solve_bsr3(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                                  d[4].data_ptr(), x.data_ptr(), iters, tol, 8, ws.data_ptr(), ws.numel(), *outs, 0)
    else:
        rc = L.chx_pcg_solve_bsr3_scratch(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(),
                                          d[4].data_ptr(), x.data_ptr(), iters, tol, 8, *outs, 0)
    assert rc == 0, L.chx_last_error_message()
    torch.cuda.synchronize()
    return x.cpu().numpy(), it.value, rr.value, bb.value


def test_scratch_solve_equals_workspace_solve(gpu):
    big = _system(96, seed=41)    # 9216 rows
    small = _system(40, seed=42)  # 1600 rows
    for sysm, iters, tol in ((big, 200, 1e-6), (small, 60, 0.0)):
        n = len(sysm[1]) - 1
        x0 = np.random.default_rng(n).standard_normal((n, 3)).astype(np.float32) * 1e-4
        xw, itw, rrw, bbw = _solve("workspace", sysm, iters, tol, x0)
        xs, its, rrs, bbs = _solve("scratch", sysm, iters, tol, x0)
        print(f"n={n}: iterations {itw} / {its}, rr {rrw:.6e} / {rrs:.6e}", flush=True)
        assert np.array_equal(xw, xs) and itw == its and rrw == rrs and bbw == bbs, n


def test_scratch_solve_rejects_oversized_systems(gpu):
    from chysx import abi

    L = abi.lib()
    rc = L.chx_pcg_solve_bsr3_scratch(65537, 0, 0, 0, 0, 0, 0, 10, 0.0, 8, None, None, None, 0)
    assert rc == 1 and b"CHX_PCG_SCRATCH_MAX_ROWS" in L.chx_last_error_message()
