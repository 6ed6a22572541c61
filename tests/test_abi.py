"""The C-ABI library loads and exports every entry point declared in
include/chysx_hip.h; argument validation answers without a device."""
import ctypes

import pytest

from chysx import abi


def test_exports_every_declared_symbol():
    syms = abi.declared_symbols()
    assert len(syms) >= 20
    L = abi.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(abi._SIGS), set(syms) ^ set(abi._SIGS)


def test_validation_errors_without_device():
    L = abi.lib()
    rc = L.chx_spmv_bsr3(-1, None, None, None, None, None, None, 1.0, 0.0, 8, 0)
    assert rc == 1 and b"negative" in L.chx_last_error_message()
    rc = L.chx_spmv_bsr3(0, None, None, None, None, None, None, 1.0, 0.0, 8, 0)
    assert rc == 0  # empty system is a no-op
    with pytest.raises(ValueError):
        abi.spmv_bsr3(4, 0, 0, 0, 0, 0, 0)
    rc = L.chx_triangle_rest_shape(None, None, -2, 1.0, 0.0, None, None, 0)
    assert rc == 1
    rc = L.chx_cloth_step(None, 0)
    assert rc == 1 and b"null batch" in L.chx_last_error_message()
    assert abi.pcg_workspace_size(1000) > 1000 * 12 * 5


def test_pcg_workspace_size_monotone():
    assert abi.pcg_workspace_size(0) <= abi.pcg_workspace_size(1) < abi.pcg_workspace_size(100000)


def test_profile_ranges_without_device():
    """The roctx ranges (the reference's NVTX ranges) need no GPU: nested
    push / pop succeed, a null name is an argument error."""
    L = abi.lib()
    assert L.chx_profile_range_push(b"chysx::cloth::step") == 0
    assert L.chx_profile_range_push(b"step::hessian") == 0
    assert L.chx_profile_range_pop() == 0
    assert L.chx_profile_range_pop() == 0
    assert L.chx_profile_range_push(None) == 1 and b"null name" in L.chx_last_error_message()
