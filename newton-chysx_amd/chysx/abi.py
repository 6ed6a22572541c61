"""ctypes view of the C ABI (``include/chysx_hip.h``) in ``libchysx_hip.so``.

Used for direct kernel-level calls (standalone SpMV / PCG) and to check that
the library exports every declared entry point.  Loading the library does
not touch the GPU; the HIP runtime initialises on the first device call.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchysx_hip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "chysx_hip.h")

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_U = ctypes.c_uint64
_S = ctypes.c_size_t

_SIGS = {
    "chx_last_error_message": (ctypes.c_char_p, []),
    "chx_device_count": (_I, [_P]),
    "chx_device_info": (_I, [_P, _P, _P]),
    "chx_pcg_grid_limit": (_I, [_I, _I, _I, _P]),
    "chx_malloc": (_I, [_P, _S]),
    "chx_free": (_I, [_P]),
    "chx_get_device": (_I, [_P]),
    "chx_host_malloc": (_I, [_P, _S]),
    "chx_host_free": (_I, [_P]),
    "chx_memcpy_h2d_async": (_I, [_P, _P, _S, _U]),
    "chx_event_create": (_I, [_P]),
    "chx_event_destroy": (_I, [_P]),
    "chx_event_record": (_I, [_P, _U]),
    "chx_event_synchronize": (_I, [_P]),
    "chx_memcpy_h2d": (_I, [_P, _P, _S, _U]),
    "chx_memcpy_d2h": (_I, [_P, _P, _S, _U]),
    "chx_memcpy_d2d": (_I, [_P, _P, _S, _U]),
    "chx_zero": (_I, [_P, _S, _U]),
    "chx_stream_synchronize": (_I, [_U]),
    "chx_profile_range_push": (_I, [ctypes.c_char_p]),
    "chx_profile_range_pop": (_I, []),
    "chx_device_synchronize": (_I, []),
    "chx_triangle_rest_shape": (_I, [_P, _P, _I, _F, _F, _P, _P, _U]),
    "chx_dihedral_rest_angle": (_I, [_P, _P, _I, _P, _U]),
    "chx_mass_area_weighted": (_I, [_P, _P, _P, _P, _I, _F, _P, _U]),
    "chx_batch_energy": (_I, [_P, _I, _P, _U]),
    "chx_spmv_bsr3": (_I, [_I, _P, _P, _P, _P, _P, _P, _F, _F, _I, _U]),
    "chx_pcg_workspace_size": (_I, [_I, _P]),
    "chx_pcg_solve_bsr3": (_I, [_I, _P, _P, _P, _P, _P, _P, _I, _D, _I, _P, _S, _P, _P, _P, _U]),
    "chx_pcg_solve_bsr3_scratch": (_I, [_I, _P, _P, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _U]),
    "chx_contact_workspace_size": (_I, [_I, _I, _P]),
    "chx_contact_bake_diag": (_I, [_I, _P, _P, _F, _P, _S, _U]),
    "chx_contact_spmv": (_I, [_I, _P, _P, _P, _F, _P, _S, _U]),
    "chx_pcg_solve_bsr3_contact": (_I, [_I, _P, _P, _P, _P, _P, _P, _I, _D, _I, _P, _S, _P, _P, _S, _P, _P, _P,
                                        _U]),
    "chx_cloth_step": (_I, [_P, _U]),
    "chx_cloth_assemble": (_I, [_P, _U]),
    "chx_cloth_solve": (_I, [_P, _I, _U]),
    "chx_batch_check": (_I, [_P, _U]),
    "chx_pcg_last_stats": (_I, [_I, _P, _S, _P, _P, _P, _P, _U]),
    "chx_cloth_step_graph": (_I, [_P, _U, _P]),
    "chx_graph_launch": (_I, [_P, _U]),
    "chx_graph_destroy": (_I, [_P]),
}

_lib = None


class ChxError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        from ._runtime import share_torch_hip_runtime

        share_torch_hip_runtime()
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"chysx: {LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header: str = HEADER):
    """Names of every function declared in the public header."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chx_[a-z0-9_]+)\s*\(", text)))


def check(rc: int, what: str = "chysx"):
    if rc != 0:
        msg = lib().chx_last_error_message().decode()
        if rc == 1:
            raise ValueError(f"{what}: {msg}")
        if rc == 2:
            raise IndexError(f"{what}: {msg}")
        raise ChxError(f"{what}: {msg}")


def spmv_bsr3(n, diag, row_offsets, col_indices, values, x, y, alpha=1.0, beta=0.0, lanes_per_row=8, stream=0):
    """y = alpha*A*x + beta*y; all array arguments are device pointers (int)."""
    check(lib().chx_spmv_bsr3(n, diag, row_offsets, col_indices, values, x, y, alpha, beta, lanes_per_row, stream),
          "spmv_bsr3")


def pcg_workspace_size(n) -> int:
    out = ctypes.c_size_t()
    check(lib().chx_pcg_workspace_size(n, ctypes.byref(out)), "pcg_workspace_size")
    return out.value


def pcg_solve_bsr3(n, diag, row_offsets, col_indices, values, b, x, max_iterations, tol, workspace, workspace_bytes,
                   lanes_per_row=8, stream=0, sync=True):
    """Block-Jacobi PCG (device pointers); returns (iterations, rr, bb) when sync."""
    it, rr, bb = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    args = (ctypes.byref(it), ctypes.byref(rr), ctypes.byref(bb)) if sync else (None, None, None)
    check(lib().chx_pcg_solve_bsr3(n, diag, row_offsets, col_indices, values, b, x, max_iterations, tol,
                                   lanes_per_row, workspace, workspace_bytes, *args, stream), "pcg_solve_bsr3")
    return (it.value, rr.value, bb.value) if sync else None


class ContactOp(ctypes.Structure):
    """chx_contact_op: device pointers of the contact sidecar (pairs [m][4]
    int32, weights [m][8] float32, count one device int32, slips [m][4]
    float32 or 0) and its scalars."""
    _fields_ = [("pairs", _P), ("weights", _P), ("count", _P), ("slips", _P), ("max_contacts", _I),
                ("stiffness", _F), ("friction_mu", _F), ("friction_eps", _F)]


def contact_workspace_size(n, max_contacts) -> int:
    out = ctypes.c_size_t()
    check(lib().chx_contact_workspace_size(n, max_contacts, ctypes.byref(out)), "contact_workspace_size")
    return out.value


def contact_bake_diag(n, diag, op: ContactOp, alpha, workspace, workspace_bytes, stream=0):
    """diag ([n][9] row-major blocks) += alpha * same-slot contact terms."""
    check(lib().chx_contact_bake_diag(n, diag, ctypes.byref(op), alpha, workspace, workspace_bytes, stream),
          "contact_bake_diag")


def contact_spmv(n, op: ContactOp, x, y, alpha, workspace, workspace_bytes, stream=0):
    """y += alpha * cross-slot contact terms * x."""
    check(lib().chx_contact_spmv(n, ctypes.byref(op), x, y, alpha, workspace, workspace_bytes, stream),
          "contact_spmv")


def pcg_solve_bsr3_contact(n, diag, row_offsets, col_indices, values, b, x, max_iterations, tol, workspace,
                           workspace_bytes, op: ContactOp, contact_workspace, contact_workspace_bytes,
                           lanes_per_row=8, stream=0):
    """Block-Jacobi PCG on (A + C_off); returns (iterations, rr, bb)."""
    it, rr, bb = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    check(lib().chx_pcg_solve_bsr3_contact(n, diag, row_offsets, col_indices, values, b, x, max_iterations, tol,
                                           lanes_per_row, workspace, workspace_bytes, ctypes.byref(op),
                                           contact_workspace, contact_workspace_bytes, ctypes.byref(it),
                                           ctypes.byref(rr), ctypes.byref(bb), stream), "pcg_solve_bsr3_contact")
    return it.value, rr.value, bb.value


def device_info():
    """(compute units, persistent PCG grid cap, runtime occupancy per CU)."""
    cu, mb, occ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().chx_device_info(ctypes.byref(cu), ctypes.byref(mb), ctypes.byref(occ)), "device_info")
    return cu.value, mb.value, occ.value


def pcg_grid_limit(lanes_per_row: int, bf16: bool = False, lds_staging: bool = True) -> int:
    """Co-resident workgroups of that form of the persistent PCG kernel."""
    mb = ctypes.c_int()
    check(lib().chx_pcg_grid_limit(int(lanes_per_row), int(bool(bf16)), int(not lds_staging), ctypes.byref(mb)),
          "pcg_grid_limit")
    return mb.value
