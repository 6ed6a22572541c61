"""One HIP runtime per process.

PyTorch for ROCm ships its own ``libamdhip64.so`` and its extension libraries
name it without a version suffix, while ``libchysx_hip.so`` links the
system ROCm's ``libamdhip64.so.7``.  If the system runtime is loaded first,
a later ``import torch`` maps a second HIP runtime into the process and one
of the two cannot open the device.  Importing torch (when installed) before
the native library makes the dynamic loader resolve ``libamdhip64.so.7`` to
the runtime torch already loaded, so chysx kernels and torch tensors share
one runtime, one device context and the same streams.  Without torch the
system runtime is used.
"""
from __future__ import annotations

import importlib.util


def share_torch_hip_runtime() -> bool:
    if importlib.util.find_spec("torch") is None:
        return False
    import torch  # noqa: F401

    return True
