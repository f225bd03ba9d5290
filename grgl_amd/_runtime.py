"""One HIP runtime per process.

PyTorch-ROCm wheels bundle their own libamdhip64 (same SONAME, libamdhip64.so.7, as the
system ROCm's).  libtorch_hip links it by its unversioned name, so if libgrgl_hip.so
loads the system runtime first, importing torch later brings in a second HIP/HSA runtime
and torch finds no device.  Importing torch first makes libgrgl_hip.so bind to the
runtime torch already loaded, so both share one device context and torch tensors can be
passed to the C ABI.  Set GRGL_AMD_NO_TORCH=1 to skip this (pure C-ABI users).
"""

import os


def share_hip_runtime_with_torch():
    if os.environ.get("GRGL_AMD_NO_TORCH") == "1":
        return False
    try:
        import torch  # noqa: F401
    except Exception:
        return False
    return True
