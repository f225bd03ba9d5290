"""One HIP runtime per process, without importing torch.

PyTorch-ROCm wheels bundle their own libamdhip64 (torch/lib/libamdhip64.so, SONAME
libamdhip64.so.7, the same SONAME as the system ROCm's).  libtorch_hip names it by its
unversioned file name, so if libgrgl_hip.so loaded the system runtime first, a later
`import torch` would bring in a second HIP/HSA runtime and torch tensors could not be passed
to the C ABI.

Loading torch's bundled runtime file (by path, globally) before libgrgl_hip.so avoids that
without importing torch: libgrgl_hip.so's libamdhip64.so.7 then resolves to the runtime
already loaded, and a later `import torch` finds that same file loaded.  torch is located
with importlib (its package is not executed), so numpy callers pay nothing for it.  Set
GRGL_AMD_NO_TORCH=1 to use the system runtime regardless (pure C-ABI or numpy users that
never touch torch in the process).
"""

import ctypes
import importlib.util
import os

_done = None


def torch_hip_runtime_path():
    """torch's bundled libamdhip64.so, or None when torch (ROCm build) is not installed."""
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.origin:
        return None
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    return path if os.path.isfile(path) else None


def share_hip_runtime_with_torch():
    """Load torch's HIP runtime globally (once) if torch is installed; returns whether it did."""
    global _done
    if _done is not None:
        return _done
    _done = False
    if os.environ.get("GRGL_AMD_NO_TORCH") == "1":
        return _done
    path = torch_hip_runtime_path()
    if path is not None:
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        _done = True
    return _done
