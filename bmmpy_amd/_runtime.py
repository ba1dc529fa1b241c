"""Selects the process's HIP runtime before libbmmpy_hip.so loads.

torch ships its own HIP runtime under ROCm's soname.  When torch is installed
it is imported first, so the library binds to that single runtime and device
pointers from torch tensors work with the library's kernels; loading the
library first would put two HIP runtimes into one process and leave torch
without a device.  Set BMMPY_NO_TORCH_PRELOAD=1 to skip this.
"""
import os

if not os.environ.get("BMMPY_NO_TORCH_PRELOAD"):
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
