"""One HIP runtime per process.

PyTorch wheels bundle their own HIP/HSA runtime.  Device pointers and streams
belong to one runtime instance, so a process that uses both torch and this
library must map a single runtime.  libmagnus_hip.so needs the runtime by its
SONAME (libamdhip64.so.7); if torch's copy is already mapped, the dynamic
loader binds the library to it.  This module therefore loads torch (when it is
installed) before the native library is loaded.  Without torch, the library
uses the system ROCm runtime.
"""
from __future__ import annotations


def bind_runtime() -> None:
    try:
        import torch  # noqa: F401  (maps torch's HIP runtime first)
    except ImportError:
        pass
