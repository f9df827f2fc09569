"""lighthugger_amd: the lighthugger GPU-driven frame, rebuilt for MI355X.

Native parts: liblh.so (HIP kernels for gfx950 behind the C ABI in
include/lh/lh.h) and liblh_scene.so (C++ scene producer, include/lh/lh_scene.h).
Python here is plumbing: loading the libraries, moving arrays to device
memory (via torch) and calling the C ABI.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
