"""MI355X-native (gfx950) implementation of the FZGPUModules stage-execution hot path.

The product is ``libfzh.so``: hand-written HIP kernels behind the C ABI in
``include/fzh.h`` plus the C++ ``fz::`` Stage/Pipeline layer.  This Python
package only builds and binds it (ctypes); it has no CPU fallback.
"""
from ._fzh import DeviceBuffer, FzhError, lib, workspace  # noqa: F401

__all__ = ["DeviceBuffer", "FzhError", "lib", "workspace"]
