"""MI355X-native position-based-dynamics cloth substep.

The product is libpbd.so (HIP kernels for gfx950 + C++ host code behind the C
ABI in include/pbd.h).  `capi` is the ctypes binding used by Python callers
and the test-suite; `build` compiles the library in-tree.
"""
__all__ = ["capi", "build"]
