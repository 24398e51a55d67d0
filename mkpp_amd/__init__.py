"""MI355X-native per-cell stiff chemistry integrator (MKPP hot path).

Subpackages:
  mkpp_amd.gen     mechanism compiler (OpenAtmos JSON -> plans -> HIP kernels)
  mkpp_amd.native  ctypes binding of the C ABI (libmkpp_amd.so)
  mkpp_amd.build   in-tree hipcc build
"""

__all__ = ["native"]
