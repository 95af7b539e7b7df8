"""MI355X-native pks2map: Peak Patch halo->map painting on gfx950.

The compute lives in libp2m.so (HIP kernels behind the C ABI of
include/pks2map_gpu.h); `abi.Painter` drives it from Python, the Fortran
executables under bin/ drive it as the drop-in pks2map.
"""
__all__ = ["abi", "formats", "catalog"]
