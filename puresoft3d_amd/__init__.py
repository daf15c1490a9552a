"""puresoft3d_amd: MI355X (gfx950) raster hot path of Puresoft3D.

The compute path is libpsgpu.so (hand-written HIP kernels behind the C ABI of
include/psgpu.h); `native` binds it with ctypes, `scenes` builds the seeded
benchmark scenes.  Importing this package does not touch the GPU.
"""
__all__ = ["native", "scenes"]
