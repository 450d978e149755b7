"""asend3d_amd -- MI355X (gfx950) solve path for ASenD3D elastic models.

The compute lives in libasend_gpu.so (HIP kernels behind the C ABI of
include/asend_gpu.h); `asend3d_amd.api` binds it with ctypes and raises at
import when the library is not built.  `meshgen` and `problems` synthesise the
structured workloads.
"""
__all__ = ["api", "meshgen", "problems", "build"]
