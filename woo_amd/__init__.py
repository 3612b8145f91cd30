"""woo_amd: MI355X (gfx950) DEM step engine behind Woo's engine/functor surface.

The compute path is libwoo_amd.so (hand-written HIP kernels behind the C ABI in
include/woo_amd.h); ``woo_amd._native`` binds it. There is no CPU fallback.
"""
__version__ = "0.1.0"
