"""tarang_amd -- MI355X (gfx950) pseudo-spectral hot path of Tarang.

The compute lives in libtarang_amd.so (HIP kernels behind the C ABI of
include/tarang_gpu.h).  tarang_amd.api mirrors the reference's basis / field
/ solver classes on top of that ABI.
"""
__version__ = "0.1.0"


def load_library():
    from . import _lib
    return _lib.load()
