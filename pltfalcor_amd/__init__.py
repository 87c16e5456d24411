"""pltfalcor_amd — MI355X-native PLT generalized-ray (spectral, polarimetric, wave-optics)
path-tracing hot path.

The compute lives in libplt_amd.so (hand-written HIP for gfx950, C ABI in include/plt_abi.h).
This package is the host-side mirror of the reference's interfaces over that ABI.
"""
from . import _lib  # noqa: F401
from ._lib import PLTError, load  # noqa: F401

__all__ = ["PLTError", "load"]
