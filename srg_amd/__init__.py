"""MI355X time-domain backprojection engine (HIP, gfx950).

``srg_amd.tdbp``  -- ctypes binding of libsrg_tdbp.so (include/tdbp.h).
``srg_amd.synth`` -- synthetic Sentinel-1-shaped bursts for tests and benches.
"""
from . import tdbp  # noqa: F401

__all__ = ["tdbp"]
