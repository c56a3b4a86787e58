"""ptrans -- host side of the MI355X phonon Monte Carlo trace.

The trace itself runs in libptrans.so (HIP, gfx950) behind the C-ABI declared in
include/ptrans.h; this package reads and writes P-TRANS's inputs (.mcp, .anphon),
builds the flattened geometry and calls the library through ctypes.
"""
from . import anphon, constants, gen, geometry, mcp  # noqa: F401

__all__ = ["anphon", "constants", "gen", "geometry", "mcp"]
