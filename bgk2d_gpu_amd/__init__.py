"""bgk2d_gpu_amd: MI355X (gfx950) implementation of BGK2D's time-step loop.

The product is the C-ABI library libbgk2d_hip.so (include/bgk2d.h) and the
Fortran driver that calls it through ISO_C_BINDING; this package holds their
sources (csrc/, driver/), the build recipe and a ctypes view used by the
tests, the benchmark and the serving loop.
"""
from .build import build_all, build_lib, have_lib  # noqa: F401
from .solver import Solver, BgkError, load, EXPORTS  # noqa: F401

__all__ = ["Solver", "BgkError", "load", "build_all", "build_lib", "have_lib", "EXPORTS"]
