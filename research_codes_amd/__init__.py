"""MI355X (gfx950) SPH interaction hot path for FDPS-style stellar-collision runs.

The compute lives in libsphk.so (HIP kernels + plain-C ABI, include/sphk.h);
this package holds the host-side Python mirror of the reference's multiwalk
functor interface (device.py), the host tree/list builder binding (tree.py),
the initial-condition generator (ic.py) and the step driver (step.py).
"""
from ._lib import LIB_PATH, SphkError, lib  # noqa: F401

__all__ = ["LIB_PATH", "SphkError", "lib"]
