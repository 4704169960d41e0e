"""MI355X (gfx950) population scoring and FCC clustering for gdock.

The compute path is libgdock_hip.so (hand-written HIP kernels behind the C
ABI in include/gdock_hip.h); the Python modules mirror the reference's Rust
modules (structure, fitness, restraints, population, clustering, constants).
"""
from . import constants
from ._lib import ATOM_DTYPE, GdockError, lib

__all__ = ["constants", "ATOM_DTYPE", "GdockError", "lib"]
