"""astaroth_amd -- MI355X-native (gfx950) implementation of Astaroth's single-device MHD hot
path: 6th-order solve, 2N-RK3 substep, periodic halo fill and device reductions, behind the
reference's device C API (include/astaroth.h). `astaroth_amd.astaroth` is the ctypes mirror of
that API."""
from . import astaroth  # noqa: F401
from .astaroth import load, HostMesh, Device  # noqa: F401
