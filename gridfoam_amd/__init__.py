"""gridfoam_amd -- MI355X-native pressure-velocity path of gridfoam.

Octree face-neighbour operator, FV assembly and Krylov solves run as
hand-written gfx950 HIP kernels (libgridfoam_hip.so, C ABI in
include/gridfoam_hip.h); this package is the host-side mirror of the
reference's Python interfaces (grid, fields, FvMatrix, fvm/fvc, solvers,
SIMPLE/PISO drivers).
"""

from .config import GridfoamConfig, SolverConfig, load_config  # noqa: F401

__all__ = ["GridfoamConfig", "SolverConfig", "load_config"]
