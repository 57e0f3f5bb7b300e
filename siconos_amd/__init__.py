"""siconos_amd — the 3-D friction-contact NSGS hot path of Siconos numerics,
rebuilt for one MI355X (gfx950): a C host layer with the reference's
FrictionContactProblem / SolverOptions / fc3d_driver interface, a new solver
id SICONOS_FRICTION_3D_NSGS_HIP (540), and hand-written HIP kernels behind a
thin C-ABI (include/fc3d_numerics.h, include/fc3d_hip.h).
"""
from . import capi  # noqa: F401

__all__ = ["capi", "numerics"]
