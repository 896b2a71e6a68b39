"""fvs2d_amd -- MI355X-native residual / time-step path of the fvs2d 2-D
unstructured finite-volume Euler solver (hand-written gfx950 HIP, FP64)."""
from .meshgen import Mesh, jittered_square, naca_ogrid, read_grid, write_grid  # noqa: F401
from .solver import BatchSolver, HostMesh, Scheme, Solver, compute_flux_invscid, read_input, read_vortex  # noqa: F401
from ._lib import FvsError, RK_CLASSIC, RK_SSP3, RK_SSP42  # noqa: F401
