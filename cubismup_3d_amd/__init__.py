"""cubismup_3d_amd -- CubismUP_3D's per-timestep periodic fluid update on AMD Instinct
MI355X: host C++ Simulation/Operator classes over a C-ABI (include/cup.h) of hand-written
HIP kernels for gfx950, with Python bindings mirroring the reference's ``cubismup3d``."""
from ._lib import lib, CupError  # noqa: F401
from .grid import Grid, to_blocks, from_blocks  # noqa: F401
from .simulation import Simulation, SimulationData, Sphere, read_field_file, read_xdmf  # noqa: F401

__all__ = ["lib", "CupError", "Grid", "to_blocks", "from_blocks", "Simulation", "SimulationData", "Sphere",
           "read_field_file", "read_xdmf"]
