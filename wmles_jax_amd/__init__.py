"""wmles_jax_amd -- MI355X-native explicit time step of a compressible wall-modelled
LES finite-volume solver, drop-in for the reference's `wmles` step API
(reference `wmles/__init__.py:1-9`).

The step (ghost sweep, fluxes, gradients, SGS, wall model, SSP-RK3, CFL dt, mass-flux
constraint) runs in hand-written HIP kernels for gfx950 (libwmles_amd.so, C ABI in
include/wmles_amd.h).  Importing this package without the built library raises.
"""
from . import _lib  # noqa: F401  (fails loudly if libwmles_amd.so is missing)
from .backend import HAS_JAX
from .gas import Gas
from .mesh import Block, Mesh, Geom, Interface, FACE_NAMES, sl
from .bc import BC, SlipWall, NoSlipWall, Dirichlet, CharacteristicFarfield, Extrapolate, WallModel
from .wall_model import EquilibriumWallModel
from .learned_wm import LearnedWallModel
from .schemes import KEP, KEEP, MUSCLHLLC, Hybrid
from .sgs import make_sgs
from .solver import Solver, ConstantMassFlux
from . import smooth_ramp
from .smooth_ramp import SmoothRampInflow, SmoothRampSponge, flow_properties
from .ensemble import EnsembleSolver
from .stats import PlaneStats
from . import io_vtk
from .plot3d import read_plot3d, write_plot3d
from .nmf import read_nmf, respan, extrude_2d, build_mesh, describe
from . import autotune

__all__ = ["HAS_JAX", "Gas", "Block", "Mesh", "Geom", "Interface", "FACE_NAMES", "sl", "BC",
           "SlipWall", "NoSlipWall", "Dirichlet", "CharacteristicFarfield", "Extrapolate",
           "WallModel", "EquilibriumWallModel", "LearnedWallModel", "KEP", "KEEP", "MUSCLHLLC", "Hybrid", "make_sgs",
           "Solver", "ConstantMassFlux", "smooth_ramp", "SmoothRampInflow", "SmoothRampSponge",
           "flow_properties", "EnsembleSolver", "PlaneStats", "io_vtk", "autotune", "read_plot3d", "write_plot3d",
           "read_nmf", "respan", "extrude_2d", "build_mesh", "describe"]
