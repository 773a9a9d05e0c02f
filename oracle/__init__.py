"""CPU FP64 oracle for the explicit-step hot path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy restatement of the reference solver's explicit time step
(`wmles/gas.py`, `wmles/mesh.py`, `wmles/bc.py`, `wmles/schemes.py`,
`wmles/viscous.py`, `wmles/sgs.py`, `wmles/wall_model.py`, `wmles/solver.py`,
`wmles/smooth_ramp.py:116-221`).  Every function cites the reference lines it
follows.  It is the checker the HIP product is compared against; it is never the
thing measured or shipped.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import it.  The product package
(`wmles_jax_amd`) never imports it and fails loudly when its HIP library is absent.

How each function is pinned (a plausible mistake fails at least one check):

* composed RHS, ghost sweep, wall flux, advance, max_dt, wall_stats, inflow:
  golden vectors written from the reference itself (NumPy backend) by
  `tests/golden/make_golden.py` -> `tests/golden/*.npz`;
* geometry: closure (sum of a cell's outward face vectors = 0), exact box volumes,
  total volume, free-stream preservation on skewed periodic grids;
* exchange / orientation: unique-id multiblock == single-block ghost maps (exact);
* far field: the four known-answer cases of the reference test (rtol 1e-12);
* Spalding Newton: bisection inversion (|du_tau| < 1e-6);
* Kader, SGS, Ducros, KG flux: closed-form special cases (pure shear, pure
  rotation, pure dilatation, constant state -> Euler flux);
* MUSCL-HLLC: Sod shock-tube plateau values against the exact Riemann solution;
* viscous flux: exact stress / heat flux for linear fields on a Cartesian grid.

`KEEP` (reference `wmles/schemes.py:25-55,72-98`, not exported by the reference's
`__init__`) is pinned by the golden case `single_keep4_wmadi` and by the
constant-state and free-stream checks.
"""
from .gas import Gas
from .mesh import Block, Mesh, Interface, Geom, compute_geometry, exchange, FACE_NAMES, sl
from .bc import (SlipWall, NoSlipWall, Dirichlet, CharacteristicFarfield, Extrapolate,
                 WallModel)
from .numerics import (EquilibriumWallModel, KEP, KEEP, MUSCLHLLC, Hybrid, make_sgs,
                       green_gauss, ducros, viscous_flux, hllc, kg_flux)
from .solver import Solver, ConstantMassFlux
from .ramp import SmoothRampInflow, SmoothRampSponge, flow_properties

__all__ = [
    "Gas", "Block", "Mesh", "Interface", "Geom", "compute_geometry", "exchange",
    "FACE_NAMES", "sl", "SlipWall", "NoSlipWall", "Dirichlet", "CharacteristicFarfield",
    "Extrapolate", "WallModel", "EquilibriumWallModel", "KEP", "KEEP", "MUSCLHLLC",
    "Hybrid", "make_sgs", "green_gauss", "ducros", "viscous_flux", "hllc", "kg_flux",
    "Solver", "ConstantMassFlux", "SmoothRampInflow", "SmoothRampSponge",
    "flow_properties",
]
