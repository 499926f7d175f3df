"""Host-side mirror of JustPIC.jl's particle hot-path API (v0.6.7) over the
MI355X HIP C-ABI in ``libjustpic_hip.so``.

Reference names map as ``f!`` -> ``f_``: ``advection_``, ``move_particles_``,
``particle2grid_``, ``grid2particle_``, ``inject_particles_``; containers are
``Particles``/``CellArray``; integrators ``Euler``, ``RungeKutta2``,
``RungeKutta4``.  ``init_particles(..., form="split")`` / ``to_split`` give the
FP32-split position form, accepted by the same functions.  ``Ensemble`` runs
many independent grids in one launch; ``rrule_advection_p2g`` is the reverse
mode of advection + particle2grid.  Importing this package does not load the
library; the first ``Context`` does, and raises if it is not built.
"""
from ._abi import HipLibraryMissing, load_library
from .adjoint import AdvectionP2G, flatten, l2_loss, rrule_advection_p2g
from .api import (AbstractAdvectionIntegrator, Euler, RungeKutta2, RungeKutta4, advection_,
                  grid2particle_, inject_particles_, move_particles_, nodal_field,
                  particle2grid_)
from .device import Context, DeviceArray, JustPICError
from .ensemble import Ensemble, random_ensemble
from .grids import Axis, expand_range, linrange, staggered_grid, steprange
from .particles import (CellArray, Particles, SplitParticles, init_cell_arrays,
                        init_particles, particles_from_host, to_f64, to_split)

__all__ = [
    "AbstractAdvectionIntegrator", "AdvectionP2G", "flatten", "l2_loss", "rrule_advection_p2g",
    "Axis", "CellArray", "Context", "DeviceArray", "Ensemble",
    "Euler", "random_ensemble",
    "HipLibraryMissing", "JustPICError", "Particles", "RungeKutta2", "RungeKutta4",
    "advection_", "expand_range", "grid2particle_", "init_cell_arrays", "init_particles",
    "inject_particles_", "linrange", "load_library", "move_particles_", "nodal_field",
    "particle2grid_", "particles_from_host", "SplitParticles", "staggered_grid", "steprange",
    "to_f64", "to_split",
]
