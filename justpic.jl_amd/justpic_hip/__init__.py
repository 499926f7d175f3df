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
from .api import (AbstractAdvectionIntegrator, BacktrackGrid, Euler, RungeKutta2, RungeKutta4,
                  advection_, advection_LinP_, advection_MQS_, cell_index, centroid2particle_,
                  centroid_field, grid2particle_, grid2particle_flip_, inject_particles_,
                  move_particles_, nodal_field, particle2centroid_, particle2grid_,
                  semilagrangian_advection_, semilagrangian_advection_LinP_,
                  semilagrangian_advection_MQS_)
from .device import Context, DeviceArray, JustPICError
from .ensemble import Ensemble, random_ensemble
from .grids import Axis, compute_dx, expand_range, linrange, staggered_grid, steprange
from .io import (cellarray_from_host, cellarray_to_host, checkpoint_name,
                 checkpointing_particles, load_checkpoint, particles_from_host_layout,
                 particles_to_host)
from .phases import PhaseRatios, inject_particles_phase_, update_phase_ratios_
from .particles import (CellArray, Particles, SplitParticles, init_cell_arrays,
                        init_particles, particles_from_host, to_f64, to_split)

__all__ = [
    "AbstractAdvectionIntegrator", "AdvectionP2G", "Axis", "BacktrackGrid", "CellArray", "Context",
    "DeviceArray", "Ensemble", "Euler", "HipLibraryMissing", "JustPICError", "Particles",
    "PhaseRatios", "RungeKutta2", "RungeKutta4", "SplitParticles",
    "advection_", "advection_LinP_", "advection_MQS_", "cell_index", "cellarray_from_host",
    "cellarray_to_host", "centroid2particle_", "centroid_field", "checkpoint_name",
    "checkpointing_particles", "compute_dx", "expand_range", "flatten", "grid2particle_",
    "grid2particle_flip_", "init_cell_arrays", "init_particles", "inject_particles_",
    "inject_particles_phase_", "l2_loss", "linrange", "load_checkpoint", "load_library",
    "move_particles_", "nodal_field", "particle2centroid_", "particle2grid_",
    "particles_from_host", "particles_from_host_layout", "particles_to_host", "random_ensemble",
    "rrule_advection_p2g", "semilagrangian_advection_", "semilagrangian_advection_LinP_",
    "semilagrangian_advection_MQS_", "staggered_grid", "steprange", "to_f64", "to_split",
    "update_phase_ratios_",
]
