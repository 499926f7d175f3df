"""The reference's public hot-path API on ``Particles`` (trailing ``_`` marks
Julia's ``!``).  Each function performs exactly one C-ABI call (tuples of
fields on split containers: one per field); containers in the FP32-split form
(``particles.form == "split"``) go to the ``*_split`` entry points."""
from __future__ import annotations

import ctypes as C

from . import _abi
from .device import DeviceArray
from .particles import CellArray, Particles


class AbstractAdvectionIntegrator:
    code = -1


class Euler(AbstractAdvectionIntegrator):
    """``Euler()`` (src/Advection/types.jl:113-115); arguments are ignored."""
    code = _abi.JP_EULER

    def __init__(self, *args):
        self.alpha = 0.5


class RungeKutta2(AbstractAdvectionIntegrator):
    """``RungeKutta2(α = 0.5)`` (src/Advection/types.jl:125-136)."""
    code = _abi.JP_RK2

    def __init__(self, alpha=0.5):
        if not (0 < alpha < 1):
            raise ValueError("Only 0 < α < 1 is supported")
        self.alpha = float(alpha)


class RungeKutta4(AbstractAdvectionIntegrator):
    """``RungeKutta4()`` (src/Advection/types.jl:143-145)."""
    code = _abi.JP_RK4

    def __init__(self, *args):
        self.alpha = 0.5


def _ptrs(arrs):
    return _abi.ptr_array([a.ptr for a in arrs])


def advection_(particles: Particles, method: AbstractAdvectionIntegrator, V, dt):
    """``advection!(particles, method, V, dt)``
    (src/Particles/Advection/advection.jl:21-62).  V: tuple of DeviceArrays with
    the shapes of ``particles.xi_vel[c]``."""
    D = particles.D
    if len(V) != D:
        raise ValueError(f"expected {D} velocity components")
    for c in range(D):
        want = tuple(a.n for a in particles.xi_vel[c])
        if tuple(V[c].shape) != want:
            raise ValueError(f"V[{c}] has shape {V[c].shape}, expected {want}")
    ctx = particles.ctx
    fn = ctx.lib.jp_advection_split if particles.form == "split" else ctx.lib.jp_advection
    ctx.check(fn(ctx.h, particles.grid_ref(), particles.coord_ptrs(),
                                   particles.index.ptr, _ptrs(V), particles.velgrid_ref(),
                                   method.code, method.alpha, float(dt)), "advection!")


def move_particles_(particles: Particles, args=(), stats: bool = False):
    """``move_particles!(particles, args)`` (src/Particles/move_safe.jl:21-49).
    Returns the outcome counters when ``stats`` (costs one host sync)."""
    ctx = particles.ctx
    st = _abi.jp_stats() if stats else None
    fn = (ctx.lib.jp_move_particles_split if particles.form == "split"
          else ctx.lib.jp_move_particles)
    ctx.check(fn(ctx.h, particles.grid_ref(), particles.coord_ptrs(),
                                        particles.index.ptr, len(args), _ptrs(args),
                                        C.byref(st) if st is not None else None),
              "move_particles!")
    return st.as_dict() if st is not None else None


def _check_nodal(F, particles):
    want = tuple(n + 1 for n in particles.ni)
    if tuple(F.shape) != want:
        raise ValueError(f"nodal field has shape {F.shape}, expected {want}")


def particle2grid_(F, Fp, particles: Particles):
    """``particle2grid!(F, Fp, particles)`` (src/Interpolations/particle_to_grid.jl:23-28);
    F/Fp may be tuples (tuple methods :70-109, :153-199)."""
    ctx = particles.ctx
    if isinstance(F, (tuple, list)) and particles.form == "split":
        if len(F) != len(Fp):
            raise ValueError("F and Fp tuples differ in length")
        for f, fp in zip(F, Fp):
            particle2grid_(f, fp, particles)
        return
    if isinstance(F, (tuple, list)):
        if len(F) != len(Fp):
            raise ValueError("F and Fp tuples differ in length")
        for f in F:
            _check_nodal(f, particles)
        ctx.check(ctx.lib.jp_particle2grid_multi(ctx.h, particles.grid_ref(), len(F), _ptrs(F),
                                                 _ptrs(Fp), particles.coord_ptrs(),
                                                 particles.index.ptr), "particle2grid!")
    else:
        _check_nodal(F, particles)
        fn = (ctx.lib.jp_particle2grid_split if particles.form == "split"
              else ctx.lib.jp_particle2grid)
        ctx.check(fn(ctx.h, particles.grid_ref(), F.ptr, Fp.ptr,
                                           particles.coord_ptrs(), particles.index.ptr),
                  "particle2grid!")


def grid2particle_(Fp, F, particles: Particles):
    """``grid2particle!(Fp, F, particles)`` (src/Interpolations/grid_to_particle.jl:26-35)."""
    ctx = particles.ctx
    if isinstance(F, (tuple, list)):
        for fp, f in zip(Fp, F):
            grid2particle_(fp, f, particles)
        return
    _check_nodal(F, particles)
    fn = (ctx.lib.jp_grid2particle_split if particles.form == "split"
          else ctx.lib.jp_grid2particle)
    ctx.check(fn(ctx.h, particles.grid_ref(), Fp.ptr, F.ptr,
                                       particles.coord_ptrs(), particles.index.ptr),
              "grid2particle!")


def inject_particles_(particles: Particles, args=(), stats: bool = False):
    """``inject_particles!(particles, args)`` (src/Particles/injection.jl:19-53).
    Each call advances ``particles.inject_counter`` (the Philox counter)."""
    ctx = particles.ctx
    st = _abi.jp_stats() if stats else None
    fn = (ctx.lib.jp_inject_particles_split if particles.form == "split"
          else ctx.lib.jp_inject_particles)
    ctx.check(fn(ctx.h, particles.grid_ref(), particles.coord_ptrs(),
                                          particles.index.ptr, len(args), _ptrs(args),
                                          particles.min_xcell, particles.seed,
                                          particles.inject_counter,
                                          C.byref(st) if st is not None else None),
              "inject_particles!")
    particles.inject_counter += 1
    return st.as_dict() if st is not None else None


def nodal_field(particles: Particles, values=None, fill=0.0) -> DeviceArray:
    """A nodal (vertex) field of shape ni+1, e.g. the ``T`` of the examples."""
    shape = tuple(n + 1 for n in particles.ni)
    if values is not None:
        return DeviceArray.from_host(particles.ctx, values, dtype="float64")
    return DeviceArray(particles.ctx, shape).fill(fill)


__all__ = ["Euler", "RungeKutta2", "RungeKutta4", "AbstractAdvectionIntegrator", "advection_",
           "move_particles_", "particle2grid_", "grid2particle_", "inject_particles_",
           "nodal_field", "CellArray"]
