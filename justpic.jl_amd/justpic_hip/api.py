"""The reference's public hot-path API on ``Particles`` (trailing ``_`` marks
Julia's ``!``).  Each function performs exactly one C-ABI call (tuples of
fields on split containers: one per field); containers in the FP32-split form
(``particles.form == "split"``) go to the ``*_split`` entry points."""
from __future__ import annotations

import ctypes as C

import numpy as np

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


def _advection_scheme(particles: Particles, method, V, dt, scheme: int, name: str):
    if particles.form == "split":
        raise ValueError(f"{name} runs on FP64 positions (to_f64 first)")
    D = particles.D
    if len(V) != D:
        raise ValueError(f"expected {D} velocity components")
    for c in range(D):
        want = tuple(a.n for a in particles.xi_vel[c])
        if tuple(V[c].shape) != want:
            raise ValueError(f"V[{c}] has shape {V[c].shape}, expected {want}")
    ctx = particles.ctx
    ctx.check(ctx.lib.jp_advection_scheme(ctx.h, particles.grid_ref(), particles.coord_ptrs(),
                                          particles.index.ptr, _ptrs(V),
                                          particles.velgrid_ref(), method.code, method.alpha,
                                          float(dt), scheme), name)


def advection_MQS_(particles: Particles, method: AbstractAdvectionIntegrator, V, dt):
    """``advection_MQS!(particles, method, V, dt)``
    (src/Particles/Advection/advection_MQS.jl:16-69): ``advection!`` with the
    monotonic quadratic spline velocity interpolation (src/Interpolations/MQS.jl)."""
    _advection_scheme(particles, method, V, dt, _abi.JP_INTERP_MQS, "advection_MQS!")


def advection_LinP_(particles: Particles, method: AbstractAdvectionIntegrator, V, dt):
    """``advection_LinP!(particles, method, V, dt)``
    (src/Particles/Advection/advection_LinP.jl:13-55): ``advection!`` with the
    linear-plus-pressure-point velocity interpolation."""
    _advection_scheme(particles, method, V, dt, _abi.JP_INTERP_LINP, "advection_LinP!")


class BacktrackGrid:
    """Device copies of the grids of ``semilagrangian_advection!``: the nodal
    (vertex) axes ``grid`` and the velocity axes ``grid_vi``, both with the
    ``compute_dx`` spacing rule (src/Particles/utils.jl:46-49).  Built once
    and reused; the functions below also accept the host axes and build one
    per call."""

    def __init__(self, ctx, grid_vi, grid):
        from .grids import compute_dx
        from .particles import DeviceAxis
        self.ctx = ctx
        self.D = len(grid)
        if len(grid_vi) != self.D or any(len(g) != self.D for g in grid_vi):
            raise ValueError("grid_vi must hold D axes for each of the D components")
        self.grid = tuple(DeviceAxis(ctx, compute_dx(a)) for a in grid)
        self.grid_vi = tuple(tuple(DeviceAxis(ctx, compute_dx(a)) for a in g) for g in grid_vi)
        self.shape = tuple(a.n for a in self.grid)
        self._g = (_abi.jp_axis * 3)(*[a.abi() for a in self.grid])
        self._vg = _abi.jp_velgrid()
        for c in range(self.D):
            for d in range(self.D):
                self._vg.ax[c][d] = self.grid_vi[c][d].abi()


def _semilagrangian(F, F0, method, V, grid_vi, grid, dt, scheme, name, ctx=None):
    tup = isinstance(F, (tuple, list))
    Fs, F0s = (tuple(F), tuple(F0)) if tup else ((F,), (F0,))
    if len(Fs) != len(F0s) or not Fs:
        raise ValueError(f"{name}: F and F0 must hold the same number of fields")
    if method.code == _abi.JP_EULER:
        raise ValueError(f"{name} takes RungeKutta2 or RungeKutta4")
    ctx = ctx or Fs[0].ctx
    bg = grid if isinstance(grid, BacktrackGrid) else BacktrackGrid(ctx, grid_vi, grid)
    for f in Fs + F0s:
        if tuple(f.shape) != bg.shape:
            raise ValueError(f"{name}: field shape {f.shape}, expected {bg.shape}")
    for c in range(bg.D):
        want = tuple(a.n for a in bg.grid_vi[c])
        if tuple(V[c].shape) != want:
            raise ValueError(f"{name}: V[{c}] has shape {V[c].shape}, expected {want}")
    # the NTuple linear method takes the spacing at the node (backtracking.jl:79)
    at_node = int(tup and scheme == _abi.JP_INTERP_LINEAR)
    ctx.check(ctx.lib.jp_semilagrangian(ctx.h, bg.D, C.cast(bg._g, C.c_void_p), C.byref(bg._vg),
                                        _ptrs(V), len(Fs), _ptrs(Fs), _ptrs(F0s), method.code,
                                        method.alpha, float(dt), scheme, at_node), name)


def semilagrangian_advection_(F, F0, method: AbstractAdvectionIntegrator, V, grid_vi, grid, dt):
    """``semilagrangian_advection!(F, F0, method, V, grid_vi, grid, dt)``
    (src/Particles/Advection/backtracking.jl:15-101).  F, F0: nodal
    DeviceArrays (or tuples of them) on ``grid``; grid_vi, grid: host
    :class:`Axis` tuples or a :class:`BacktrackGrid` as ``grid``.  Every
    method reads F0 (the reference's scalar method reads F while writing it);
    a backtracked point outside the grid raises."""
    _semilagrangian(F, F0, method, V, grid_vi, grid, dt, _abi.JP_INTERP_LINEAR,
                    "semilagrangian_advection!")


def semilagrangian_advection_MQS_(F, F0, method: AbstractAdvectionIntegrator, V, grid_vi, grid,
                                  dt):
    """``semilagrangian_advection_MQS!`` (src/Particles/Advection/backtracking_MQS.jl)"""
    _semilagrangian(F, F0, method, V, grid_vi, grid, dt, _abi.JP_INTERP_MQS,
                    "semilagrangian_advection_MQS!")


def semilagrangian_advection_LinP_(F, F0, method: AbstractAdvectionIntegrator, V, grid_vi, grid,
                                   dt):
    """``semilagrangian_advection_LinP!`` (src/Particles/Advection/backtracking_LinP.jl)"""
    _semilagrangian(F, F0, method, V, grid_vi, grid, dt, _abi.JP_INTERP_LINP,
                    "semilagrangian_advection_LinP!")


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


def _fields(x):
    return tuple(x) if isinstance(x, (tuple, list)) else (x,)


def grid2particle_flip_(Fp, F, F0, particles: Particles, alpha=0.0):
    """``grid2particle_flip!(Fp, xvi, F, F0, particles; α)``
    (src/Interpolations/grid_to_particle.jl:125-207): PIC/FLIP blend of the
    nodal F and F0 into Fp (α = 1 PIC, α = 0 FLIP); tuples field by field."""
    if particles.form != "f64":
        raise ValueError("grid2particle_flip! needs FP64 particle coordinates")
    Fp, F, F0 = _fields(Fp), _fields(F), _fields(F0)
    if not (len(Fp) == len(F) == len(F0)):
        raise ValueError("Fp, F and F0 differ in length")
    for f in F + F0:
        _check_nodal(f, particles)
    ctx = particles.ctx
    ctx.check(ctx.lib.jp_grid2particle_flip(ctx.h, particles.grid_ref(), len(Fp), _ptrs(Fp),
                                            _ptrs(F), _ptrs(F0), particles.coord_ptrs(),
                                            particles.index.ptr, float(alpha)),
              "grid2particle_flip!")


def _check_centred(F, particles):
    if tuple(F.shape) != tuple(particles.ni):
        raise ValueError(f"cell-centred field has shape {F.shape}, expected {particles.ni}")


def particle2centroid_(F, Fp, particles: Particles):
    """``particle2centroid!(F, Fp, particles)``
    (src/Interpolations/particle_to_grid_centroid.jl:12-128): cell-centred F
    (shape ni) from the particles of each cell; a tuple selects the
    reference's NTuple method (inverse-distance weights in 2D)."""
    if particles.form != "f64":
        raise ValueError("particle2centroid! needs FP64 particle coordinates")
    tup = isinstance(F, (tuple, list))
    F, Fp = _fields(F), _fields(Fp)
    if len(F) != len(Fp):
        raise ValueError("F and Fp differ in length")
    for f in F:
        _check_centred(f, particles)
    ctx = particles.ctx
    ctx.check(ctx.lib.jp_particle2centroid(ctx.h, particles.grid_ref(),
                                           particles.centre_axes_ref(), len(F), _ptrs(F),
                                           _ptrs(Fp), particles.coord_ptrs(), 1 if tup else 0),
              "particle2centroid!")


def centroid2particle_(Fp, F, particles: Particles):
    """``centroid2particle!(Fp, F, particles)``
    (src/Interpolations/centroid_to_particle.jl:15-76)."""
    if particles.form != "f64":
        raise ValueError("centroid2particle! needs FP64 particle coordinates")
    F, Fp = _fields(F), _fields(Fp)
    if len(F) != len(Fp):
        raise ValueError("F and Fp differ in length")
    for f in F:
        _check_centred(f, particles)
    ctx = particles.ctx
    ctx.check(ctx.lib.jp_centroid2particle(ctx.h, particles.grid_ref(),
                                           particles.centre_axes_ref(), len(F), _ptrs(Fp),
                                           _ptrs(F), particles.coord_ptrs()),
              "centroid2particle!")


def centroid_field(particles: Particles, values=None, fill=0.0) -> DeviceArray:
    """A cell-centred field of shape ni."""
    if values is not None:
        return DeviceArray.from_host(particles.ctx, values, dtype="float64")
    return DeviceArray(particles.ctx, tuple(particles.ni)).fill(fill)


def cell_index(x, xv, ctx=None):
    """``cell_index(x::NTuple, xv::NTuple)`` (src/particles.jl:205-224) for
    arrays of points: per dimension, ``abs(trunc((x - xv[1]) / |xv[2] - xv[1]|)) + 1``
    (1-based, as the reference returns it).  ``x``: tuple of float64
    DeviceArrays of equal length; ``xv``: tuple of host :class:`Axis` (or
    ``DeviceAxis``).  Returns a tuple of int64 DeviceArrays."""
    from .particles import DeviceAxis
    if not isinstance(x, (tuple, list)) or len(x) != len(xv) or not 1 <= len(x) <= 3:
        raise ValueError("x and xv must be tuples of the same length (1 to 3)")
    ctx = ctx or x[0].ctx
    n = x[0].size
    if any(a.size != n for a in x):
        raise ValueError("all coordinate arrays must have the same length")
    dax = [a if isinstance(a, DeviceAxis) else DeviceAxis(ctx, a) for a in xv]
    axes = (_abi.jp_axis * len(dax))(*[a.abi() for a in dax])
    out = tuple(DeviceArray(ctx, (n,), np.int64) for _ in x)
    ctx.check(ctx.lib.jp_cell_index(ctx.h, len(x), axes, n, _ptrs(x), _ptrs(out)),
              "cell_index")
    return out


def nodal_field(particles: Particles, values=None, fill=0.0) -> DeviceArray:
    """A nodal (vertex) field of shape ni+1, e.g. the ``T`` of the examples."""
    shape = tuple(n + 1 for n in particles.ni)
    if values is not None:
        return DeviceArray.from_host(particles.ctx, values, dtype="float64")
    return DeviceArray(particles.ctx, shape).fill(fill)


__all__ = ["Euler", "RungeKutta2", "RungeKutta4", "AbstractAdvectionIntegrator", "advection_",
           "move_particles_", "particle2grid_", "grid2particle_", "inject_particles_",
           "nodal_field", "cell_index", "CellArray"]
