"""Material phases on particles: ``inject_particles_phase!`` and
``update_phase_ratios!`` with the ``PhaseRatios`` container.

``PhaseRatios`` (src/PhaseRatios/PhaseRatios.jl:12-33, constructors.jl:16-43)
holds, per grid location, the fraction of each phase among the particles
around it: cell centres, vertices, the staggered velocity faces and, in 3D,
the edge midpoints.  Each field is a device CellArray with ``nphases``
values per location (location l, phase k at ``data[l + L*k]``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .device import DeviceArray
from .particles import CellArray, Particles

FIELDS_2D = ("center", "vertex", "Vx", "Vy")
FIELDS_3D = ("center", "vertex", "Vx", "Vy", "Vz", "yz", "xz", "xy")


def ratio_dims(ni):
    """Location dims of every PhaseRatios field (constructors.jl:16-43)."""
    D = len(ni)
    dims = {"center": tuple(ni), "vertex": tuple(n + 1 for n in ni)}
    for d, name in enumerate(("Vx", "Vy", "Vz")[:D]):
        dims[name] = tuple(n + (1 if e == d else 0) for e, n in enumerate(ni))
    if D == 3:
        dims["yz"] = (ni[0], ni[1] + 1, ni[2] + 1)
        dims["xz"] = (ni[0] + 1, ni[1], ni[2] + 1)
        dims["xy"] = (ni[0] + 1, ni[1] + 1, ni[2])
    return dims


class PhaseRatios:
    """``PhaseRatios(backend, nphases, ni)``: zero-filled phase fractions."""

    def __init__(self, ctx, nphases: int, ni):
        if not 1 <= int(nphases) <= _abi.JP_MAX_PHASES:
            raise ValueError(f"nphases must be in 1..{_abi.JP_MAX_PHASES}")
        self.ctx = ctx
        self.nphases = int(nphases)
        self.ni = tuple(int(n) for n in ni)
        self.dims = ratio_dims(self.ni)
        names = FIELDS_2D if len(self.ni) == 2 else FIELDS_3D
        self.fields = {k: DeviceArray(ctx, (int(np.prod(self.dims[k])) * self.nphases,)).fill(0.0)
                       for k in names}

    def __getattr__(self, name):
        f = self.__dict__.get("fields", {})
        if name in f:
            return f[name]
        raise AttributeError(name)

    def to_host(self, name) -> np.ndarray:
        """Field `name` as an array of shape dims + (nphases,)."""
        L = int(np.prod(self.dims[name]))
        a = self.fields[name].to_host().reshape(self.nphases, L).T
        return a.reshape(self.dims[name] + (self.nphases,), order="F")


def _centre_axes(p: Particles):
    return p.centre_axes_ref()


def update_phase_ratios_(phase_ratios: PhaseRatios, particles: Particles, phases: CellArray):
    """``update_phase_ratios!(phase_ratios, particles, phases)``
    (src/PhaseRatios/utils.jl:15-43)."""
    if particles.form != "f64":
        raise ValueError("update_phase_ratios! needs FP64 particle coordinates")
    if tuple(phase_ratios.ni) != tuple(particles.ni):
        raise ValueError("phase ratios and particles live on different grids")
    ctx = particles.ctx
    D = particles.D
    pr = phase_ratios.fields
    faces = _abi.ptr_array([pr[k].ptr for k in ("Vx", "Vy", "Vz")[:D]])
    mids = _abi.ptr_array([pr[k].ptr for k in ("xy", "yz", "xz")]) if D == 3 else None
    ctx.check(ctx.lib.jp_update_phase_ratios(ctx.h, particles.grid_ref(), _centre_axes(particles),
                                             particles.coord_ptrs(), phases.ptr,
                                             phase_ratios.nphases, pr["center"].ptr,
                                             pr["vertex"].ptr, faces, mids),
              "update_phase_ratios!")


def inject_particles_phase_(particles: Particles, phases: CellArray, args=(), fields=(),
                            stats: bool = False):
    """``inject_particles_phase!(particles, phases, args, fields)``
    (src/Particles/injection.jl:146-325).  ``fields[a]`` (a DeviceArray) is
    cell-centred when its shape is ``ni``, nodal when it is ``ni + 1``; each
    call advances ``particles.inject_counter``."""
    if particles.form != "f64":
        raise ValueError("inject_particles_phase! needs FP64 particle coordinates")
    if len(args) != len(fields):
        raise ValueError("args and fields differ in length")
    ni = tuple(particles.ni)
    nodal = tuple(n + 1 for n in ni)
    centred = []
    for f in fields:
        if tuple(f.shape) == ni:
            centred.append(1)
        elif tuple(f.shape) == nodal:
            centred.append(0)
        else:
            raise ValueError(f"field of shape {f.shape} is neither cell-centred {ni} "
                             f"nor nodal {nodal}")
    ctx = particles.ctx
    st = _abi.jp_stats() if stats else None
    cen = (C.c_int * max(1, len(centred)))(*centred)
    ctx.check(ctx.lib.jp_inject_particles_phase(
        ctx.h, particles.grid_ref(), _centre_axes(particles), particles.coord_ptrs(),
        particles.index.ptr, phases.ptr, len(args), _abi.ptr_array([a.ptr for a in args]),
        _abi.ptr_array([f.ptr for f in fields]), cen, particles.min_xcell, particles.seed,
        particles.inject_counter, C.byref(st) if st is not None else None),
        "inject_particles_phase!")
    particles.inject_counter += 1
    return st.as_dict() if st is not None else None
