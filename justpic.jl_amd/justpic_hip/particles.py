"""Particles container and CellArrays on the HIP device.

Mirrors ``JustPIC.Particles`` (src/particles.jl:17-50) and the GPU
``CellArray`` layout of the reference's device backends
(ext/JustPICAMDGPUExt.jl:26-35: block length 0, data shape (C, K, 1)), i.e.
element (cell c, slot k) at ``data[c + C*k]`` with c the column-major cell
index.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from .device import Context, DeviceArray
from .grids import Axis


class CellArray:
    """K slots per cell over ``ni`` cells (B = 0 planes).  ``data`` is the flat
    device buffer; ``to_host()`` returns a (K, C) numpy array (slot-major)."""

    def __init__(self, ctx: Context, ni, K, dtype=np.float64, fill=None):
        self.ctx = ctx
        self.ni = tuple(int(n) for n in ni)
        self.K = int(K)
        self.C = int(np.prod(self.ni))
        self.data = DeviceArray(ctx, (self.C * self.K,), dtype)
        if fill is not None:
            self.data.fill(fill)

    @property
    def ptr(self):
        return self.data.ptr

    @property
    def dtype(self):
        return self.data.dtype

    def to_host(self) -> np.ndarray:
        return self.data.to_host().reshape(self.K, self.C)

    def copy_from(self, a):
        self.data.copy_from(np.asarray(a).reshape(-1))

    def copy(self):
        o = CellArray.__new__(CellArray)
        o.ctx, o.ni, o.K, o.C = self.ctx, self.ni, self.K, self.C
        o.data = self.data.copy()
        return o

    def cell(self, *I):
        """Host copy of the K values of cell I (0-based)."""
        c = int(np.ravel_multi_index(I, self.ni, order="F"))
        return self.to_host()[:, c]


class DeviceAxis:
    """An :class:`Axis` with device copies of x, dx and inv_dx."""

    def __init__(self, ctx: Context, ax: Axis):
        self.host = ax
        self.x = DeviceArray.from_host(ctx, ax.x)
        self.dx = DeviceArray.from_host(ctx, ax.dx)
        self.inv_dx = DeviceArray.from_host(ctx, ax.inv_dx)
        self.n = ax.n

    def abi(self):
        return _abi.jp_axis(self.x.ptr, self.dx.ptr, self.inv_dx.ptr, self.n)


class Particles:
    """Device particle container (``Particles{HIPBackend}``).

    Fields follow the reference: ``coords`` (D CellArrays), ``index`` (Bool
    CellArray), ``nxcell``, ``max_xcell``, ``min_xcell``, ``np``, grids ``xvi``,
    ``xci``, ``xi_vel`` (host :class:`Axis` objects; ``di``/``_di`` are their
    ``dx``/``inv_dx``).  ``seed`` and ``inject_counter`` key the Philox stream
    that replaces Julia's ``rand()``.
    """

    def __init__(self, ctx, coords, index, nxcell, max_xcell, min_xcell, xvi, xci, xi_vel,
                 seed=0):
        self.ctx = ctx
        self.coords = tuple(coords)
        self.index = index
        self.nxcell = int(nxcell)
        self.max_xcell = int(max_xcell)
        self.min_xcell = int(min_xcell)
        self.ni = tuple(a.n - 1 for a in xvi)
        self.np = self.max_xcell * int(np.prod(self.ni))
        self.xvi = tuple(xvi)
        self.xci = tuple(xci)
        self.xi_vel = tuple(tuple(g) for g in xi_vel)
        self.seed = int(seed)
        self.inject_counter = 0
        self._dxv = [DeviceAxis(ctx, a) for a in self.xvi]
        self._dvel = [[DeviceAxis(ctx, a) for a in g] for g in self.xi_vel]
        self._grid = self._make_grid()
        self._velgrid = self._make_velgrid()

    @property
    def D(self):
        return len(self.xvi)

    @property
    def C(self):
        return int(np.prod(self.ni))

    @property
    def di(self):
        return {"vertex": tuple(a.dx for a in self.xvi),
                "center": tuple(a.dx for a in self.xci),
                "velocity": tuple(tuple(a.dx for a in g) for g in self.xi_vel)}

    def _make_grid(self):
        g = _abi.jp_cellgrid()
        g.ndim = self.D
        g.max_xcell = self.max_xcell
        for d in range(self.D):
            g.ni[d] = self.ni[d]
            g.xv[d] = self._dxv[d].abi()
        return g

    def _make_velgrid(self):
        vg = _abi.jp_velgrid()
        for c in range(self.D):
            for d in range(self.D):
                vg.ax[c][d] = self._dvel[c][d].abi()
        return vg

    def grid_ref(self):
        return C.byref(self._grid)

    def velgrid_ref(self):
        return C.byref(self._velgrid)

    def coord_ptrs(self):
        return _abi.ptr_array([c.ptr for c in self.coords])

    def centre_axes_ref(self):
        """Device copies of ``xci`` (with di.center) as a ``jp_axis`` array."""
        if getattr(self, "_dxc", None) is None:
            self._dxc = [DeviceAxis(self.ctx, a) for a in self.xci]
        return (_abi.jp_axis * self.D)(*[a.abi() for a in self._dxc])

    def active(self) -> int:
        return int(self.index.to_host().sum())

    def host_state(self):
        """(coords list of flat arrays, index flat array), library layout."""
        return ([c.data.flat() for c in self.coords], self.index.data.flat())


class SplitParticles(Particles):
    """Particles in the FP32-split form (no reference counterpart).

    ``coords`` hold D float32 CellArrays of offsets u in cell units: slot
    (c, k) of a cell with first vertex x_I and spacing h sits at
    ``x_I + u*h``.  Built from an FP64 container by :func:`to_split`, converted
    back by :func:`to_f64`; the hot-path functions of :mod:`.api` accept both
    forms.  Requires uniform vertex axes and the standard velocity staggering.
    """

    form = "split"

    @property
    def off(self):
        return self.coords


Particles.form = "f64"


def _check_uniform(p: Particles):
    for d, a in enumerate(p.xvi):
        if a.dx.size and not np.all(a.dx == a.dx[0]):
            raise ValueError(f"FP32-split form needs a uniform vertex axis (axis {d} is not)")


def to_split(p: Particles) -> SplitParticles:
    """FP32-split copy of an FP64 container (same index, seed and counter)."""
    if p.form != "f64":
        raise ValueError("to_split expects an FP64 Particles container")
    _check_uniform(p)
    ctx = p.ctx
    off = [CellArray(ctx, p.ni, p.max_xcell, dtype=np.float32) for _ in range(p.D)]
    index = p.index.copy()
    s = SplitParticles(ctx, off, index, p.nxcell, p.max_xcell, p.min_xcell, p.xvi, p.xci,
                       p.xi_vel, seed=p.seed)
    s.inject_counter = p.inject_counter
    ctx.check(ctx.lib.jp_split_from_f64(ctx.h, s.grid_ref(), p.coord_ptrs(), index.ptr,
                                        s.coord_ptrs()), "to_split")
    return s


def to_f64(s: SplitParticles) -> Particles:
    """FP64 container with the absolute coordinates of a split container
    (inactive slots NaN, as everywhere else)."""
    if s.form != "split":
        raise ValueError("to_f64 expects a SplitParticles container")
    ctx = s.ctx
    coords = [CellArray(ctx, s.ni, s.max_xcell) for _ in range(s.D)]
    index = s.index.copy()
    p = Particles(ctx, coords, index, s.nxcell, s.max_xcell, s.min_xcell, s.xvi, s.xci,
                  s.xi_vel, seed=s.seed)
    p.inject_counter = s.inject_counter
    ctx.check(ctx.lib.jp_split_to_f64(ctx.h, s.grid_ref(), s.coord_ptrs(), index.ptr,
                                      p.coord_ptrs()), "to_f64")
    return p


def _split_grids(xi_vel):
    D = len(xi_vel)
    xvi = tuple(xi_vel[i][i] for i in range(D))
    if D == 2:
        xci = (xi_vel[1][0].inner(), xi_vel[0][1].inner())
    else:
        xci = (xi_vel[1][0].inner(), xi_vel[0][1].inner(), xi_vel[0][2].inner())
    return xvi, xci


def particles_from_host(ctx: Context, xi_vel, coords, index, nxcell, max_xcell, min_xcell,
                        seed=0):
    """Device Particles holding a given host state (library layout arrays),
    e.g. a restart, or the same state an oracle run starts from."""
    xi_vel = tuple(tuple(g) for g in xi_vel)
    xvi, xci = _split_grids(xi_vel)
    ni = tuple(a.n - 1 for a in xvi)
    cs = []
    for a in coords:
        ca = CellArray(ctx, ni, max_xcell)
        ca.copy_from(np.asarray(a, dtype=np.float64))
        cs.append(ca)
    idx = CellArray(ctx, ni, max_xcell, dtype=np.uint8)
    idx.copy_from(np.asarray(index, dtype=np.uint8))
    return Particles(ctx, cs, idx, nxcell, max_xcell, min_xcell, xvi, xci, xi_vel, seed=seed)


def init_particles(ctx: Context, nxcell, max_xcell, min_xcell, *xi_vel, seed: int = 0,
                   form: str = "f64"):
    """``init_particles(backend, nxcell, max_xcell, min_xcell, xi_vel...)``
    (src/Particles/particles_utils.jl:30-166): nxcell rounded up to a multiple
    of 2^D, K = max(nxcell, max_xcell), quadrant-stratified random fill.
    ``form="split"`` returns the FP32-split form of the same fill."""
    if form not in ("f64", "split"):
        raise ValueError("form must be 'f64' or 'split'")
    if form == "split":
        p = init_particles(ctx, nxcell, max_xcell, min_xcell, *xi_vel, seed=seed)
        s = to_split(p)
        for c in p.coords:
            c.data.free()
        return s
    if len(xi_vel) == 1 and isinstance(xi_vel[0], tuple) and xi_vel and \
            isinstance(xi_vel[0][0], tuple):
        xi_vel = xi_vel[0]
    if len(xi_vel) == 0:
        raise ValueError("The velocity grid cannot be empty")
    D = len(xi_vel)
    if D not in (2, 3):
        raise ValueError("The dimension of the problem must be either 2 or 3")
    xi_vel = tuple(tuple(g) for g in xi_vel)
    # xvi[i] = xi_vel[i][i]; xci = interior of the staggered axes (:56-74)
    xvi, xci = _split_grids(xi_vel)
    NQ = 4 if D == 2 else 8
    npq = math.ceil(nxcell / NQ)
    nxcell = npq * NQ
    max_xcell = max(nxcell, max_xcell)
    if max_xcell > 64:
        raise ValueError("max_xcell must be at most 64 (one wavefront of slots)")
    ni = tuple(a.n - 1 for a in xvi)
    coords = [CellArray(ctx, ni, max_xcell) for _ in range(D)]
    index = CellArray(ctx, ni, max_xcell, dtype=np.uint8)
    p = Particles(ctx, coords, index, nxcell, max_xcell, min_xcell, xvi, xci, xi_vel, seed=seed)
    ctx.check(ctx.lib.jp_init_particles(ctx.h, p.grid_ref(), p.coord_ptrs(), index.ptr, npq,
                                        int(seed)), "init_particles")
    return p


def init_cell_arrays(particles: Particles, n: int, fill=0.0):
    """``init_cell_arrays(particles, Val(n))`` (particles_utils.jl:294-301)."""
    return tuple(CellArray(particles.ctx, particles.ni, particles.max_xcell, fill=fill)
                 for _ in range(n))
