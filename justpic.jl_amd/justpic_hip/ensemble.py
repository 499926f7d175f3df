"""Batched ensemble of independent 2D grids (no reference counterpart).

An :class:`Ensemble` holds ``n`` FP32-split particle containers, each with its
own grid, staggered velocity and time step.  Advection of all of them is one
persistent launch (``jp_ensemble_advect``); a full time step (advection, then
per grid ``move_particles!`` with its fields and ``particle2grid!``) can be
captured once into a hipGraph and replayed.  :meth:`Ensemble.tune` times the
launch variants on the ensemble itself and keeps the fastest.

The per-grid calls of a step are independent, and most grids are too small to
fill the GPU on their own, so they are spread over a few asynchronous
sub-contexts (each its own stream and scratch; largest grid first onto the
least-loaded stream), forked after the advection with an event and joined
before the step returns.  A captured step records the same fork and join.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .api import RungeKutta2
from .device import Context, DeviceArray


class Ensemble:
    """Members are tuples ``(particles, V, dt)``: a :class:`SplitParticles`
    container, its two velocity components (DeviceArrays on the staggered
    grids) and its time step.  The arrays must outlive the ensemble."""

    def __init__(self, ctx: Context, members, method=None, streams: int = 8):
        method = method or RungeKutta2()
        self.ctx = ctx
        self.members = list(members)
        n = len(self.members)
        if n == 0:
            raise ValueError("empty ensemble")
        for p, V, dt in self.members:
            if getattr(p, "form", None) != "split" or p.D != 2:
                raise ValueError("ensemble members must be 2D FP32-split containers")
            if len(V) != 2:
                raise ValueError("two velocity components per member")
        grids = (_abi.jp_cellgrid * n)(*[p._grid for p, _, _ in self.members])
        vgs = (_abi.jp_velgrid * n)(*[p._velgrid for p, _, _ in self.members])
        off = _abi.ptr_array([o.ptr for p, _, _ in self.members for o in p.off])
        idx = _abi.ptr_array([p.index.ptr for p, _, _ in self.members])
        vel = _abi.ptr_array([v.ptr for _, V, _ in self.members for v in V])
        dts = (C.c_double * n)(*[float(dt) for _, _, dt in self.members])
        h = C.c_void_p()
        ctx.check(ctx.lib.jp_ensemble_create(ctx.h, n, grids, vgs, off, idx, vel, dts,
                                             method.code, method.alpha, C.byref(h)),
                  "ensemble")
        self.h = h
        self._keep = (grids, vgs, off, idx, vel, dts)
        # sub-contexts for the per-grid calls, members balanced by slot count
        ns = max(1, min(int(streams), n))
        self.subs = [Context(ctx.device, synchronous=False) for _ in range(ns)]
        load = [0] * ns
        self.lane = [0] * n
        for i in sorted(range(n), key=lambda i: -self.members[i][0].C * self.members[i][0].max_xcell):
            j = min(range(ns), key=lambda j: load[j])
            self.lane[i] = j
            p = self.members[i][0]
            load[j] += p.C * p.max_xcell
        self._fork = ctx.event()
        self._join = [sc.event() for sc in self.subs]

    @property
    def n(self):
        return len(self.members)

    def particles(self) -> int:
        return sum(p.active() for p, _, _ in self.members)

    def advect(self):
        """One persistent launch advecting every member (the tuned variant)."""
        self.ctx.check(self.ctx.lib.jp_ensemble_advect(self.ctx.h, self.h), "ensemble advect")

    def set_variant(self, v: int):
        self.ctx.check(self.ctx.lib.jp_ensemble_set_variant(self.ctx.h, self.h, int(v)),
                       "ensemble variant")

    def tune(self, reps: int = 3):
        """Time every launch variant (positions unchanged) and keep the
        fastest.  Returns (best variant, per-variant milliseconds)."""
        best = C.c_int()
        ms = (C.c_float * _abi.JP_ENS_VARIANTS)()
        self.ctx.check(self.ctx.lib.jp_ensemble_tune(self.ctx.h, self.h, int(reps),
                                                     C.byref(best), ms), "ensemble tune")
        return best.value, list(ms)

    def step(self, fields, nodal):
        """advection! for all members, then per member move_particles!(p,
        (field,)) and particle2grid!(F, field, p), the members' calls on the
        sub-context streams."""
        self._step(fields, nodal, sync=self.ctx.synchronous)

    def _step(self, fields, nodal, sync):
        ctx, lib = self.ctx, self.ctx.lib
        self.advect()
        ctx.record(self._fork)
        for sc in self.subs:
            sc.wait(self._fork)
        for i, ((p, _, _), f, F) in enumerate(zip(self.members, fields, nodal)):
            sc = self.subs[self.lane[i]]
            sc.check(lib.jp_move_particles_split(sc.h, p.grid_ref(), p.coord_ptrs(), p.index.ptr,
                                                 1, _abi.ptr_array([f.ptr]), None),
                     "move_particles!")
            sc.check(lib.jp_particle2grid_split(sc.h, p.grid_ref(), F.ptr, f.ptr, p.coord_ptrs(),
                                                p.index.ptr), "particle2grid!")
        for sc, ev in zip(self.subs, self._join):
            sc.record(ev)
            ctx.wait(ev)
        if sync:
            ctx.sync()

    def capture_step(self, fields, nodal):
        """hipGraph of one :meth:`step` (replay with ``ctx.graph_launch``).
        Run one eager step first so that the context's scratch is sized."""
        self.ctx.graph_begin()
        try:
            self._step(fields, nodal, sync=False)
        finally:
            g = self.ctx.graph_end()
        return g

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.ctx.lib.jp_ensemble_destroy(self.ctx.h, self.h)
            self.h = C.c_void_p()
        for sc, ev in zip(getattr(self, "subs", ()), getattr(self, "_join", ())):
            sc.sync()
            sc.lib.jp_event_destroy(sc.h, ev)
            sc.close()
        if getattr(self, "_fork", None) is not None:
            self.ctx.lib.jp_event_destroy(self.ctx.h, self._fork)
            self._fork = None
        self.subs, self._join = [], []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_ensemble(ctx: Context, n: int, seed: int = 0,
                    sizes=(64, 128, 192, 256, 384, 512), amp=(50.0, 400.0),
                    nxcell=24, max_xcell=48, min_xcell=12, streams: int = 8):
    """The ensemble workload: per member a cell count drawn from ``sizes``
    (squared), the README stream function at an amplitude drawn uniform in
    ``amp``, its own CFL time step, seeded positions in the FP32-split form.
    Returns (ensemble, fields, nodal fields)."""
    from .configs import gaussian_T, stream2d
    from .particles import CellArray, init_particles
    from .api import grid2particle_, nodal_field
    rng = np.random.default_rng(seed)
    members, fields, nodal = [], [], []
    for i in range(n):
        m = int(rng.choice(sizes))
        a = float(rng.uniform(*amp))
        cfg = stream2d(m, amp=a, readme=False)
        p = init_particles(ctx, nxcell, max_xcell, min_xcell, *cfg["xi_vel"],
                           seed=seed * 1000 + i, form="split")
        V = tuple(DeviceArray.from_host(ctx, v) for v in cfg["V"])
        T = DeviceArray.from_host(ctx, gaussian_T(cfg["xv"]))
        f = CellArray(ctx, p.ni, p.max_xcell, fill=float("nan"))
        grid2particle_(f, T, p)
        members.append((p, V, cfg["dt"]))
        fields.append(f)
        nodal.append(nodal_field(p))
    return Ensemble(ctx, members, streams=streams), fields, nodal
