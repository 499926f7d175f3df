"""Reverse mode of ``advection!`` + ``particle2grid!`` (no reference counterpart).

The differentiated map is, on a flat list of particles (positions ``x0``,
tracer values ``Tp``)::

    F = particle2grid(advect^nsteps(x0; V, dt), Tp)

with the reference's integrators (src/Particles/Advection/advection.jl:66-91),
its domain-exit deletion (src/Particles/move_safe.jl:99-103) and its nodal
gather (src/Interpolations/particle_to_grid.jl:37-68, 113-151).  A flat list
has no slot capacity, so particles that ``move_particles!`` would drop for a
full destination cell are not modelled.

:func:`rrule_advection_p2g` follows ChainRulesCore's ``rrule`` convention: it
returns the primal output and a pullback that maps the cotangent of ``F`` to
the cotangents of ``(V, x0, Tp)``.  The forward pass keeps the positions at
the start of every block of ``every`` steps; the pullback recomputes each
block from its checkpoint (``jp_adjoint_forward`` / ``jp_adjoint_backward``).
"""
from __future__ import annotations

import numpy as np

from . import _abi
from .api import AbstractAdvectionIntegrator, RungeKutta2
from .device import DeviceArray
from .particles import CellArray, Particles


def flatten(particles: Particles, field: CellArray | None = None):
    """Active particles of a container as a flat list: D position arrays and
    the tracer taken from ``field`` (zeros without one).  Order: groups of 64
    consecutive cells, slot-major inside a group, so that a wavefront holds 64
    different cells and a workgroup a compact patch (the pullback accumulates
    a workgroup's velocity cotangents in LDS).  Host round trip; set-up only.
    Any order gives the same model; only the speed of the pullback differs."""
    if getattr(particles, "form", "f64") != "f64":
        raise ValueError("flatten takes an FP64 container (convert with to_f64)")
    coords, index = particles.host_state()
    C, K = particles.C, particles.max_xcell
    e = np.nonzero(index.astype(bool))[0]
    c, k = e % C, e // C
    e = e[np.lexsort((c % 64, k, c // 64))]
    ctx = particles.ctx
    x = tuple(DeviceArray.from_host(ctx, np.ascontiguousarray(a[e])) for a in coords)
    T = field.data.flat()[e] if field is not None else np.zeros(e.size)
    return x, DeviceArray.from_host(ctx, np.ascontiguousarray(T, dtype=np.float64))


def _as_device(ctx, a, shape=None):
    if isinstance(a, DeviceArray):
        return a
    a = np.asarray(a, dtype=np.float64)
    return DeviceArray.from_host(ctx, a if shape is None else a.reshape(shape, order="F"))


class AdvectionP2G:
    """``F = particle2grid(advect^nsteps(x0; V), Tp)`` on the grids of
    ``particles`` (its vertex axes and staggered velocity axes).

    ``form``: "f64" runs the trajectories in FP64 (``jp_adjoint_forward``),
    "split" in the FP32-split form (``jp_adjoint_forward_split``: a cell index
    and a float32 offset per particle, the split advection's arithmetic; the
    checkpoints hold the FP64 positions of the split trajectory, from which
    the FP64 pullback recomputes each block)."""

    def __init__(self, particles: Particles, method: AbstractAdvectionIntegrator | None,
                 dt: float, nsteps: int, every: int = 10, form: str = "f64"):
        if getattr(particles, "form", "f64") != "f64":
            raise ValueError("the adjoint runs on FP64 grids (an FP64 container)")
        if form not in ("f64", "split"):
            raise ValueError("form must be 'f64' or 'split'")
        self.form = form
        if not 1 <= int(every) <= _abi.JP_ADJ_MAX_EVERY:
            raise ValueError(f"every must be in 1..{_abi.JP_ADJ_MAX_EVERY}")
        if int(nsteps) < 0:
            raise ValueError("nsteps must be >= 0")
        self.p = particles
        self.ctx = particles.ctx
        self.method = method or RungeKutta2()
        self.dt = float(dt)
        self.nsteps = int(nsteps)
        self.every = int(every)
        self.D = particles.D
        self.nodes = tuple(n + 1 for n in particles.ni)
        self._fw = None

    def forward(self, V, x0, Tp) -> DeviceArray:
        """Run the forward pass; returns the nodal field F (shape ni+1).  The
        buffers of one object are reused: the next forward pass with the same
        particle count overwrites F and the checkpoints."""
        ctx, D = self.ctx, self.D
        V = tuple(_as_device(ctx, v) for v in V)
        x0 = tuple(_as_device(ctx, x) for x in x0)
        Tp = _as_device(ctx, Tp)
        if len(V) != D or len(x0) != D:
            raise ValueError(f"need {D} velocity components and {D} position arrays")
        for c in range(D):
            want = int(np.prod([a.n for a in self.p.xi_vel[c]]))
            if V[c].size != want:
                raise ValueError(f"velocity component {c} has {V[c].size} values, its grid {want}")
        n = Tp.size
        if any(x.size != n for x in x0):
            raise ValueError("x0 and Tp sizes differ")
        nblk = max(1, -(-self.nsteps // self.every))
        if self._fw is None or self._fw["n"] != n:
            ckpt = DeviceArray(ctx, (nblk * D * max(n, 1),))
            alive = DeviceArray(ctx, (max(n, 1),), np.uint8)
            xN = tuple(DeviceArray(ctx, (max(n, 1),)) for _ in range(D))
            F = DeviceArray(ctx, self.nodes)
            W = DeviceArray(ctx, self.nodes)
        else:  # the same buffers: a new forward pass overwrites the previous one
            fw = self._fw
            ckpt, alive, xN, F, W = fw["ckpt"], fw["alive"], fw["xN"], fw["F"], fw["W"]
        entry = ctx.lib.jp_adjoint_forward if self.form == "f64" else ctx.lib.jp_adjoint_forward_split
        ctx.check(entry(
            ctx.h, self.p.grid_ref(), self.p.velgrid_ref(), _abi.ptr_array([v.ptr for v in V]),
            self.method.code, self.method.alpha, self.dt, self.nsteps, self.every, n,
            _abi.ptr_array([x.ptr for x in x0]), Tp.ptr, ckpt.ptr, alive.ptr,
            _abi.ptr_array([x.ptr for x in xN]), F.ptr, W.ptr), "adjoint forward")
        self._fw = dict(V=V, Tp=Tp, n=n, ckpt=ckpt, alive=alive, xN=xN, F=F, W=W)
        return F

    @property
    def alive(self) -> DeviceArray:
        return self._fw["alive"]

    @property
    def xN(self):
        return self._fw["xN"]

    @property
    def W(self) -> DeviceArray:
        return self._fw["W"]

    def pullback(self, Fbar, out=None):
        """Cotangents ``(Vbar, x0bar, Tpbar)`` for the cotangent ``Fbar`` of F
        (finite everywhere: give 0 at nodes the loss ignores, e.g. NaN nodes).
        Deleted particles get zero cotangents.  ``out`` (a previous result)
        is overwritten instead of allocating new arrays."""
        if self._fw is None:
            raise RuntimeError("pullback before forward")
        ctx, D, fw = self.ctx, self.D, self._fw
        Fbar = _as_device(ctx, Fbar, self.nodes)
        if Fbar.size != fw["F"].size:
            raise ValueError("Fbar must have the nodal shape ni+1")
        n = fw["n"]
        if out is None:
            Vbar = tuple(DeviceArray(ctx, v.shape) for v in fw["V"])
            x0bar = tuple(DeviceArray(ctx, (max(n, 1),)) for _ in range(D))
            Tpbar = DeviceArray(ctx, (max(n, 1),))
        else:
            Vbar, x0bar, Tpbar = out
            if (any(a.size != v.size for a, v in zip(Vbar, fw["V"]))
                    or any(x.size < n for x in x0bar) or Tpbar.size < n):
                raise ValueError("out arrays do not match this forward pass")
        ctx.check(ctx.lib.jp_adjoint_backward(
            ctx.h, self.p.grid_ref(), self.p.velgrid_ref(),
            _abi.ptr_array([v.ptr for v in fw["V"]]), self.method.code, self.method.alpha,
            self.dt, self.nsteps, self.every, n, fw["ckpt"].ptr, fw["alive"].ptr, fw["Tp"].ptr,
            fw["F"].ptr, fw["W"].ptr, Fbar.ptr, _abi.ptr_array([v.ptr for v in Vbar]),
            _abi.ptr_array([x.ptr for x in x0bar]), Tpbar.ptr), "adjoint backward")
        return Vbar, x0bar, Tpbar


def rrule_advection_p2g(particles: Particles, method, V, dt, nsteps, x0, Tp, every: int = 10,
                        form: str = "f64"):
    """ChainRulesCore-style ``rrule``: ``(F, pullback)`` with
    ``pullback(Fbar) -> (Vbar, x0bar, Tpbar)``; ``form`` as in AdvectionP2G."""
    op = AdvectionP2G(particles, method, dt, nsteps, every, form)
    F = op.forward(V, x0, Tp)
    return F, op.pullback


def l2_loss(F: DeviceArray, F_target) -> tuple[float, np.ndarray]:
    """0.5*sum((F - F_target)^2) over nodes where both are finite, and its
    cotangent (zero elsewhere, flat column-major), on the host."""
    a = F.flat()
    b = (F_target.flat() if isinstance(F_target, DeviceArray)
         else np.asarray(F_target, dtype=np.float64).ravel(order="F"))
    if a.size != b.size:
        raise ValueError("F and F_target sizes differ")
    d = a - b
    d[~np.isfinite(d)] = 0.0
    return 0.5 * float(np.sum(d * d)), d


__all__ = ["AdvectionP2G", "rrule_advection_p2g", "flatten", "l2_loss"]
