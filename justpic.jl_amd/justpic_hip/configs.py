"""Synthetic workloads of the reference's examples, built on the host.

* ``stream2d``: the README rotating cell (README.md:30-79): vertices
  ``range(0, 1, n+1)``, staggered Vx on (xv, expanded yc), Vy on (expanded xc,
  yv), ``vx = A sin(pi x) cos(pi y)``, ``vy = -A cos(pi x) sin(pi y)``,
  ``dt = min(dx/max|Vx|, dy/max|Vy|)``.
* ``gaussian_T``: ``T = exp(-((x-0.5)^2 + (y-0.75)^2)/0.01)`` on the vertices.
* ``cellular3d``: ``(sin pi x cos pi y cos pi z, cos pi x sin pi y cos pi z,
  -2 cos pi x cos pi y sin pi z) * A`` on the 3D staggered grids
  (scripts/temperature_advection3D.jl layout).
* ``inverse2d``: the adjoint inverse problem: the rotating cell as the initial
  guess, a "true" velocity with a seeded smooth divergence-free perturbation
  of 10% of the amplitude, a Gaussian tracer (sigma 0.1 at (0.5, 0.7)).
"""
from __future__ import annotations

import numpy as np

from .grids import Axis, staggered_grid


def vertex_axes(ncells, lengths=None, readme=False):
    lengths = lengths or (1.0,) * len(ncells)
    mk = Axis.range if readme else Axis.lin
    return tuple(mk(0.0, L, n + 1) for n, L in zip(ncells, lengths))


def sample(f, axes):
    """f evaluated on the tensor grid of `axes`, Julia column-major shape."""
    mesh = np.meshgrid(*[a.x for a in axes], indexing="ij")
    return np.asfortranarray(f(*mesh))


def stream2d(n=256, amp=250.0, readme=True):
    xv = vertex_axes((n, n), readme=readme)
    xi_vel = staggered_grid(xv, readme=readme)
    Vx = sample(lambda x, y: amp * np.sin(np.pi * x) * np.cos(np.pi * y), xi_vel[0])
    Vy = sample(lambda x, y: -amp * np.cos(np.pi * x) * np.sin(np.pi * y), xi_vel[1])
    dx = xv[0].x[1] - xv[0].x[0]
    dy = xv[1].x[1] - xv[1].x[0]
    dt = min(dx / np.max(np.abs(Vx)), dy / np.max(np.abs(Vy)))
    return {"xv": xv, "xi_vel": xi_vel, "V": (Vx, Vy), "dt": float(dt)}


def smooth_perturbation(xi_vel, amp, rel=0.1, seed=0, modes=3):
    """Divergence-free, boundary-tangential velocity perturbation: modes
    (m, n) <= `modes` with vx = a n sin(m pi x) cos(n pi y), vy = -a m cos(m pi
    x) sin(n pi y), seeded coefficients, scaled to max|dv| = rel * amp."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((modes, modes))
    def vx(x, y):
        return sum(a[m - 1, n - 1] * n * np.sin(m * np.pi * x) * np.cos(n * np.pi * y)
                   for m in range(1, modes + 1) for n in range(1, modes + 1))
    def vy(x, y):
        return sum(-a[m - 1, n - 1] * m * np.cos(m * np.pi * x) * np.sin(n * np.pi * y)
                   for m in range(1, modes + 1) for n in range(1, modes + 1))
    dVx, dVy = sample(vx, xi_vel[0]), sample(vy, xi_vel[1])
    s = rel * amp / max(np.max(np.abs(dVx)), np.max(np.abs(dVy)))
    return np.asfortranarray(dVx * s), np.asfortranarray(dVy * s)


def inverse2d(n=256, amp=250.0, rel=0.1, seed=0):
    """Guess (the rotating cell), true velocity (guess + smooth 10%
    perturbation), CFL time step of the true field, tracer T(x, y)."""
    cfg = stream2d(n, amp=amp, readme=False)
    dVx, dVy = smooth_perturbation(cfg["xi_vel"], amp, rel, seed)
    Vt = (np.asfortranarray(cfg["V"][0] + dVx), np.asfortranarray(cfg["V"][1] + dVy))
    dx = cfg["xv"][0].x[1] - cfg["xv"][0].x[0]
    dy = cfg["xv"][1].x[1] - cfg["xv"][1].x[0]
    dt = min(dx / np.max(np.abs(Vt[0])), dy / np.max(np.abs(Vt[1])))
    cfg.update(V_true=Vt, dt=float(dt),
               T=lambda x, y: np.exp(-((x - 0.5) ** 2 + (y - 0.7) ** 2) / (2 * 0.1 ** 2)))
    return cfg


def gaussian_T(xv, x0=0.5, y0=0.75, w=0.01):
    return sample(lambda x, y: np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / w), xv)


def cellular3d(n=128, amp=100.0):
    xv = vertex_axes((n, n, n))
    xi_vel = staggered_grid(xv)
    f = [
        lambda x, y, z: amp * np.sin(np.pi * x) * np.cos(np.pi * y) * np.cos(np.pi * z),
        lambda x, y, z: amp * np.cos(np.pi * x) * np.sin(np.pi * y) * np.cos(np.pi * z),
        lambda x, y, z: -2 * amp * np.cos(np.pi * x) * np.cos(np.pi * y) * np.sin(np.pi * z),
    ]
    V = tuple(sample(f[c], xi_vel[c]) for c in range(3))
    dts = [(xv[d].x[1] - xv[d].x[0]) / np.max(np.abs(V[d])) for d in range(3)]
    return {"xv": xv, "xi_vel": xi_vel, "V": V, "dt": float(min(dts))}


def uniform_field(shape_cells_K, seed=1):
    rng = np.random.default_rng(seed)
    return rng.random(shape_cells_K)
