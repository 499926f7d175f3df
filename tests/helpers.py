"""Shared builders for the tests: small grids and host particle states."""
import numpy as np

from justpic_hip.grids import Axis, staggered_grid
from oracle import oracle as O


def grid2d(n=8, m=None, L=1.0, kind="range"):
    m = m or n
    if kind == "range":
        xv = (Axis.lin(0.0, L, n + 1), Axis.lin(0.0, L, m + 1))
    else:
        rng = np.random.default_rng(n * 100 + m)
        xs = np.concatenate([[0.0], np.sort(rng.uniform(0.05, 0.95, n - 1)), [1.0]])
        ys = np.concatenate([[0.0], np.sort(rng.uniform(0.05, 0.95, m - 1)), [1.0]])
        xv = (Axis(xs, "vector"), Axis(ys, "vector"))
    return xv, staggered_grid(xv)


def grid3d(n=6, L=1.0):
    xv = tuple(Axis.lin(0.0, L, n + 1) for _ in range(3))
    return xv, staggered_grid(xv)


def sample(f, axes):
    mesh = np.meshgrid(*[a.x for a in axes], indexing="ij")
    return np.asfortranarray(f(*mesh))


def stream_V(xi_vel, amp=250.0):
    Vx = sample(lambda x, y: amp * np.sin(np.pi * x) * np.cos(np.pi * y), xi_vel[0])
    Vy = sample(lambda x, y: -amp * np.cos(np.pi * x) * np.sin(np.pi * y), xi_vel[1])
    return Vx, Vy


def cellular_V(xi_vel, amp=100.0):
    f = [
        lambda x, y, z: amp * np.sin(np.pi * x) * np.cos(np.pi * y) * np.cos(np.pi * z),
        lambda x, y, z: amp * np.cos(np.pi * x) * np.sin(np.pi * y) * np.cos(np.pi * z),
        lambda x, y, z: -2 * amp * np.cos(np.pi * x) * np.cos(np.pi * y) * np.sin(np.pi * z),
    ]
    return tuple(sample(f[c], xi_vel[c]) for c in range(3))


def host_particles(xv, xi_vel, K, npq, seed=1):
    hp = O.HostParticles(xv, K, xi_vel)
    O.init_particles(hp, npq, seed)
    return hp


def cfl_dt(xv, V):
    return min((xv[d].x[1] - xv[d].x[0]) / np.max(np.abs(V[d])) for d in range(len(xv)))


def place(hp, cell, slot, p, args=None, vals=None):
    """Put one particle at (cell tuple, slot) with coordinates p."""
    c = int(np.ravel_multi_index(cell, hp.ni, order="F"))
    e = c + hp.C * slot
    for d in range(hp.D):
        hp.coords[d][e] = p[d]
    hp.index[e] = 1
    if args is not None:
        for a, v in zip(args, vals):
            a[e] = v
    return e


def clear(hp):
    for d in range(hp.D):
        hp.coords[d][:] = np.nan
    hp.index[:] = 0


def elem(hp, cell, slot):
    return int(np.ravel_multi_index(cell, hp.ni, order="F")) + hp.C * slot


# ---- adjoint cases -------------------------------------------------------

def flat_active(hp):
    """Active particles as a flat list (slot-major storage order) and the mask."""
    m = hp.index.astype(bool)
    return [c[m].copy() for c in hp.coords], m


def blob(x0):
    return np.exp(-((x0[0] - 0.5) ** 2 + (x0[1] - 0.7) ** 2) / 0.02)


def adjoint_case(D, outflow=False, npq=2):
    """Small adjoint workload: host particles, velocity, time step.  With
    `outflow` a uniform velocity carries part of the particles out."""
    if D == 2:
        xv, xi = grid2d(12, 10)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 32, npq, seed=3)
    else:
        xv, xi = grid3d(5)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 40, npq, seed=3)
    if outflow:
        u = (3.0, 1.0, -0.5)
        V = tuple(sample(lambda *X, c=c: np.full_like(X[0], u[c]), xi[c]) for c in range(D))
    dt = cfl_dt(xv, V) * (0.2 if outflow else 0.7)
    return hp, V, dt
