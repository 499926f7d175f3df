"""grid2particle_flip!, particle2centroid! and centroid2particle! on the GPU
against the oracle, bit for bit (runs on an MI355X)."""
import numpy as np
import pytest

import justpic_hip as jp
from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V
from test_gpu_parity import assert_field, dev_V, to_dev

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-transfers]", *a, flush=True)


def moved_case(D, kind="range"):
    """particles after one advection + move (uneven cells, NaN holes)"""
    if D == 2:
        xv, xi = grid2d(30, 26, kind=kind)
        V = stream_V(xi, 5.0 if kind == "vector" else 250.0)
        hp = host_particles(xv, xi, 24, 4, seed=8)
    else:
        xv, xi = grid3d(7)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 40, 3, seed=8)
    dt = 0.5 * min(np.min(a.dx) for a in xv) / max(np.abs(v).max() for v in V)
    O.advection(hp, V, O.JP_RK2, 0.5, dt)
    O.move_particles(hp, [])
    return xv, xi, hp


def field(axes, seed):
    rng = np.random.default_rng(seed)
    mesh = np.meshgrid(*[a.x for a in axes], indexing="ij")
    return np.asfortranarray(np.sin(3 * mesh[0] + seed) * np.cos(2 * mesh[-1]) +
                             0.01 * rng.random(mesh[0].shape))


@pytest.mark.parametrize("alpha", [0.0, 0.3, 1.0])
@pytest.mark.parametrize("D,kind,nf", [(2, "range", 1), (2, "vector", 2), (3, "range", 3)])
def test_flip_matches_oracle(ctx, D, kind, nf, alpha):
    say("flip", D, kind, nf, alpha)
    xv, xi, hp = moved_case(D, kind)
    rng = np.random.default_rng(nf)
    Fp = [np.where(hp.index == 1, rng.random(hp.S), np.nan) for _ in range(nf)]
    F = [field(xv, 1 + f) for f in range(nf)]
    F0 = [field(xv, 7 + f) for f in range(nf)]
    p = to_dev(ctx, hp)
    Fpd = []
    for a in Fp:
        ca = jp.CellArray(ctx, hp.ni, hp.K)
        ca.copy_from(a)
        Fpd.append(ca)
    Fd = [jp.DeviceArray.from_host(ctx, f) for f in F]
    F0d = [jp.DeviceArray.from_host(ctx, f) for f in F0]
    if nf == 1:
        jp.grid2particle_flip_(Fpd[0], Fd[0], F0d[0], p, alpha=alpha)
    else:
        jp.grid2particle_flip_(tuple(Fpd), tuple(Fd), tuple(F0d), p, alpha=alpha)
    O.grid2particle_flip(hp, Fp, F, F0, alpha)
    for a, b in zip(Fpd, Fp):
        assert_field(a.data.flat(), b, exact=True)


@pytest.mark.parametrize("tup", [False, True])
@pytest.mark.parametrize("D,kind", [(2, "range"), (2, "vector"), (3, "range")])
def test_particle2centroid_matches_oracle(ctx, D, kind, tup):
    say("p2c", D, kind, tup)
    xv, xi, hp = moved_case(D, kind)
    nf = 2 if tup else 1
    rng = np.random.default_rng(5)
    Fp = [np.where(hp.index == 1, rng.random(hp.S), np.nan) for _ in range(nf)]
    p = to_dev(ctx, hp)
    Fpd = []
    for a in Fp:
        ca = jp.CellArray(ctx, hp.ni, hp.K)
        ca.copy_from(a)
        Fpd.append(ca)
    Fd = [jp.centroid_field(p, fill=-3.0) for _ in range(nf)]
    jp.particle2centroid_(tuple(Fd) if tup else Fd[0], tuple(Fpd) if tup else Fpd[0], p)
    want = O.particle2centroid(hp, Fp, tuple_method=tup)
    for a, b in zip(Fd, want):
        assert_field(a.flat(), b, exact=True)


@pytest.mark.parametrize("D,kind,nf", [(2, "range", 1), (2, "vector", 2), (3, "range", 2)])
def test_centroid2particle_matches_oracle(ctx, D, kind, nf):
    say("c2p", D, kind, nf)
    xv, xi, hp = moved_case(D, kind)
    xc = O.centre_axes(hp)
    F = [field(xc, 3 + f) for f in range(nf)]
    p = to_dev(ctx, hp)
    Fpd = [jp.CellArray(ctx, hp.ni, hp.K, fill=-1.0) for _ in range(nf)]
    Fd = [jp.DeviceArray.from_host(ctx, f) for f in F]
    if nf == 1:
        jp.centroid2particle_(Fpd[0], Fd[0], p)
    else:
        jp.centroid2particle_(tuple(Fpd), tuple(Fd), p)
    Fp = [np.full(hp.S, -1.0) for _ in range(nf)]
    O.centroid2particle(hp, Fp, F)
    for a, b in zip(Fpd, Fp):
        assert np.array_equal(a.data.flat(), b), np.max(np.abs(a.data.flat() - b))
