"""GPU checks of the FP32-split position form (runs on an MI355X).

The split form has no reference counterpart; it is pinned two ways:
  * where every quantity is exactly representable (dyadic grids, offsets and
    displacements) its move and injection must reproduce the FP64 oracle
    bit for bit, because the slot rules are the same and integer cell
    assignment agrees with the strict in-cell test away from ties;
  * on the benchmark flows its positions must stay within 1e-6 of the domain
    size of the FP64 oracle after many steps, with the cell-assignment
    disagreements counted by particle identity.
"""
import numpy as np
import pytest

import justpic_hip as jp
from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V

pytestmark = pytest.mark.gpu

POS_TOL = 1e-6  # of the domain size (the domains here are unit boxes)


def say(*a):
    print("[gpu-split]", *a, flush=True)


def to_dev(ctx, hp, min_xcell=1, seed=0):
    return jp.particles_from_host(ctx, hp.xi_vel, hp.coords, hp.index, hp.K, hp.K, min_xcell,
                                  seed=seed)


def dev_V(ctx, V):
    return tuple(jp.DeviceArray.from_host(ctx, v) for v in V)


def split_state(s):
    """(coords, index) of a split container, absolute FP64 coordinates."""
    p = jp.to_f64(s)
    return p.host_state()


def by_id(coords, index, ids, C):
    e = np.nonzero(index == 1)[0]
    idv = ids[e].astype(np.int64)
    o = np.argsort(idv)
    pos = np.stack([c[e] for c in coords], 1)
    return idv[o], (e % C)[o], pos[o]


def compare_by_id(a, b, C):
    """max position error over common particles, and disagreements: particles
    present in one state only plus common ones assigned to different cells."""
    ia, ca, pa = by_id(*a, C)
    ib, cb, pb = by_id(*b, C)
    common, ja, jb = np.intersect1d(ia, ib, assume_unique=True, return_indices=True)
    err = float(np.max(np.abs(pa[ja] - pb[jb]))) if common.size else 0.0
    dis = int(len(ia) + len(ib) - 2 * common.size + np.count_nonzero(ca[ja] != cb[jb]))
    return err, dis, common.size


# ------------------------------------------------------------------ conversion
@pytest.mark.parametrize("D", [2, 3])
def test_split_roundtrip(ctx, D):
    say("roundtrip", D)
    xv, xi = grid2d(40, 33) if D == 2 else grid3d(9)
    p = jp.init_particles(ctx, 4 * 2 ** D, 6 * 2 ** D, 1, *xi, seed=3)
    s = jp.to_split(p)
    assert s.form == "split" and s.off[0].dtype == np.float32
    c0, i0 = p.host_state()
    c1, i1 = split_state(s)
    assert np.array_equal(i0, i1)
    m = i0 == 1
    for d in range(D):
        u = s.off[d].data.flat()
        assert np.all((u[m] >= 0) & (u[m] < 1)), "active offsets leave [0, 1)"
        assert np.all(np.isnan(u[~m]))
        h = xv[d].x[1] - xv[d].x[0]
        assert np.max(np.abs(c1[d][m] - c0[d][m])) <= h * 2.0 ** -24 + 4e-16
        assert np.all(np.isnan(c1[d][~m]))


def test_split_rejects_nonuniform_axes(ctx):
    xv, xi = grid2d(8, kind="vector")
    p = jp.init_particles(ctx, 4, 8, 1, *xi, seed=1)
    with pytest.raises(ValueError):
        jp.to_split(p)


# ------------------------------------------------------------------ exact move / inject
def dyadic_state(D, n, K, seed):
    """Oracle state on a 1/n grid (n a power of two) whose offsets are odd multiples of 1/64 of a
    cell and whose displacements are at most one cell: every coordinate and
    every offset is exact in FP32 and FP64, and no particle sits on a face."""
    assert n & (n - 1) == 0
    xv, xi = grid2d(n) if D == 2 else grid3d(n)
    hp = O.HostParticles(xv, K, xi)
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    occ = rng.random(hp.S) < 0.6
    hp.index[:] = occ.astype(np.uint8)
    cells = np.arange(hp.S) % hp.C
    I = np.stack(np.unravel_index(cells, hp.ni, order="F"), 0)
    for d in range(D):
        u = (2 * rng.integers(0, 32, hp.S) + 1) / 64.0
        shift = rng.choice([-1, 0, 0, 0, 1], hp.S)
        x = (I[d] + u + shift) * h
        hp.coords[d][:] = np.where(occ, x, np.nan)
    return xv, xi, hp


@pytest.mark.parametrize("D", [2, 3])
def test_split_move_exact(ctx, D):
    say("move exact", D)
    xv, xi, hp = dyadic_state(D, 16 if D == 2 else 8, 10, seed=11 + D)
    T = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    s = jp.to_split(p)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    st_d = jp.move_particles_(s, (Td,), stats=True)
    st_o = O.move_particles(hp, [T])
    say("stats", st_d, st_o)
    assert st_o["moved"] > 0 and st_o["deleted"] > 0
    assert st_d == st_o
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    for d in range(D):
        assert np.array_equal(coords[d], hp.coords[d], equal_nan=True)
    assert np.array_equal(Td.data.flat(), T, equal_nan=True)


@pytest.mark.parametrize("D", [2, 3])
def test_split_inject_matches_oracle(ctx, D):
    say("inject", D)
    xv, xi, hp = dyadic_state(D, 16 if D == 2 else 8, 16, seed=21 + D)
    O.move_particles(hp, [])  # settle every particle into its cell
    T = np.where(hp.index == 1, np.random.default_rng(5).random(hp.S), np.nan)
    minx = 10 if D == 2 else 12
    s = jp.to_split(to_dev(ctx, hp, min_xcell=minx, seed=77))
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    st_d = jp.inject_particles_(s, (Td,), stats=True)
    st_o = O.inject_particles(hp, [T], minx, seed=77, counter=0)
    say("stats", st_d, st_o)
    assert st_o["injected"] > 0
    assert st_d == st_o
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    m = hp.index == 1
    h = 1.0 / hp.ni[0]
    for d in range(D):
        # new positions: the same Philox draw, rounded to an FP32 offset
        assert np.max(np.abs(coords[d][m] - hp.coords[d][m])) <= h * 2.0 ** -24 + 4e-16
    # nearest-neighbour copies pick the same source particle
    assert np.array_equal(Td.data.flat(), T, equal_nan=True)


# ------------------------------------------------------------------ p2g / g2p
@pytest.mark.parametrize("D", [2, 3])
def test_split_p2g_g2p(ctx, D):
    say("p2g/g2p", D)
    if D == 2:
        xv, xi = grid2d(24, 20)
        hp = host_particles(xv, xi, 12, 2, seed=31)
    else:
        xv, xi = grid3d(7)
        hp = host_particles(xv, xi, 16, 1, seed=31)
    rng = np.random.default_rng(9)
    Fp = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    s = jp.to_split(to_dev(ctx, hp))
    Fpd = jp.CellArray(ctx, hp.ni, hp.K)
    Fpd.copy_from(Fp)
    F = jp.nodal_field(s)
    jp.particle2grid_(F, Fpd, s)
    want = O.particle2grid(hp, Fp)
    # float32 weights: relative weight error ~1e-6 moves a weighted mean of
    # values in [0, 1) by at most ~1e-6
    np.testing.assert_allclose(F.flat(), want, rtol=0, atol=2e-6)
    Tn = np.asfortranarray(np.cos(3 * np.add.outer(xv[0].x, xv[1].x)) if D == 2 else
                           np.cos(3 * np.add.outer(np.add.outer(xv[0].x, xv[1].x), xv[2].x)))
    pT = jp.CellArray(ctx, hp.ni, hp.K, fill=0.0)
    jp.grid2particle_(pT, jp.DeviceArray.from_host(ctx, Tn), s)
    want = O.grid2particle(hp, Tn, np.zeros(hp.S))
    np.testing.assert_allclose(pT.data.flat(), want, rtol=0, atol=1e-6)


# ------------------------------------------------------------------ many steps vs FP64
LOOPS = [
    ("2d_rk2", 2, "rk2", 20),
    ("2d_rk4", 2, "rk4", 10),
    ("3d_rk4", 3, "rk4", 8),
]


@pytest.mark.parametrize("name,D,meth,steps", LOOPS)
def test_split_error_vs_fp64_oracle(ctx, name, D, meth, steps):
    """advection + move for several steps; particles tracked by an id field."""
    say("loop", name)
    if D == 2:
        xv, xi = grid2d(128)
        V = stream_V(xi, 250.0)
        hp = host_particles(xv, xi, 48, 6, seed=41)
    else:
        xv, xi = grid3d(24)
        V = cellular_V(xi, 100.0)
        hp = host_particles(xv, xi, 54, 3, seed=41)
    method, code = (jp.RungeKutta2(), O.JP_RK2) if meth == "rk2" else \
        (jp.RungeKutta4(), O.JP_RK4)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    s = jp.to_split(to_dev(ctx, hp))
    idd = jp.CellArray(ctx, hp.ni, hp.K)
    idd.copy_from(ids)
    Vd = dev_V(ctx, V)
    dt = cfl_dt(xv, V)
    for _ in range(steps):
        jp.advection_(s, method, Vd, dt)
        jp.move_particles_(s, (idd,))
        O.advection(hp, V, code, 0.5, dt)
        O.move_particles(hp, [ids])
    coords, index = split_state(s)
    err, dis, n = compare_by_id((coords, index, idd.data.flat()), (hp.coords, hp.index, ids),
                                hp.C)
    say(f"{name}: {n} particles, {steps} steps, max |x_split - x_fp64| = {err:.3e}, "
        f"cell disagreements = {dis}")
    assert err <= POS_TOL
    assert dis <= 1e-4 * n
