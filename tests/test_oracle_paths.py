"""Pins of the oracle's hot-path functions by closed forms, invariants, the
reference's own integration tests and hand-built cases (no GPU)."""
import math

import numpy as np
import pytest

from justpic_hip.grids import Axis, staggered_grid
from oracle import oracle as O

from helpers import (cellular_V, cfl_dt, clear, elem, grid2d, grid3d, host_particles, place,
                     sample, stream_V)


# ---------------------------------------------------------------- init
def test_init_doctest_count():
    """docs/src/index.md:29-46: nxcell=6 (rounded to 8), K=12 on 7x7 cells -> 392."""
    xv = (Axis.lin(0, 1, 8), Axis.lin(0, 1, 8))
    hp = O.HostParticles(xv, 12, staggered_grid(xv))
    O.init_particles(hp, math.ceil(6 / 4), seed=3)
    assert int(hp.index.sum()) == 392


@pytest.mark.parametrize("D", [2, 3])
def test_init_quadrant_stratified(D):
    xv, xi = grid2d(5) if D == 2 else grid3d(4)
    npq, K = 3, 3 * 2 ** D + 5
    hp = host_particles(xv, xi, K, npq, seed=11)
    nq = 2 ** D
    for c in range(hp.C):
        I = np.unravel_index(c, hp.ni, order="F")
        for l in range(K):
            e = c + hp.C * l
            if l >= npq * nq:
                assert hp.index[e] == 0 and all(np.isnan(hp.coords[d][e]) for d in range(D))
                continue
            q = l // npq
            for d in range(D):
                lo = xv[d].x[I[d]] + ((q >> d) & 1) * xv[d].dx[I[d]] / 2
                assert lo <= hp.coords[d][e] <= lo + xv[d].dx[I[d]] / 2


# ---------------------------------------------------------------- advection
def _linear_setup(D, a, n=8):
    """v_c = a_c * (x_c - 1/2) on every staggered grid: bilinear interpolation
    is exact and each component is an independent linear ODE y' = a y."""
    xv, xi = grid2d(n) if D == 2 else grid3d(n)
    V = []
    for c in range(D):
        V.append(sample(lambda *X, c=c: a[c] * (X[c] - 0.5), xi[c]))
    return xv, xi, V


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("method,alpha,poly", [
    (O.JP_EULER, 0.5, lambda h: 1 + h),
    (O.JP_RK2, 0.5, lambda h: 1 + h + h * h / 2),
    (O.JP_RK2, 2 / 3, lambda h: 1 + h + h * h / 2),
    (O.JP_RK4, 0.5, lambda h: 1 + h + h ** 2 / 2 + h ** 3 / 6 + h ** 4 / 24),
])
def test_advection_linear_field_closed_form(D, method, alpha, poly):
    """For y' = a y every RK method reproduces its stability polynomial:
    y1 = y0 * R(a dt) (a wrong stage weight or a dropped stage fails).  All
    particles stay inside the velocity grids (|displacement| < 0.02)."""
    a = (-0.8, -0.6, -0.4)[:D]  # contracting: nothing leaves the grids
    xv, xi, V = _linear_setup(D, a)
    hp = host_particles(xv, xi, 8 * 2 ** D, 2, seed=5)
    x0 = [c.copy() for c in hp.coords]
    dt = 0.05
    O.advection(hp, V, method, alpha, dt)
    act = hp.index.astype(bool)
    for d in range(D):
        want = 0.5 + (x0[d][act] - 0.5) * poly(a[d] * dt)
        assert np.allclose(hp.coords[d][act], want, rtol=1e-13, atol=1e-14)
        assert np.all(np.isnan(hp.coords[d][~act]))


def test_advection_outside_velocity_grid_is_inf():
    """interp_velocity2particle returns +Inf outside the component grid
    (src/Particles/Advection/advection.jl:104-108); the particle then leaves."""
    xv, xi = grid2d(4)
    V = stream_V(xi)
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    e = place(hp, (3, 1), 0, (1.0 + 1e-9, 0.3))  # beyond Vx's x-axis
    O.advection(hp, V, O.JP_EULER, 0.5, 1e-3)
    assert np.isinf(hp.coords[0][e]) or np.isnan(hp.coords[0][e])


# ---------------------------------------------------------------- g2p / p2g
@pytest.mark.parametrize("D", [2, 3])
def test_g2p_reproduces_linear_field(D):
    """test/test_interpolation_kernels.jl:92-94,159-161: T = last coordinate."""
    xv, xi = grid2d(4) if D == 2 else grid3d(4)
    hp = host_particles(xv, xi, 3 * 2 ** D, 1, seed=2)
    T = sample(lambda *X: X[D - 1], xv)
    pT = O.grid2particle(hp, T)
    act = hp.index.astype(bool)
    assert np.allclose(pT[act], hp.coords[D - 1][act], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("D", [2, 3])
def test_p2g_reference_error_bound(D):
    """test/test_interpolation_kernels.jl:101-105,168-171: g2p then p2g of a
    linear field, norm(T2 - T)/length(T) < 0.1."""
    xv, xi = grid2d(4) if D == 2 else grid3d(4)
    hp = host_particles(xv, xi, 12 if D == 3 else 8, 1 if D == 3 else 2, seed=4)
    T = sample(lambda *X: X[D - 1], xv)
    pT = O.grid2particle(hp, T)
    T2 = O.particle2grid(hp, pT)
    assert np.linalg.norm(T2 - T.ravel(order="F")) / T2.size < 0.1


def test_p2g_inverse_distance_square_weights():
    """One particle per cell on a 2x2 grid: every node is the 1/d^2-weighted
    mean of the particles of the cells touching it; corners copy their cell."""
    xv, xi = grid2d(2)
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    pts = {(0, 0): (0.2, 0.1), (1, 0): (0.7, 0.3), (0, 1): (0.4, 0.9), (1, 1): (0.55, 0.6)}
    vals = {(0, 0): 1.0, (1, 0): -2.0, (0, 1): 5.0, (1, 1): 3.5}
    Fp = np.full(hp.S, np.nan)
    for cell, p in pts.items():
        e = place(hp, cell, 1, p)
        Fp[e] = vals[cell]
    F = O.particle2grid(hp, Fp).reshape(3, 3, order="F")
    for i in range(3):
        for j in range(3):
            node = np.array([xv[0].x[i], xv[1].x[j]])
            cells = [(a, b) for a in (i - 1, i) for b in (j - 1, j) if 0 <= a < 2 and 0 <= b < 2]
            w = np.array([1 / np.sum((node - np.array(pts[c])) ** 2) for c in cells])
            f = np.array([vals[c] for c in cells])
            assert F[i, j] == pytest.approx(np.sum(w * f) / np.sum(w), rel=1e-14)
    assert F[0, 0] == pytest.approx(1.0, rel=2e-16) and F[2, 2] == pytest.approx(3.5, rel=2e-16)


def test_p2g_empty_node_is_nan_and_constant_field():
    xv, xi = grid2d(3)
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    e = place(hp, (0, 0), 2, (0.1, 0.2))
    Fp = np.zeros(hp.S)
    Fp[e] = 4.25
    F = O.particle2grid(hp, Fp).reshape(4, 4, order="F")
    assert F[0, 0] == 4.25 and F[1, 1] == 4.25
    assert np.isnan(F[3, 3]) and np.isnan(F[2, 2])


@pytest.mark.parametrize("D", [2, 3])
def test_p2g_multi_matches_scalar_within_ulps(D):
    """Tuple p2g multiplies by inv(w); scalar 2D divides: same up to 1 ulp."""
    xv, xi = grid2d(5) if D == 2 else grid3d(3)
    hp = host_particles(xv, xi, 2 ** D * 3, 2, seed=8)
    rng = np.random.default_rng(1)
    f1, f2 = rng.random(hp.S), rng.random(hp.S)
    s = O.particle2grid(hp, f1)
    m1, m2 = O.particle2grid_multi(hp, [f1, f2])
    assert np.allclose(s, m1, rtol=4e-16, atol=0)
    assert np.allclose(O.particle2grid(hp, f2), m2, rtol=4e-16, atol=0)


# ---------------------------------------------------------------- move
def test_move_to_neighbour_first_free_slot():
    xv, xi = grid2d(4)
    hp = O.HostParticles(xv, 6, xi)
    clear(hp)
    arg = np.full(hp.S, np.nan)
    place(hp, (1, 1), 0, (0.30, 0.30), [arg], [1.0])   # stays
    place(hp, (2, 1), 0, (0.60, 0.30), [arg], [2.0])   # occupant of slot 0 in (2,1)
    place(hp, (1, 1), 3, (0.55, 0.30), [arg], [3.0])   # moves right into (2,1)
    st = O.move_particles(hp, [arg])
    assert st["moved"] == 1 and st["deleted"] == 0 and st["dropped"] == 0
    e = elem(hp, (2, 1), 1)
    assert hp.index[e] == 1 and hp.coords[0][e] == 0.55 and arg[e] == 3.0
    src = elem(hp, (1, 1), 3)
    assert hp.index[src] == 0 and np.isnan(hp.coords[0][src]) and np.isnan(arg[src])


def test_move_starting_point_carries_across_destinations():
    """starting_point is per source cell and shared between destinations
    (src/Particles/move_safe.jl:83,115-117): after inserting at slot 3 of one
    neighbour, the next mover goes to slot >= 3 of another neighbour."""
    xv, xi = grid2d(4)
    hp = O.HostParticles(xv, 6, xi)
    clear(hp)
    for s in range(3):  # right neighbour (2,1): slots 0..2 busy
        place(hp, (2, 1), s, (0.60 + 0.01 * s, 0.30))
    place(hp, (1, 1), 0, (0.55, 0.30))  # -> (2,1) slot 3
    place(hp, (1, 1), 1, (0.30, 0.55))  # -> (1,2), empty, but lands in slot 3
    O.move_particles(hp)
    assert hp.index[elem(hp, (2, 1), 3)] == 1
    assert hp.index[elem(hp, (1, 2), 3)] == 1
    assert hp.index[elem(hp, (1, 2), 0)] == 0


def test_move_full_destination_drops_and_domain_exit_deletes():
    xv, xi = grid2d(3)
    K = 3
    hp = O.HostParticles(xv, K, xi)
    clear(hp)
    for s in range(K):
        place(hp, (2, 1), s, (0.7 + 0.01 * s, 0.5))
    place(hp, (1, 1), 0, (0.68, 0.5))   # -> full (2,1): dropped
    place(hp, (0, 0), 0, (-0.01, 0.1))  # out of the domain: deleted
    place(hp, (0, 2), 0, (0.1, 1.0))    # on the domain boundary (strict): deleted
    st = O.move_particles(hp)
    assert st == {"moved": 0, "deleted": 2, "dropped": 1, "far": 0, "injected": 0, "orphan": 0}
    assert int(hp.index.sum()) == K


def test_move_left_vertex_tie_reinserts_in_same_cell():
    """p == xv[i] fails the strict isincell test but bisection seeded at i
    returns i: the particle is re-slotted into its own cell."""
    xv, xi = grid2d(4)
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    place(hp, (2, 2), 2, (0.5, 0.6))  # x exactly on the left vertex line
    st = O.move_particles(hp)
    assert st["moved"] == 1
    assert hp.index[elem(hp, (2, 2), 0)] == 1 and hp.coords[0][elem(hp, (2, 2), 0)] == 0.5


@pytest.mark.parametrize("D", [2, 3])
def test_move_invariants_random_flow(D):
    """Conservation (active_after = active_before - deleted - dropped), every
    active particle inside its cell's closed bracket, ids carried with coords."""
    if D == 2:
        xv, xi = grid2d(16)
        V = stream_V(xi)
        K, npq = 12, 2
    else:
        xv, xi = grid3d(6)
        V = cellular_V(xi)
        K, npq = 16, 1
    hp = host_particles(xv, xi, K, npq, seed=9)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    dt = 0.9 * cfl_dt(xv, V)
    for _ in range(3):
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        pos = {int(ids[e]): tuple(hp.coords[d][e] for d in range(D))
               for e in np.flatnonzero(hp.index)}
        before = int(hp.index.sum())
        st = O.move_particles(hp, [ids])
        assert int(hp.index.sum()) == before - st["deleted"] - st["dropped"]
        for e in np.flatnonzero(hp.index):
            c, k = e % hp.C, e // hp.C
            I = np.unravel_index(c, hp.ni, order="F")
            p = tuple(hp.coords[d][e] for d in range(D))
            assert pos[int(ids[e])] == p
            for d in range(D):
                assert xv[d].x[I[d]] <= p[d] <= xv[d].x[I[d] + 1]
        assert np.all(np.isnan(ids[hp.index == 0]))


# ---------------------------------------------------------------- injection
def test_inject_fills_deficient_first_quadrant_and_copies_nearest():
    xv, xi = grid2d(4)
    hp = O.HostParticles(xv, 8, xi)
    clear(hp)
    T = np.full(hp.S, np.nan)
    # cell (1,1) is empty; the only neighbours sit in (0,1) and (2,2)
    place(hp, (0, 1), 0, (0.24, 0.26), [T], [7.0])
    place(hp, (2, 2), 0, (0.74, 0.74), [T], [9.0])
    for c in [(a, b) for a in range(4) for b in range(4)]:
        if c in ((0, 1), (2, 2), (1, 1)):
            continue
        for s in range(4):
            place(hp, c, 4 + s, (xv[0].x[c[0]] + 0.03 + 0.12 * (s & 1),
                                 xv[1].x[c[1]] + 0.03 + 0.12 * (s >> 1)), [T], [0.0])
    st = O.inject_particles(hp, [T], min_xcell=8, seed=1, counter=0)
    # min per quadrant = cld(8, 4) = 2; cell (1,1) gets 2 per quadrant
    got = [e for e in (elem(hp, (1, 1), s) for s in range(8)) if hp.index[e]]
    assert len(got) == 8
    for s, e in enumerate(got):
        q = s // 2
        for d in range(2):
            lo = xv[d].x[1] + 0.125 * ((q >> d) & 1)
            assert lo < hp.coords[d][e] < lo + 0.125
    # the first injected particle (quadrant c11, next to (0,1)) copies 7.0
    assert T[got[0]] == 7.0
    assert st["injected"] >= 8


def test_inject_breaks_on_first_sufficient_quadrant():
    """src/Particles/injection.jl:100 breaks out of the quadrant loop: a full
    first quadrant stops injection even if later quadrants are empty."""
    xv, xi = grid2d(2)
    hp = O.HostParticles(xv, 8, xi)
    clear(hp)
    for s in range(3):
        place(hp, (0, 0), s, (0.05 + 0.1 * s, 0.1))  # quadrant c11 of cell (0,0)
    for c in [(1, 0), (0, 1), (1, 1)]:  # two particles, both in quadrant c11
        for q in range(2):
            place(hp, c, q, (xv[0].x[c[0]] + 0.1 + 0.1 * q, xv[1].x[c[1]] + 0.1))
    st = O.inject_particles(hp, [], min_xcell=8, seed=1, counter=0)
    assert int(hp.index[[elem(hp, (0, 0), s) for s in range(8)]].sum()) == 3
    assert st["injected"] == 0


def test_inject_orphan_when_no_neighbour():
    """An empty grid: the very first injected particle has no neighbour (the
    reference reads out of bounds there); its field keeps its old value and
    every later particle copies a neighbour."""
    xv, xi = grid2d(2)
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    T = np.full(hp.S, -1.0)
    st = O.inject_particles(hp, [T], min_xcell=4, seed=2, counter=0)
    assert st["injected"] == 16 and st["orphan"] == 1
    assert np.all(T == -1.0) and int(hp.index.sum()) == 16


def test_reference_miniapp_conservation_2d():
    """test/test_2D.jl:482-538 (advection_test_2D): 64^2 cells, nxcell 25->28,
    K 50, min 10, RK2(2/3), 25 iterations of p2g/advect/move/inject/g2p;
    |sum T - sum T0| / sum T0 < 1e-2."""
    xv = (Axis.lin(0, 1, 65), Axis.lin(0, 1, 65))
    xi = staggered_grid(xv)
    V = stream_V(xi)
    hp = host_particles(xv, xi, 50, 7, seed=21)
    T = sample(lambda x, y: y, xv)
    dt = cfl_dt(xv, V) / 2
    pT = O.grid2particle(hp, T)
    sumT = T.sum()
    Tf = T.ravel(order="F")
    for it in range(25):
        Tf = O.particle2grid(hp, pT)
        O.advection(hp, V, O.JP_RK2, 2 / 3, dt)
        O.move_particles(hp, [pT])
        O.inject_particles(hp, [pT], 10, seed=21, counter=it)
        pT = O.grid2particle(hp, Tf, pT)
    assert abs(sumT - Tf.sum()) / sumT < 1e-2
