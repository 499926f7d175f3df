"""GPU parity of the FP64 drop-in against the oracle (runs on an MI355X).

Bar: masks and cell indices bit-exact; positions and fields within 1e-12
relative (the kernels are written to be bit-identical; equality is asserted
separately where it holds by construction).
"""
import numpy as np
import pytest

import justpic_hip as jp
from oracle import oracle as O

from helpers import (cellular_V, cfl_dt, clear, elem, grid2d, grid3d, host_particles, place,
                     stream_V)

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def say(*a):
    print("[gpu-parity]", *a, flush=True)


def to_dev(ctx, hp, nxcell=None, min_xcell=1, seed=0):
    return jp.particles_from_host(ctx, hp.xi_vel, hp.coords, hp.index, nxcell or hp.K, hp.K,
                                  min_xcell, seed=seed)


def dev_V(ctx, V):
    return tuple(jp.DeviceArray.from_host(ctx, v) for v in V)


def assert_state_equal(p, hp, exact=True):
    coords, index = p.host_state()
    assert np.array_equal(index, hp.index), "mask mismatch"
    for d in range(hp.D):
        a, b = coords[d], hp.coords[d]
        assert np.array_equal(np.isnan(a), np.isnan(b)), f"NaN pattern mismatch dim {d}"
        m = ~np.isnan(b)
        if exact:
            assert np.array_equal(a[m], b[m]), \
                f"dim {d}: max |diff| {np.max(np.abs(a[m] - b[m])) if m.any() else 0}"
        else:
            np.testing.assert_allclose(a[m], b[m], rtol=RTOL, atol=0)


def assert_field(a, b, exact=True):
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    if exact:
        assert np.array_equal(a[m], b[m]), f"max |diff| {np.max(np.abs(a[m] - b[m]))}"
    else:
        np.testing.assert_allclose(a[m], b[m], rtol=RTOL, atol=0)


# ------------------------------------------------------------------ init
@pytest.mark.parametrize("D", [2, 3])
def test_init_bit_exact(ctx, D):
    say("init", D)
    xv, xi = grid2d(9, 7) if D == 2 else grid3d(5)
    npq, K = 3, 3 * 2 ** D + 4
    p = jp.init_particles(ctx, npq * 2 ** D, K, 1, *xi, seed=42)
    hp = host_particles(xv, xi, K, npq, seed=42)
    assert_state_equal(p, hp)


# ------------------------------------------------------------------ advection
CASES = [
    ("euler", jp.Euler(), O.JP_EULER, 0.5),
    ("rk2", jp.RungeKutta2(), O.JP_RK2, 0.5),
    ("rk2_23", jp.RungeKutta2(2 / 3), O.JP_RK2, 2 / 3),
    ("rk4", jp.RungeKutta4(), O.JP_RK4, 0.5),
]


@pytest.mark.parametrize("name,method,code,alpha", CASES)
@pytest.mark.parametrize("D", [2, 3])
def test_advection_parity(ctx, D, name, method, code, alpha):
    say("advection", D, name)
    if D == 2:
        xv, xi = grid2d(24, 20)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 24, 4, seed=3)
    else:
        xv, xi = grid3d(8)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 27, 2, seed=3)
    dt = cfl_dt(xv, V)
    p = to_dev(ctx, hp)
    Vd = dev_V(ctx, V)
    for _ in range(2):
        jp.advection_(p, method, Vd, dt)
        O.advection(hp, V, code, alpha, dt)
    assert_state_equal(p, hp, exact=True)


def test_advection_vector_grid_parity(ctx):
    """Non-uniform (vector) axes: diff(x) spacing, bisection over refined axes."""
    say("advection vector grid")
    xv, xi = grid2d(12, 10, kind="vector")
    V = stream_V(xi, amp=5.0)
    hp = host_particles(xv, xi, 16, 2, seed=5)
    p = to_dev(ctx, hp)
    dt = 0.5 * min(np.min(xv[0].dx), np.min(xv[1].dx)) / 5.0
    jp.advection_(p, jp.RungeKutta2(), dev_V(ctx, V), dt)
    O.advection(hp, V, O.JP_RK2, 0.5, dt)
    assert_state_equal(p, hp, exact=True)


def test_advection_gridline_ties(ctx):
    """Particles exactly on staggered-grid lines take the reference bisection
    path (ties resolved by the search path, not by rounding)."""
    say("advection ties")
    xv, xi = grid2d(8)
    V = stream_V(xi)
    hp = O.HostParticles(xv, 8, xi)
    clear(hp)
    ys = xi[0][1].x  # Vx's y-axis (cell centers + ghosts)
    s = 0
    for i in range(1, 7):
        for j in range(1, 7):
            place(hp, (i, j), s % 8, (xv[0].x[i], ys[j + 1] if s % 2 else ys[j]))
            s += 1
    p = to_dev(ctx, hp)
    dt = 0.3 * cfl_dt(xv, V)
    jp.advection_(p, jp.RungeKutta2(), dev_V(ctx, V), dt)
    O.advection(hp, V, O.JP_RK2, 0.5, dt)
    assert_state_equal(p, hp, exact=True)


# ------------------------------------------------------------------ move
@pytest.mark.parametrize("D", [2, 3])
def test_move_parity(ctx, D):
    say("move", D)
    if D == 2:
        xv, xi = grid2d(30, 26)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 14, 3, seed=7)
    else:
        xv, xi = grid3d(9)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 20, 2, seed=7)
    T = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    Vd = dev_V(ctx, V)
    dt = cfl_dt(xv, V)
    for it in range(3):
        jp.advection_(p, jp.RungeKutta2(), Vd, dt)
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        st_d = jp.move_particles_(p, (Td,), stats=True)
        st_o = O.move_particles(hp, [T])
        say("move stats", it, st_d, st_o)
        assert st_d == st_o
        assert st_o["moved"] > 0
        assert_state_equal(p, hp, exact=True)
        assert_field(Td.data.flat(), T, exact=True)


@pytest.mark.parametrize("D,K,npq", [(2, 64, 12), (2, 54, 10), (3, 64, 6), (3, 54, 5)])
def test_move_parity_wide_K(ctx, D, K, npq):
    """Wide cells: K = 64 (every bit of the 64-bit occupancy words) and a K
    that is not a multiple of 8, with crowded cells, through slot groups and
    their per-cell byte stores."""
    say("move wide K", D, K)
    if D == 2:
        xv, xi = grid2d(20, 18)
        V = stream_V(xi)
    else:
        xv, xi = grid3d(7)
        V = cellular_V(xi)
    hp = host_particles(xv, xi, K, npq, seed=29)
    T = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    Vd = dev_V(ctx, V)
    dt = cfl_dt(xv, V)
    for it in range(3):
        jp.advection_(p, jp.RungeKutta2(), Vd, dt)
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        st_d = jp.move_particles_(p, (Td,), stats=True)
        st_o = O.move_particles(hp, [T])
        assert st_d == st_o
        assert st_o["moved"] > 0
        assert_state_equal(p, hp, exact=True)
        assert_field(Td.data.flat(), T, exact=True)


def test_move_small_K_drops_and_deletes(ctx):
    """Crowded cells (K small) force drops; a large dt pushes particles out of
    the domain (deletions)."""
    say("move drops")
    xv, xi = grid2d(10)
    Vx, Vy = stream_V(xi)
    V = (np.asfortranarray(Vx + 120.0), Vy)  # drift through the x = 1 wall
    hp = host_particles(xv, xi, 8, 2, seed=11)
    p = to_dev(ctx, hp)
    dt = 0.95 * cfl_dt(xv, V)
    Vd = dev_V(ctx, V)
    tot = {"dropped": 0, "deleted": 0}
    for it in range(4):
        jp.advection_(p, jp.Euler(), Vd, dt)
        O.advection(hp, V, O.JP_EULER, 0.5, dt)
        st_d = jp.move_particles_(p, (), stats=True)
        st_o = O.move_particles(hp)
        assert st_d == st_o
        for k in tot:
            tot[k] += st_o[k]
        assert_state_equal(p, hp, exact=True)
    say("move drops totals", tot)
    assert tot["dropped"] > 0 and tot["deleted"] > 0


@pytest.mark.parametrize("D", [2, 3])
def test_move_vertex_ties(ctx, D):
    """Particles exactly on vertex lines (h = 0.125, exact in binary).  From
    cell 1, a particle on the right vertex of cell 2 goes to cell 2, fails its
    strict in-cell test, and is re-slotted there when cell 2's colour phase
    comes later (an arrival that moves on: its origin is read back).  From
    cell 2, one on the right vertex of cell 3 stays in cell 3, whose phase has
    already run; the starting point carried from the re-slot puts it in slot 1.
    A particle on its own cell's left vertex is re-slotted in that cell.
    Compared with the oracle bit for bit, the field included."""
    say("move vertex ties", D)
    xv, xi = grid2d(8) if D == 2 else grid3d(8)
    hp = O.HostParticles(xv, 6, xi)
    clear(hp)
    T = np.full(hp.S, np.nan)
    t1, t2, y = (1,) * (D - 1), (2,) * (D - 1), (0.2,) * (D - 1)
    place(hp, (1,) + t1, 0, (0.375,) + y, [T], [1.0])
    place(hp, (1,) + t1, 1, (0.2,) + y, [T], [2.0])
    place(hp, (2,) + t1, 0, (0.3,) + y, [T], [3.0])
    place(hp, (2,) + t1, 2, (0.5,) + y, [T], [4.0])
    place(hp, (2,) + t2, 1, (0.25,) + (0.3,) * (D - 1), [T], [5.0])
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    st_d = jp.move_particles_(p, (Td,), stats=True)
    st_o = O.move_particles(hp, [T])
    say("vertex tie stats", st_d, st_o)
    assert st_d == st_o and st_o["moved"] == 4
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), T, exact=True)
    # the arrival re-slotted in cell 2 and the one kept in cell 3
    assert T[elem(hp, (2,) + t1, 1)] == 1.0 and T[elem(hp, (3,) + t1, 1)] == 4.0


def test_move_far_jumps_serial_path(ctx):
    """CFL > 1: movers jump more than one cell; the phases run serially and
    still match the oracle."""
    say("move far")
    xv, xi = grid2d(8)
    V = stream_V(xi)
    hp = host_particles(xv, xi, 16, 2, seed=13)
    p = to_dev(ctx, hp)
    dt = 2.5 * cfl_dt(xv, V)
    jp.advection_(p, jp.RungeKutta2(), dev_V(ctx, V), dt)
    O.advection(hp, V, O.JP_RK2, 0.5, dt)
    st_d = jp.move_particles_(p, (), stats=True)
    st_o = O.move_particles(hp)
    assert st_o["far"] > 0
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)


# ------------------------------------------------------------------ p2g / g2p
@pytest.mark.parametrize("D", [2, 3])
def test_p2g_g2p_parity(ctx, D):
    say("p2g/g2p", D)
    if D == 2:
        xv, xi = grid2d(20, 17)
        hp = host_particles(xv, xi, 12, 2, seed=17)
    else:
        xv, xi = grid3d(7)
        hp = host_particles(xv, xi, 16, 1, seed=17)
    rng = np.random.default_rng(2)
    Fp = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    Fp2 = np.where(hp.index == 1, rng.random(hp.S) - 0.5, np.nan)
    # knock out a region so some nodes get NaN (0/0)
    hp.index[: hp.C // 3] = 0
    p = to_dev(ctx, hp)
    Fpd = jp.CellArray(ctx, hp.ni, hp.K)
    Fpd.copy_from(Fp)
    Fpd2 = jp.CellArray(ctx, hp.ni, hp.K)
    Fpd2.copy_from(Fp2)
    F = jp.nodal_field(p)
    jp.particle2grid_(F, Fpd, p)
    assert_field(F.flat(), O.particle2grid(hp, Fp), exact=True)
    F1, F2 = jp.nodal_field(p), jp.nodal_field(p)
    jp.particle2grid_((F1, F2), (Fpd, Fpd2), p)
    m1, m2 = O.particle2grid_multi(hp, [Fp, Fp2])
    assert_field(F1.flat(), m1, exact=True)
    assert_field(F2.flat(), m2, exact=True)
    # g2p of a smooth nodal field
    Tn = np.asfortranarray(np.cos(3 * np.add.outer(xv[0].x, xv[1].x)) if D == 2 else
                           np.cos(3 * np.add.outer(np.add.outer(xv[0].x, xv[1].x), xv[2].x)))
    Td = jp.DeviceArray.from_host(ctx, Tn)
    pT = jp.CellArray(ctx, hp.ni, hp.K, fill=0.0)
    jp.grid2particle_(pT, Td, p)
    want = O.grid2particle(hp, Tn, np.zeros(hp.S))
    assert_field(pT.data.flat(), want, exact=True)


# ------------------------------------------------------------------ inject
@pytest.mark.parametrize("D", [2, 3])
def test_inject_parity(ctx, D):
    say("inject", D)
    if D == 2:
        xv, xi = grid2d(18, 15)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 16, 2, seed=19)
        minx = 10
    else:
        xv, xi = grid3d(7)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 20, 1, seed=19)
        minx = 12
    rng = np.random.default_rng(4)
    # thin the population so many quadrants are deficient
    drop = (rng.random(hp.S) < 0.45) & (hp.index == 1)
    hp.index[drop] = 0
    for d in range(hp.D):
        hp.coords[d][drop] = np.nan
    T = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    p = to_dev(ctx, hp, min_xcell=minx, seed=99)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    st_d = jp.inject_particles_(p, (Td,), stats=True)
    st_o = O.inject_particles(hp, [T], minx, seed=99, counter=0)
    say("inject stats", st_d, st_o)
    assert st_o["injected"] > 0
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), T, exact=True)


# ------------------------------------------------------------------ full step
@pytest.mark.parametrize("D", [2, 3])
def test_time_loop_parity(ctx, D):
    """advect -> move -> inject -> p2g -> g2p for several steps (the loop of
    test/test_2D.jl:516-524)."""
    say("time loop", D)
    if D == 2:
        xv, xi = grid2d(32)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 24, 3, seed=23)
        minx, steps = 8, 6
    else:
        xv, xi = grid3d(8)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 24, 2, seed=23)
        minx, steps = 10, 3
    Tn = np.asfortranarray(np.random.default_rng(1).random(tuple(a.n for a in xv)))
    pT = O.grid2particle(hp, Tn, np.full(hp.S, np.nan))
    p = to_dev(ctx, hp, min_xcell=minx, seed=5)
    pTd = jp.CellArray(ctx, hp.ni, hp.K)
    pTd.copy_from(pT)
    Vd = dev_V(ctx, V)
    Fd = jp.nodal_field(p)
    dt = cfl_dt(xv, V)
    for it in range(steps):
        jp.advection_(p, jp.RungeKutta2(), Vd, dt)
        jp.move_particles_(p, (pTd,))
        jp.inject_particles_(p, (pTd,))
        jp.particle2grid_(Fd, pTd, p)
        jp.grid2particle_(pTd, Fd, p)
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        O.move_particles(hp, [pT])
        O.inject_particles(hp, [pT], minx, seed=5, counter=it)
        F = O.particle2grid(hp, pT)
        O.grid2particle(hp, F, pT)
        assert_state_equal(p, hp, exact=True)
        assert_field(Fd.flat(), F, exact=True)
        assert_field(pTd.data.flat(), pT, exact=True)


def test_determinism(ctx):
    say("determinism")
    xv, xi = grid2d(40)
    V = stream_V(xi)
    outs = []
    for _ in range(2):
        p = jp.init_particles(ctx, 24, 48, 12, *xi, seed=8)
        Vd = dev_V(ctx, V)
        T = jp.CellArray(ctx, p.ni, p.max_xcell, fill=1.0)
        for _ in range(3):
            jp.advection_(p, jp.RungeKutta2(), Vd, cfl_dt(xv, V))
            jp.move_particles_(p, (T,))
            jp.inject_particles_(p, (T,))
        outs.append(p.host_state())
    assert np.array_equal(outs[0][1], outs[1][1])
    for d in range(2):
        assert np.array_equal(outs[0][0][d], outs[1][0][d], equal_nan=True)


def test_error_paths(ctx):
    say("errors")
    with pytest.raises(ValueError):
        jp.RungeKutta2(1.1)
    xv, xi = grid2d(4)
    p = jp.init_particles(ctx, 4, 8, 2, *xi, seed=1)
    with pytest.raises(ValueError):
        jp.advection_(p, jp.RungeKutta2(), (jp.DeviceArray(ctx, (3, 3)),) * 2, 0.1)
    F = jp.DeviceArray(ctx, (4, 4))
    with pytest.raises(ValueError):
        jp.particle2grid_(F, jp.CellArray(ctx, p.ni, p.max_xcell), p)
