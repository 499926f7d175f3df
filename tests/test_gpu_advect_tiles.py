"""FP64 advection with the velocity staged in LDS per tile of cells, against
the oracle and against the slot kernel (runs on an MI355X).

The tiles are 64x4 (2D) and 16x4x4 (3D) cells; a component's window covers
the RK stage points of a step of up to one cell.  A particle whose stage
point leaves the window, or sits on a velocity node, is advected again on the
slot kernel's path.  The cases: grids spanning several tiles with partial
tiles at the ends, steps of up to one cell (everything from the windows),
steps of about three cells (many stage points outside), refined (vector)
axes (the window guess misses), particles on velocity-grid lines, slot
groups, and all four integrators.  Bit for bit.
"""
import contextlib

import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip import _abi
from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V
from test_gpu_parity import assert_state_equal, dev_V, to_dev

pytestmark = pytest.mark.gpu

METHODS = [
    ("euler", jp.Euler(), O.JP_EULER, 0.5),
    ("rk2", jp.RungeKutta2(), O.JP_RK2, 0.5),
    ("rk2_23", jp.RungeKutta2(2 / 3), O.JP_RK2, 2 / 3),
    ("rk4", jp.RungeKutta4(), O.JP_RK4, 0.5),
]


def say(*a):
    print("[gpu-advect-tiles]", *a, flush=True)


@contextlib.contextmanager
def option(ctx, opt, value):
    ctx.set_option(opt, value)
    try:
        yield
    finally:
        ctx.set_option(opt, 0)


def case(D, kind="range"):
    if D == 2:
        xv, xi = grid2d(150, 13, kind=kind)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 12, 2, seed=9)
    else:
        xv, xi = grid3d(20)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 10, 1, seed=9)
    return xv, xi, V, hp


@pytest.mark.parametrize("name,method,code,alpha", METHODS)
@pytest.mark.parametrize("cfl", [0.9, 3.0])
@pytest.mark.parametrize("D", [2, 3])
def test_tiled_advection_fp64(ctx, D, cfl, name, method, code, alpha):
    say("tiles", D, cfl, name)
    xv, xi, V, hp = case(D)
    dt = cfl * cfl_dt(xv, V)
    p = to_dev(ctx, hp)
    q = to_dev(ctx, hp)
    Vd = dev_V(ctx, V)
    jp.advection_(p, method, Vd, dt)
    with option(ctx, _abi.JP_OPT_ADVECT_SLOTS, 1):
        jp.advection_(q, method, Vd, dt)
    O.advection(hp, V, code, alpha, dt)
    assert_state_equal(p, hp, exact=True)
    assert_state_equal(q, hp, exact=True)


@pytest.mark.parametrize("span", [8, 16])
def test_tiled_advection_slot_groups(ctx, span):
    """Slot groups on blockIdx.y (each stages its own windows)."""
    say("tiles span", span)
    xv, xi, V, hp = case(3)
    dt = cfl_dt(xv, V)
    p = to_dev(ctx, hp)
    with option(ctx, _abi.JP_OPT_SLOT_SPAN, span):
        jp.advection_(p, jp.RungeKutta4(), dev_V(ctx, V), dt)
    O.advection(hp, V, O.JP_RK4, 0.5, dt)
    assert_state_equal(p, hp, exact=True)


def test_tiled_advection_vector_axes(ctx):
    """Refined (vector) axes: the window guess assumes uniform spacing and
    misses often; every miss takes the slot kernel's path."""
    say("tiles vector axes")
    xv, xi, V, hp = case(2, kind="vector")
    V = stream_V(xi, amp=5.0)
    dt = 0.5 * min(np.min(xv[0].dx), np.min(xv[1].dx)) / 5.0
    p = to_dev(ctx, hp)
    jp.advection_(p, jp.RungeKutta2(), dev_V(ctx, V), dt)
    O.advection(hp, V, O.JP_RK2, 0.5, dt)
    assert_state_equal(p, hp, exact=True)


def test_tiled_advection_node_lines(ctx):
    """Particles on the x lines of Vx's nodes and the y lines of Vy's."""
    say("tiles node lines")
    xv, xi = grid2d(70, 9)
    V = stream_V(xi)
    hp = O.HostParticles(xv, 8, xi)
    rng = np.random.default_rng(2)
    cells = np.arange(hp.S) % hp.C
    I = np.stack(np.unravel_index(cells, hp.ni, order="F"), 0)
    occ = rng.random(hp.S) < 0.5
    hp.index[:] = occ.astype(np.uint8)
    h = [1.0 / 70, 1.0 / 9]
    for d in range(2):
        u = rng.random(hp.S)
        u = np.where(rng.random(hp.S) < 0.3, 0.0 if d == 0 else 0.5, u)
        hp.coords[d][:] = np.where(occ, (I[d] + u) * h[d], np.nan)
    p = to_dev(ctx, hp)
    dt = 0.5 * cfl_dt(xv, V)
    jp.advection_(p, jp.RungeKutta4(), dev_V(ctx, V), dt)
    O.advection(hp, V, O.JP_RK4, 0.5, dt)
    assert_state_equal(p, hp, exact=True)
