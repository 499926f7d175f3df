"""The 2D move's colour phases in one launch over tiles with halos, against
the oracle and against one launch per phase (runs on an MI355X).

The tiles are 32 x 16 cells with a halo of 4/6 cells in x and 10/18 in y, so
every grid here spans several tiles.  The states are built to make the slot
assignment depend on the phase order across tile edges: crowded cells (few
free slots, so destinations fill up and particles are dropped), displacements
of up to one cell in both directions, a drift along y (the long side of the
dependency cone), and on the dyadic grids particles exactly on vertex lines
(arrivals that fail the strict test of their new cell and are re-slotted
there in a later phase).
"""
import contextlib

import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip import _abi
from oracle import oracle as O

from helpers import grid2d
from test_gpu_parity import assert_field, assert_state_equal, to_dev
from test_gpu_split import split_state

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-move-tiles]", *a, flush=True)


@contextlib.contextmanager
def phases_option(ctx, value):
    ctx.set_option(_abi.JP_OPT_MOVE_PHASES, value)
    try:
        yield
    finally:
        ctx.set_option(_abi.JP_OPT_MOVE_PHASES, 0)


def shifted_state(n0, n1, K, fill, seed, drift_y=0.0, on_vertex=0.0):
    """Particles in a fraction `fill` of the slots, each displaced by -1, 0 or
    +1 cells per dimension (+1 in y more often when drift_y > 0).  With
    on_vertex > 0 the grid must be dyadic: that fraction of the coordinates
    sits exactly on a vertex line."""
    xv, xi = grid2d(n0, n1)
    hp = O.HostParticles(xv, K, xi)
    rng = np.random.default_rng(seed)
    occ = rng.random(hp.S) < fill
    hp.index[:] = occ.astype(np.uint8)
    cells = np.arange(hp.S) % hp.C
    I = np.stack(np.unravel_index(cells, hp.ni, order="F"), 0)
    for d, n in enumerate((n0, n1)):
        u = (2 * rng.integers(0, 32, hp.S) + 1) / 64.0
        if on_vertex:
            u = np.where(rng.random(hp.S) < on_vertex, 0.0, u)
        shift = rng.choice([-1, 0, 0, 1], hp.S)
        if d == 1 and drift_y:
            shift = np.where(rng.random(hp.S) < drift_y, 1, shift)
        # keep inside the domain (strictly: a particle on the wall is deleted,
        # which the oracle and the kernels agree on, but it moves nothing)
        x = np.clip(I[d] + u + shift, 1.0 / 64, n - 1.0 / 64) / n
        hp.coords[d][:] = np.where(occ, x, np.nan)
    return hp


CASES = [
    # (n0, n1, K, fill, drift_y, on_vertex)
    (64, 128, 16, 0.75, 0.0, 0.0),
    (64, 64, 8, 0.8, 0.7, 0.0),
    (128, 32, 24, 0.6, 0.3, 0.1),
    (32, 256, 12, 0.7, 0.9, 0.05),
    (97, 83, 16, 0.75, 0.4, 0.0),
    (5, 150, 10, 0.8, 0.8, 0.0),
]


@pytest.mark.parametrize("phases", [0, 1])
@pytest.mark.parametrize("case", CASES)
def test_move_tiles_fp64(ctx, case, phases):
    """FP64 move: masks, positions, the field (particle ids) and the outcome
    counts bit for bit against the oracle, tiled (0) and per-phase (1)."""
    n0, n1, K, fill, drift, vert = case
    say("fp64", case, "phases", phases)
    hp = shifted_state(n0, n1, K, fill, seed=n0 * 7 + n1, drift_y=drift, on_vertex=vert)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    with phases_option(ctx, phases):
        st_d = jp.move_particles_(p, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    say("stats", st_d, st_o)
    assert st_o["moved"] > 0
    assert st_o["dropped"] > 0 or fill < 0.7
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), ids, exact=True)


@pytest.mark.parametrize("case", [c for c in CASES if c[0] & (c[0] - 1) == 0
                                  and c[1] & (c[1] - 1) == 0 and c[5] == 0.0])
def test_move_tiles_split(ctx, case):
    """FP32-split move on the same dyadic states (exact in float32): bit for
    bit against the FP64 oracle."""
    n0, n1, K, fill, drift, _ = case
    say("split", case)
    hp = shifted_state(n0, n1, K, fill, seed=n0 * 5 + n1, drift_y=drift)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    s = jp.to_split(to_dev(ctx, hp))
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    st_d = jp.move_particles_(s, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    say("stats", st_d, st_o)
    assert st_d == st_o
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    for d in range(2):
        assert np.array_equal(coords[d], hp.coords[d], equal_nan=True)
    assert np.array_equal(Td.data.flat(), ids, equal_nan=True)


def test_move_tiles_repeated_steps(ctx):
    """Five moves in a row on one state (each move starts from the occupancy
    the previous one left, holes included), tiled against the oracle."""
    say("repeated")
    hp = shifted_state(96, 80, 16, 0.7, seed=3, drift_y=0.5)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    rng = np.random.default_rng(8)
    for it in range(5):
        st_d = jp.move_particles_(p, (Td,), stats=True)
        st_o = O.move_particles(hp, [ids])
        assert st_d == st_o, it
        assert_state_equal(p, hp, exact=True)
        assert_field(Td.data.flat(), ids, exact=True)
        # displace the survivors again by less than one cell (y drift)
        m = hp.index == 1
        for d, n in enumerate(hp.ni):
            step = rng.choice([-0.6, 0.0, 0.9] if d == 1 else [-0.7, 0.0, 0.7], hp.S)
            x = np.clip(hp.coords[d] * n + np.where(m, step, 0.0), 1.0 / 64, n - 1.0 / 64) / n
            hp.coords[d][:] = np.where(m, x, np.nan)
        for d in range(2):
            p.coords[d].copy_from(hp.coords[d])
