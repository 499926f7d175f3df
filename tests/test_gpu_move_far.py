"""Moves in which particles jump more than one cell (runs on an MI355X).

Such a move runs the colour phases' slot loops in one workgroup, cell by cell
in levels that keep the reference's order between any two cells that can
interact (the level spacing follows the largest jump), finding the far
destinations from the staged positions; the scatter then moves the data as
in a regular move.  Compared with the oracle bit for
bit: FP64 and FP32-split, 2D and 3D, jumps of 2 to 20 cells, crowded cells
(drops), particles leaving the domain (deletions), and a regular move on the
same context afterwards (the far-jump state is cleared after use).
"""
import numpy as np
import pytest

import justpic_hip as jp
from oracle import oracle as O

from helpers import grid2d, grid3d
from test_gpu_parity import assert_field, assert_state_equal, to_dev
from test_gpu_split import split_state

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-move-far]", *a, flush=True)


def jumped_state(ni, K, fill, far_frac, R, seed, on_vertex=0.0):
    """Particles in a fraction `fill` of the slots, displaced by -1..1 cells,
    and a fraction far_frac of them by up to R cells per dimension (some
    beyond the domain).  Offsets are odd multiples of 1/64 of a cell (exact
    in float32 on power-of-two grids), or with probability on_vertex exactly
    0 (the particle sits on a vertex line)."""
    D = len(ni)
    xv, xi = grid2d(*ni) if D == 2 else grid3d(ni[0])
    hp = O.HostParticles(xv, K, xi)
    rng = np.random.default_rng(seed)
    occ = rng.random(hp.S) < fill
    hp.index[:] = occ.astype(np.uint8)
    cells = np.arange(hp.S) % hp.C
    I = np.stack(np.unravel_index(cells, hp.ni, order="F"), 0)
    far = rng.random(hp.S) < far_frac
    for d in range(D):
        u = (2 * rng.integers(0, 32, hp.S) + 1) / 64.0
        if on_vertex:
            u = np.where(rng.random(hp.S) < on_vertex, 0.0, u)
        shift = np.where(far, rng.integers(-R, R + 1, hp.S), rng.choice([-1, 0, 0, 1], hp.S))
        x = (I[d] + u + shift) / ni[d]
        hp.coords[d][:] = np.where(occ, x, np.nan)
    return hp


CASES = [
    # (ni, K, fill, far_frac, R)
    ((32, 24), 12, 0.6, 0.05, 2),
    ((40, 36), 16, 0.7, 0.02, 4),
    ((16, 64), 8, 0.8, 0.1, 3),
    ((20, 20), 10, 0.5, 0.01, 20),
    ((8, 8, 8), 12, 0.6, 0.05, 2),
    ((8, 8, 8), 16, 0.5, 0.02, 5),
]


@pytest.mark.parametrize("case", CASES)
def test_far_move_fp64(ctx, case):
    ni, K, fill, ff, R = case
    say("fp64", case)
    hp = jumped_state(ni, K, fill, ff, R, seed=sum(ni) + R)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    st_d = jp.move_particles_(p, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    say("stats", st_d, st_o)
    assert st_o["far"] > 0
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), ids, exact=True)


@pytest.mark.parametrize("case", [c for c in CASES if all(n & (n - 1) == 0 for n in c[0])])
def test_far_move_split(ctx, case):
    ni, K, fill, ff, R = case
    say("split", case)
    hp = jumped_state(ni, K, fill, ff, R, seed=3 * sum(ni) + R)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    s = jp.to_split(to_dev(ctx, hp))
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    st_d = jp.move_particles_(s, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    say("stats", st_d, st_o)
    assert st_o["far"] > 0
    assert st_d == st_o
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    for d in range(len(ni)):
        assert np.array_equal(coords[d], hp.coords[d], equal_nan=True)
    assert np.array_equal(Td.data.flat(), ids, equal_nan=True)


@pytest.mark.parametrize("D", [2, 3])
def test_far_move_vertex_lines(ctx, D):
    """FP64 far jumps onto vertex lines: a particle that lands on the edge of
    its destination fails the strict test there and is looked up again when
    that cell's phase comes later (bit for bit against the oracle)."""
    say("far vertex", D)
    ni = (32, 32) if D == 2 else (8, 8, 8)
    hp = jumped_state(ni, 12, 0.6, 0.1, 3, seed=11 + D, on_vertex=0.25)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    st_d = jp.move_particles_(p, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    say("stats", st_d, st_o)
    assert st_o["far"] > 0
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), ids, exact=True)


def test_regular_move_after_far_move(ctx):
    """A far-jump move, then a move with jumps of at most one cell on the same
    context: both bit for bit against the oracle, the second with no far
    movers (the far-jump state does not outlive its move)."""
    say("far then regular")
    hp = jumped_state((32, 32), 12, 0.6, 0.05, 3, seed=77)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(ids)
    st_d = jp.move_particles_(p, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    assert st_o["far"] > 0 and st_d == st_o
    # shift every survivor by less than one cell and move again
    rng = np.random.default_rng(5)
    m = hp.index == 1
    for d in range(2):
        step = rng.choice([-0.7, 0.0, 0.7], hp.S) / hp.ni[d]
        x = np.clip(hp.coords[d] + np.where(m, step, 0.0), 0.5 / 64 / hp.ni[d],
                    1.0 - 0.5 / 64 / hp.ni[d])
        hp.coords[d][:] = np.where(m, x, np.nan)
        p.coords[d].copy_from(hp.coords[d])
    st_d = jp.move_particles_(p, (Td,), stats=True)
    st_o = O.move_particles(hp, [ids])
    say("second move", st_d, st_o)
    assert st_o["far"] == 0 and st_o["moved"] > 0
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), ids, exact=True)
