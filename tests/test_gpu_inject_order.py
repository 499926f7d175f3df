"""The 2D injection's four colours in one launch, against the oracle and
against one launch per colour (runs on an MI355X).

In the single launch a listed cell waits for its listed neighbours of
earlier colours, so the states here are built to have many of them: random
occupancy with rectangles of empty or nearly empty cells, where every colour
is listed next to every other (chains of four dependent cells).  Compared
bit for bit with the oracle (FP64) and with the per-colour launches (FP64
and split: masks, positions and the copied field), with the default number
of persistent wavefronts, with 64 of them (each takes many cells, so a
wavefront waits on cells that other wavefronts take later in their own
order) and with one workgroup of four.  Two calls in a row on one context
check that the per-cell state of one call does not leak into the next; two
contexts inject at once; edge cases: 2 x 2 cells, two rows or columns, every
cell listed (all empty, with orphans), K = 64, and a wavefront count far
above what the device holds at once.
"""
import contextlib

import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip import _abi
from justpic_hip.grids import Axis, staggered_grid
from oracle import oracle as O

from helpers import clear, place
from test_gpu_parity import assert_state_equal, to_dev
from test_gpu_split import split_state

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-inject-order]", *a, flush=True)


@contextlib.contextmanager
def options(ctx, **kv):
    ids = {"phases": _abi.JP_OPT_INJECT_PHASES, "waves": _abi.JP_OPT_INJECT_WAVES}
    for k, v in kv.items():
        ctx.set_option(ids[k], v)
    try:
        yield
    finally:
        for k in kv:
            ctx.set_option(ids[k], 0)


def clustered_state(n0, n1, K, seed, holes=40):
    """Random occupancy (0..K particles per cell) with `holes` rectangles of
    2..9 cells a side holding at most one particle per cell."""
    xv = (Axis(np.arange(n0 + 1) / n0, "range"), Axis(np.arange(n1 + 1) / n1, "range"))
    xi = staggered_grid(xv)
    hp = O.HostParticles(xv, K, xi)
    clear(hp)
    rng = np.random.default_rng(seed)
    cnt = rng.integers(0, K + 1, (n0, n1))
    for _ in range(holes):
        w, h = rng.integers(2, 10, 2)
        i0, j0 = rng.integers(0, n0 - 1), rng.integers(0, n1 - 1)
        cnt[i0:i0 + w, j0:j0 + h] = rng.integers(0, 2, cnt[i0:i0 + w, j0:j0 + h].shape)
    for i in range(n0):
        for j in range(n1):
            for s in range(int(cnt[i, j])):
                place(hp, (i, j), s, ((i + rng.uniform(0.02, 0.98)) / n0,
                                      (j + rng.uniform(0.02, 0.98)) / n1))
    return hp


def run_pair(ctx, hp, T, minx, seed, **opt):
    """The FP64 and split forms on copies of one state, with options."""
    p = to_dev(ctx, hp, min_xcell=minx, seed=seed)
    s = jp.to_split(p)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    Ts = Td.copy()
    with options(ctx, **opt):
        st_d = jp.inject_particles_(p, (Td,), stats=True)
        st_s = jp.inject_particles_(s, (Ts,), stats=True)
    return p, s, Td, Ts, st_d, st_s


@pytest.mark.parametrize("waves", [0, 64, 1])
@pytest.mark.parametrize("shape", [(160, 96), (37, 211)])
def test_inject_one_launch(ctx, shape, waves):
    n0, n1 = shape
    say("one launch", shape, "waves", waves)
    K, minx, seed = 16, 8, 7
    hp = clustered_state(n0, n1, K, seed=n0 + n1)
    rng = np.random.default_rng(1)
    T = np.where(hp.index == 1, rng.random(hp.S), -1.0 - rng.random(hp.S))
    p, s, Td, Ts, st_d, st_s = run_pair(ctx, hp, T, minx, seed, waves=waves)
    q, r, Qd, Qs, sq_d, sq_s = run_pair(ctx, hp, T, minx, seed, phases=1)
    To = T.copy()
    st_o = O.inject_particles(hp, [To], minx, seed=seed, counter=0)
    say("stats", st_d, st_s, st_o)
    assert st_o["injected"] > 1000
    assert st_d == st_o and st_s == st_o and sq_d == st_o and sq_s == st_o
    assert_state_equal(p, hp, exact=True)
    assert_state_equal(q, hp, exact=True)
    assert np.array_equal(Td.data.flat(), To, equal_nan=True)
    assert np.array_equal(Qd.data.flat(), To, equal_nan=True)
    # split: the same masks, offsets and copies as its per-colour launches
    c1, i1 = split_state(s)
    c2, i2 = split_state(r)
    assert np.array_equal(i1, hp.index) and np.array_equal(i2, hp.index)
    for d in range(2):
        assert np.array_equal(c1[d], c2[d], equal_nan=True)
    assert np.array_equal(Ts.data.flat(), Qs.data.flat(), equal_nan=True)


def test_inject_calls_in_a_row(ctx):
    """Two states with different listed cells, one after the other on one
    context: each bit for bit against the oracle."""
    say("calls in a row")
    K, minx = 12, 8
    for it, seed in enumerate((21, 22)):
        hp = clustered_state(96, 64, K, seed=seed, holes=25)
        rng = np.random.default_rng(seed)
        T = np.where(hp.index == 1, rng.random(hp.S), np.nan)
        p = to_dev(ctx, hp, min_xcell=minx, seed=seed)
        Td = jp.CellArray(ctx, hp.ni, hp.K)
        Td.copy_from(T)
        st_d = jp.inject_particles_(p, (Td,), stats=True)
        st_o = O.inject_particles(hp, [T], minx, seed=seed, counter=0)
        say("call", it, st_d, st_o)
        assert st_d == st_o
        assert_state_equal(p, hp, exact=True)
        assert np.array_equal(Td.data.flat(), T, equal_nan=True)


def test_inject_two_streams(ctx):
    """Two contexts (two streams) inject at once with no synchronisation
    between the calls: the ordered launches on one device run one after the
    other (the second waits for the first), and each result is the
    oracle's."""
    say("two streams")
    K, minx = 16, 8
    ctxs = [jp.Context(0, synchronous=False) for _ in range(2)]
    runs = []
    for i, cx in enumerate(ctxs):
        hp = clustered_state(128, 96, K, seed=40 + i, holes=30)
        rng = np.random.default_rng(i)
        T = np.where(hp.index == 1, rng.random(hp.S), np.nan)
        p = to_dev(cx, hp, min_xcell=minx, seed=3 + i)
        Td = jp.CellArray(cx, hp.ni, hp.K)
        Td.copy_from(T)
        cx.sync()
        runs.append((cx, hp, T, p, Td, 3 + i))
    for _ in range(2):  # two rounds: each context's second call follows the other's
        for cx, hp, T, p, Td, seed in runs:
            jp.inject_particles_(p, (Td,))
    for cx, *_ in runs:
        cx.sync()
    for cx, hp, T, p, Td, seed in runs:
        st1 = O.inject_particles(hp, [T], minx, seed=seed, counter=0)
        st2 = O.inject_particles(hp, [T], minx, seed=seed, counter=1)
        say("oracle", st1, st2)
        assert st1["injected"] > 500
        assert_state_equal(p, hp, exact=True)
        assert np.array_equal(Td.data.flat(), T, equal_nan=True)


def empty_state(n0, n1, K, seed, fill=0.0):
    """n0 x n1 cells holding a fraction `fill` of particles (0: all empty)."""
    xv = (Axis(np.arange(n0 + 1) / n0, "range"), Axis(np.arange(n1 + 1) / n1, "range"))
    hp = O.HostParticles(xv, K, staggered_grid(xv))
    clear(hp)
    rng = np.random.default_rng(seed)
    for i in range(n0):
        for j in range(n1):
            for s in range(K):
                if rng.random() < fill:
                    place(hp, (i, j), s, ((i + rng.uniform(0.02, 0.98)) / n0,
                                          (j + rng.uniform(0.02, 0.98)) / n1))
    return hp


EDGE = [
    # (n0, n1, K, min_xcell, fill)
    (2, 2, 16, 8, 0.1),     # four cells, every one a neighbour of the others
    (33, 2, 12, 8, 0.2),    # two rows
    (2, 29, 12, 8, 0.2),    # two columns
    (40, 24, 16, 12, 0.0),  # every cell listed (all empty): the longest chains
    (24, 20, 64, 32, 0.3),  # K = 64
]


@pytest.mark.parametrize("case", EDGE)
def test_inject_one_launch_edges(ctx, case):
    n0, n1, K, minx, fill = case
    say("edges", case)
    hp = empty_state(n0, n1, K, seed=n0 * 7 + n1, fill=fill)
    rng = np.random.default_rng(2)
    T = np.where(hp.index == 1, rng.random(hp.S), -1.0 - rng.random(hp.S))
    p, s, Td, Ts, st_d, st_s = run_pair(ctx, hp, T, minx, 5)
    q, r, Qd, Qs, sq_d, sq_s = run_pair(ctx, hp, T, minx, 5, phases=1)
    st_o = O.inject_particles(hp, [T], minx, seed=5, counter=0)
    say("stats", st_d, st_o)
    assert st_o["injected"] > 0
    assert st_d == st_o and st_s == st_o and sq_d == st_o and sq_s == st_o
    assert_state_equal(p, hp, exact=True)
    assert np.array_equal(Td.data.flat(), T, equal_nan=True)
    c1, i1 = split_state(s)
    c2, i2 = split_state(r)
    assert np.array_equal(i1, hp.index) and np.array_equal(i2, hp.index)
    for d in range(2):
        assert np.array_equal(c1[d], c2[d], equal_nan=True)
    assert np.array_equal(Ts.data.flat(), Qs.data.flat(), equal_nan=True)


def test_inject_one_launch_more_waves_than_resident(ctx):
    """A wavefront count far above what the device holds at once: the ordered
    launch is capped at the resident count (every wait ends) and the result
    is the oracle's."""
    say("waves above residency")
    hp = clustered_state(96, 80, 16, seed=5)
    rng = np.random.default_rng(4)
    T = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    p, s, Td, Ts, st_d, st_s = run_pair(ctx, hp, T, 8, 9, waves=1 << 20)
    st_o = O.inject_particles(hp, [T], 8, seed=9, counter=0)
    assert st_d == st_o and st_s == st_o
    assert_state_equal(p, hp, exact=True)
    assert np.array_equal(Td.data.flat(), T, equal_nan=True)
