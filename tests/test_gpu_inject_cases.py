"""Injection edge cases of the 2D one-field kernels against the oracle, FP64
bit for bit and the split form with the same masks and copies (runs on an
MI355X).

* An empty 5x5 block: its middle cell's first injection has no neighbour at
  all (an orphan: the field keeps its slot's value), and the cell's later
  injections find that particle nearest and copy from it.
* Slots running out inside a quadrant: K = 6, min_xcell 12 (3 per quadrant),
  a cell holding 4 particles in its last quadrant gets the 2 free slots in
  the first one (K = 16: 3 in each of the three before the full one); a cell
  whose first quadrant is full is left alone.
* No field at all.
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.grids import Axis, staggered_grid
from oracle import oracle as O

from helpers import clear, place
from test_gpu_parity import assert_state_equal, to_dev
from test_gpu_split import split_state

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-inject-cases]", *a, flush=True)


def grid(n=12):
    xv = (Axis(np.arange(n + 1) / n, "range"), Axis(np.arange(n + 1) / n, "range"))
    return xv, staggered_grid(xv)


def edge_state(K, seed):
    """n x n cells (n = 12, dyadic-free 1/12 spacing), random particles except
    an empty 5x5 block around cell (5, 5) and hand-placed cells near (9, 2)"""
    xv, xi = grid()
    hp = O.HostParticles(xv, K, xi)
    clear(hp)
    rng = np.random.default_rng(seed)
    h = 1.0 / 12
    for i in range(12):
        for j in range(12):
            if abs(i - 5) <= 2 and abs(j - 5) <= 2:
                continue
            if (i, j) in ((9, 2), (10, 2)):
                continue
            k = int(rng.integers(0, K + 1))
            for s in range(k):
                place(hp, (i, j), s, ((i + rng.uniform(0.02, 0.98)) * h,
                                       (j + rng.uniform(0.02, 0.98)) * h))
    # (9, 2): 4 particles in quadrant 2 (upper right), none in quadrant 0
    for s in range(4):
        place(hp, (9, 2), s, ((9 + 0.6 + 0.05 * s) * h, (2 + 0.7) * h))
    # (10, 2): first quadrant full (3 particles) -> nothing injected
    for s in range(3):
        place(hp, (10, 2), s, ((10 + 0.1 + 0.1 * s) * h, (2 + 0.2) * h))
    return hp


@pytest.mark.parametrize("nf", [0, 1])
@pytest.mark.parametrize("K", [6, 16])
def test_inject_edge_cases(ctx, K, nf):
    say("edge cases K", K, "fields", nf)
    hp = edge_state(K, seed=K + nf)
    rng = np.random.default_rng(3)
    # empty slots hold a value of their own, which an orphan keeps
    T = np.where(hp.index == 1, rng.random(hp.S), -2.0 - rng.random(hp.S))
    minx = 12
    p = to_dev(ctx, hp, min_xcell=minx, seed=5)
    s = jp.to_split(p)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    Ts = Td.copy()
    args_d = (Td,) if nf else ()
    args_s = (Ts,) if nf else ()
    st_d = jp.inject_particles_(p, args_d, stats=True)
    st_s = jp.inject_particles_(s, args_s, stats=True)
    st_o = O.inject_particles(hp, [T] if nf else [], minx, seed=5, counter=0)
    say("stats", st_d, st_s, st_o)
    assert st_o["orphan"] >= 1 and st_o["injected"] > st_o["orphan"]
    assert st_d == st_o and st_s == st_o
    assert_state_equal(p, hp, exact=True)
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    if nf:
        assert np.array_equal(Td.data.flat(), T, equal_nan=True)
        assert np.array_equal(Ts.data.flat(), T, equal_nan=True)
    # the cells built by hand
    C = hp.C
    c = int(np.ravel_multi_index((9, 2), hp.ni, order="F"))
    got = [int(index[c + C * k]) for k in range(K)]
    # K = 6: the 2 free slots go to the first quadrant; K = 16: 3 to each of
    # the three quadrants before the full one
    assert sum(got) == (6 if K == 6 else 13)
    c = int(np.ravel_multi_index((10, 2), hp.ni, order="F"))
    assert sum(int(index[c + C * k]) for k in range(K)) == 3
