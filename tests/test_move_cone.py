"""The halo of the 2D move's tiled colour phases (CPU).

The tiled kernel replays the nine colour phases of a tile of cells plus a
halo in LDS.  Its validity test and its halo rest on a closed form for the
dependency cone of phase p = 3 c0 + c1 (c0 = I0 mod 3, c1 = I1 mod 3): the
offsets from a cell of phase p to the cells whose processing in an earlier
phase can change what p's processing reads lie in
    dimension 0: -c0 .. 2 c0,    dimension 1: -p .. 2 p.
Here the cone is recomputed by brute force over every chain of processings
(a processing reads and writes its cell's 3x3 neighbourhood, so it depends on
the processings of earlier phases within 2 cells, and on theirs), and the
halo constants of the kernel are read from its source.
"""
import itertools
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "justpic.jl_amd", "csrc", "jp_move.hip")


def phase(c):
    return 3 * (c[0] % 3) + c[1] % 3


def reach(p):
    """All offsets reachable from a cell of phase p by chains of processings
    of strictly earlier phases, each within 2 cells of the previous one
    (colours taken from the actual offsets on an infinite grid)."""
    base = (3 * 10 + p // 3, 3 * 10 + p % 3)  # a cell of phase p
    seen = {base}
    frontier = [base]
    while frontier:
        nxt = []
        for a in frontier:
            for d in itertools.product(range(-2, 3), repeat=2):
                b = (a[0] + d[0], a[1] + d[1])
                if phase(b) < phase(a) and b not in seen:
                    seen.add(b)
                    nxt.append(b)
        frontier = nxt
    return [(b[0] - base[0], b[1] - base[1]) for b in seen]


def test_cone_closed_form():
    for p in range(9):
        offs = reach(p)
        lo = tuple(min(o[d] for o in offs) for d in range(2))
        hi = tuple(max(o[d] for o in offs) for d in range(2))
        c0 = p // 3
        assert lo == (-c0, -p), p
        assert hi == (2 * c0, 2 * p), p


def test_kernel_halo_covers_cone():
    """Halo = widest cone + 2 (the 3x3 neighbourhood of an owned cell, and
    the neighbourhood each of their processings reads)."""
    src = open(SRC).read()
    vals = dict(re.findall(r"(kTileHalo\w+) = (\d+)", src))
    lo0 = max(-min(o[0] for o in reach(p)) for p in range(9)) + 2
    hi0 = max(max(o[0] for o in reach(p)) for p in range(9)) + 2
    lo1 = max(-min(o[1] for o in reach(p)) for p in range(9)) + 2
    hi1 = max(max(o[1] for o in reach(p)) for p in range(9)) + 2
    assert (int(vals["kTileHalo0Lo"]), int(vals["kTileHalo0Hi"]),
            int(vals["kTileHalo1Lo"]), int(vals["kTileHalo1Hi"])) == (lo0, hi0, lo1, hi1)
