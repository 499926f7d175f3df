"""The allocation-free particle2grid entry point against the allocating one
(runs on an MI355X).

jp_particle2grid_noalloc keeps its occupancy words in a fixed device array
instead of the context's scratch; its result must equal jp_particle2grid's
bit for bit.  Two calls in a row, the second on a smaller grid (the array
keeps the first call's words, which the second must not read), then a grid
over the size limit, which is rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip import _abi

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)


def case(D, seed):
    if D == 2:
        xv, xi = grid2d(96, 80)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 48, 6, seed=seed)
    else:
        xv, xi = grid3d(12)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 54, 4, seed=seed)
    return xv, xi, V, hp


def both(ctx, D, seed):
    """F from jp_particle2grid and from jp_particle2grid_noalloc on the same
    moved particles (one advection + move first, so cells are uneven)."""
    xv, xi, V, hp = case(D, seed)
    p = jp.particles_from_host(ctx, xi, hp.coords, hp.index, hp.K, hp.K, 1)
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in V)
    jp.advection_(p, jp.RungeKutta2(), Vd, cfl_dt(xv, V))
    jp.move_particles_(p, ())
    rng = np.random.default_rng(seed)
    Fp = jp.CellArray(ctx, p.ni, p.max_xcell)
    Fp.copy_from(rng.random(p.C * p.max_xcell))
    F1 = jp.nodal_field(p, fill=-1.0)
    F2 = jp.nodal_field(p, fill=-2.0)
    lib, h = ctx.lib, ctx.h
    ctx.check(lib.jp_particle2grid(h, p.grid_ref(), F1.ptr, Fp.ptr, p.coord_ptrs(), p.index.ptr),
              "particle2grid")
    ctx.check(lib.jp_particle2grid_noalloc(h, p.grid_ref(), F2.ptr, Fp.ptr, p.coord_ptrs(),
                                           p.index.ptr), "particle2grid_noalloc")
    return p, F1.flat(), F2.flat()


def test_noalloc_p2g_matches_allocating_entry_point(ctx):
    say("noalloc 2D 96x80 then 3D 12^3 (smaller)")
    p1, a1, b1 = both(ctx, 2, 5)
    p2, a2, b2 = both(ctx, 3, 6)
    assert p2.C < p1.C
    for a, b in ((a1, b1), (a2, b2)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a, b, equal_nan=True)
        assert np.count_nonzero(~np.isnan(a)) > 0.9 * a.size


def test_noalloc_rejects_grids_over_the_limit(ctx):
    """A grid of JP_NOALLOC_MAX_CELLS + 1 cells is refused with JP_ERR_ARG by
    the host check (no kernel runs; the axis pointers are never read)."""
    say("noalloc size limit")
    xv, xi = grid2d(4)
    p = jp.init_particles(ctx, 4, 4, 1, *xi, seed=1)
    g = _abi.jp_cellgrid.from_buffer_copy(p._grid)
    g.ni[0] = _abi.JP_NOALLOC_MAX_CELLS + 1
    g.ni[1] = 1
    g.xv[0].n = g.ni[0] + 1
    g.xv[1].n = 2
    F = jp.nodal_field(p)
    rc = ctx.lib.jp_particle2grid_noalloc(ctx.h, C.byref(g), F.ptr, F.ptr, p.coord_ptrs(),
                                          p.index.ptr)
    assert rc == _abi.JP_ERR_ARG
    assert "2097152" in ctx.lib.jp_last_error(ctx.h).decode()
