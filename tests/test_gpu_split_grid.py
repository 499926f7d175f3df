"""FP32-split form on grids other than the standard staggering (runs on an MI355X).

The README builds its velocity grids with ``expand_range``, which rounds the
ghost ends to 5 significant digits (README.md:48-55), so the centre axes of
Vx and Vy are uniform but shifted and stretched against the standard
staggering (by up to 0.012 cell at 256^2).  The split advection derives each
velocity axis's affine map from its actual coordinates; every split entry
point rejects non-uniform axes in C, before any kernel changes a particle.
"""
import ctypes as C

import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip import _abi
from justpic_hip.configs import stream2d
from oracle import oracle as O

from helpers import host_particles
from test_gpu_split import POS_TOL, compare_by_id, dev_V, split_state, to_dev

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-split-grid]", *a, flush=True)


def readme_case(n, npq=6, K=48, seed=51):
    cfg = stream2d(n, readme=True)
    hp = host_particles(cfg["xv"], cfg["xi_vel"], K, npq, seed=seed)
    return cfg, hp


def test_readme_grid_fp64_bit_exact(ctx):
    """The FP64 drop-in on the README grid: bisection on the actual axes, bit
    for bit against the oracle (advection + move, 6 steps)."""
    cfg, hp = readme_case(64)
    say("readme fp64 64^2")
    p = to_dev(ctx, hp)
    Vd = dev_V(ctx, cfg["V"])
    T = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    for _ in range(6):
        jp.advection_(p, jp.RungeKutta2(), Vd, cfg["dt"])
        jp.move_particles_(p, (Td,))
        O.advection(hp, cfg["V"], O.JP_RK2, 0.5, cfg["dt"])
        O.move_particles(hp, [T])
    coords, index = p.host_state()
    assert np.array_equal(index, hp.index)
    for d in range(2):
        assert np.array_equal(coords[d], hp.coords[d], equal_nan=True)
    assert np.array_equal(Td.data.flat(), T, equal_nan=True)


def test_readme_grid_split_vs_fp64_oracle(ctx):
    """Split advection + move on the README grid, 60 RK2 steps at 256^2
    (12 particles per cell), against the FP64 oracle: positions within 1e-6
    of the domain, cell-assignment disagreements within 1e-4 of the
    particles.  Reading the ghost axes as the standard staggering (0.012 cell
    off at 256^2) would move particles by up to ~6e-7 of the domain per step
    in this flow."""
    cfg, hp = readme_case(256, npq=3)
    say("readme split 256^2")
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    s = jp.to_split(to_dev(ctx, hp))
    idd = jp.CellArray(ctx, hp.ni, hp.K)
    idd.copy_from(ids)
    Vd = dev_V(ctx, cfg["V"])
    for _ in range(60):
        jp.advection_(s, jp.RungeKutta2(), Vd, cfg["dt"])
        jp.move_particles_(s, (idd,))
        O.advection(hp, cfg["V"], O.JP_RK2, 0.5, cfg["dt"])
        O.move_particles(hp, [ids])
    coords, index = split_state(s)
    err, dis, n = compare_by_id((coords, index, idd.data.flat()), (hp.coords, hp.index, ids),
                                hp.C)
    say(f"readme 256^2: {n} particles, 60 steps, max err {err:.3e}, disagreements {dis}")
    assert err <= POS_TOL
    assert dis <= 1e-4 * n


def id_errors(a, b, C):
    """per-particle position errors over the ids both states hold, the counts
    of ids held by one state only, and of common ids in different cells"""
    from test_gpu_split import by_id
    ia, ca, pa = by_id(*a, C)
    ib, cb, pb = by_id(*b, C)
    common, ja, jb = np.intersect1d(ia, ib, assume_unique=True, return_indices=True)
    err = np.max(np.abs(pa[ja] - pb[jb]), axis=1)
    return err, len(ia) - common.size, len(ib) - common.size, int(np.count_nonzero(ca[ja] != cb[jb]))


def test_readme_config_full_size(ctx):
    """The README rotating cell at its own size: 256^2 cells, nxcell 24,
    max_xcell 48, RK2, 750 steps of advection! + move_particles!, no fields,
    split form against the FP64 drop-in (bit-exact with the oracle at every
    size the oracle runs).

    Over 750 steps the two forms part ways where the flow itself separates
    trajectories: near the hyperbolic stagnation points of this cellular flow
    a position difference grows like exp(250 pi t), about 1e4 over these 750
    steps, and which particle a full cell drops depends on arrival order, so
    the two populations differ by drops.  Checked: (a) 99% of the common
    particles within 1e-6 of the domain and all within 1e-4; (b) populations
    within 1% of each other; (c) the per-step bar: one more step from the
    FP64 state converted to the split form puts every particle within 1e-6
    of the domain of the FP64 step, with at most 1e-4 of them in another
    cell."""
    cfg = stream2d(256, readme=True)
    say("readme full size 256^2, 750 steps")
    p = jp.init_particles(ctx, 24, 48, 12, *cfg["xi_vel"], seed=7)
    s = jp.to_split(p)
    ids = jp.CellArray(ctx, p.ni, p.max_xcell)
    _, idx = p.host_state()
    ids.copy_from(np.where(idx == 1, np.arange(idx.size, dtype=np.float64), np.nan))
    ids_s = ids.copy()
    Vd = dev_V(ctx, cfg["V"])
    n0 = p.active()
    for _ in range(750):
        jp.advection_(p, jp.RungeKutta2(), Vd, cfg["dt"])
        jp.move_particles_(p, (ids,))
        jp.advection_(s, jp.RungeKutta2(), Vd, cfg["dt"])
        jp.move_particles_(s, (ids_s,))
    c64, i64 = p.host_state()
    cs, is_ = split_state(s)
    err, only_s, only_64, cells = id_errors((cs, is_, ids_s.data.flat()),
                                            (c64, i64, ids.data.flat()), p.C)
    n64, ns = int(np.count_nonzero(i64)), int(np.count_nonzero(is_))
    q = np.quantile(err, [0.5, 0.99, 0.999])
    say(f"750 steps: {n0} -> fp64 {n64}, split {ns}; common {err.size}, only split {only_s}, "
        f"only fp64 {only_64}, other cell {cells}; error p50 {q[0]:.2e} p99 {q[1]:.2e} "
        f"p99.9 {q[2]:.2e} max {err.max():.2e}")
    assert q[1] <= POS_TOL
    assert err.max() <= 1e-4
    assert abs(ns - n64) <= 1e-2 * n64
    # (c) one step from the same state
    s1 = jp.to_split(p)
    ids_1 = ids.copy()
    jp.advection_(p, jp.RungeKutta2(), Vd, cfg["dt"])
    st64 = jp.move_particles_(p, (ids,), stats=True)
    jp.advection_(s1, jp.RungeKutta2(), Vd, cfg["dt"])
    st1 = jp.move_particles_(s1, (ids_1,), stats=True)
    c64, i64 = p.host_state()
    cs, is_ = split_state(s1)
    err, only_s, only_64, cells = id_errors((cs, is_, ids_1.data.flat()),
                                            (c64, i64, ids.data.flat()), p.C)
    dis = only_s + only_64 + cells
    say(f"one step: moved {st64['moved']} / {st1['moved']}, dropped {st64['dropped']} / "
        f"{st1['dropped']}, max error {err.max():.2e}, disagreements {dis} of {err.size}")
    assert err.max() <= POS_TOL
    assert dis <= 1e-4 * err.size


# ------------------------------------------------------------------ rejection in C
def _dev_axis(ctx, x, keep):
    """jp_axis over device copies of x (dx, inv_dx from diff) and the buffers
    that must stay alive."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    dx = np.diff(x)
    bufs = [jp.DeviceArray.from_host(ctx, a) for a in (x, dx, 1.0 / dx)]
    keep.extend(bufs)
    return _abi.jp_axis(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, x.size)


def _state(s):
    return [c.data.flat() for c in s.coords], s.index.data.flat()


def _same(a, b):
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a[0], b[0])) and \
        np.array_equal(a[1], b[1])


def test_split_entry_points_reject_nonuniform_axes_in_c(ctx):
    """Calls through the C-ABI with a vertex axis or a velocity axis moved off
    uniform spacing by 1e-6 of a cell: every split entry point that depends on
    uniform axes returns JP_ERR_ARG and changes nothing; synchronous mode
    reports it at once, asynchronous mode at the next jp_sync."""
    cfg = stream2d(32, readme=True)
    p = jp.init_particles(ctx, 8, 16, 6, *cfg["xi_vel"], seed=3)
    s = jp.to_split(p)
    Vd = dev_V(ctx, cfg["V"])
    F = jp.nodal_field(s)
    T = jp.CellArray(ctx, s.ni, s.max_xcell, fill=1.0)
    lib, h = ctx.lib, ctx.h
    keep = []
    # a vertex axis with one interior point moved by 1e-6 of a cell
    xb = cfg["xv"][0].x.copy()
    xb[7] += 1e-6 * (xb[1] - xb[0])
    gbad = _abi.jp_cellgrid.from_buffer_copy(s._grid)
    gbad.xv[0] = _dev_axis(ctx, xb, keep)
    # a velocity axis (the expanded centre axis of Vx along y) likewise
    vbad = _abi.jp_velgrid.from_buffer_copy(s._velgrid)
    yb = cfg["xi_vel"][0][1].x.copy()
    yb[5] += 1e-6 * (yb[1] - yb[0])
    vbad.ax[0][1] = _dev_axis(ctx, yb, keep)
    Vp = _abi.ptr_array([v.ptr for v in Vd])
    args = _abi.ptr_array([T.ptr])
    before = _state(s)

    calls = {
        "advection (vertex)": lambda g, vg: lib.jp_advection_split(
            h, C.byref(g), s.coord_ptrs(), s.index.ptr, Vp, C.byref(vg), _abi.JP_RK2, 0.5,
            cfg["dt"]),
        "move": lambda g, vg: lib.jp_move_particles_split(h, C.byref(g), s.coord_ptrs(),
                                                         s.index.ptr, 1, args, None),
        "particle2grid": lambda g, vg: lib.jp_particle2grid_split(h, C.byref(g), F.ptr, T.ptr,
                                                                 s.coord_ptrs(), s.index.ptr),
        "inject": lambda g, vg: lib.jp_inject_particles_split(h, C.byref(g), s.coord_ptrs(),
                                                             s.index.ptr, 1, args, 6, 1, 0,
                                                             None),
    }
    F.fill(-7.0)
    for name, call in calls.items():
        rc = call(gbad, s._velgrid)
        msg = lib.jp_last_error(h).decode()
        say(name, rc, msg)
        assert rc == _abi.JP_ERR_ARG, name
        assert "uniform vertex" in msg
        assert _same(_state(s), before), name
    assert np.all(F.flat() == -7.0), "particle2grid wrote F for a rejected grid"
    # velocity axis off uniform: the advection rejects it
    g_ok = s._grid
    rc = calls["advection (vertex)"](g_ok, vbad)
    assert rc == _abi.JP_ERR_ARG and "uniform velocity" in lib.jp_last_error(h).decode()
    assert _same(_state(s), before)
    # asynchronous: axes are checked (and read back) on first use, so the
    # rejection is immediate there too
    ctx.set_sync(False)
    try:
        rc = calls["advection (vertex)"](g_ok, vbad)
        assert rc == _abi.JP_ERR_ARG and "uniform velocity" in lib.jp_last_error(h).decode()
        assert lib.jp_sync(h) == 0
    finally:
        ctx.set_sync(True)
    assert _same(_state(s), before)
    # under graph capture the verdict stays on the device: the captured call
    # returns 0, its kernels skip all work, and jp_sync after the replay
    # reports the rejection once
    ctx.graph_begin()
    rc = calls["advection (vertex)"](g_ok, vbad)
    graph = ctx.graph_end()
    assert rc == 0
    ctx.graph_launch(graph)
    rc = lib.jp_sync(h)
    assert rc == _abi.JP_ERR_ARG and "uniform velocity" in lib.jp_last_error(h).decode()
    assert lib.jp_sync(h) == 0, "the report clears the sticky word"
    ctx.graph_destroy(graph)
    assert _same(_state(s), before)
    # the good grid still advects
    jp.advection_(s, jp.RungeKutta2(), Vd, cfg["dt"])
    assert not _same(_state(s), before)


def test_geometry_verdict_cached_until_free_or_recheck(ctx):
    """Axes are checked once per set of pointers: a vertex axis made
    non-uniform in place afterwards is still trusted (the documented cost of
    the cache), JP_OPT_GEOMETRY_RECHECK = 1 checks every call and rejects it,
    and any jp_free drops every cached verdict."""
    cfg = stream2d(32)
    p = jp.init_particles(ctx, 8, 16, 6, *cfg["xi_vel"], seed=4)
    s = jp.to_split(p)
    Vd = dev_V(ctx, cfg["V"])
    lib, h = ctx.lib, ctx.h
    keep = []
    x = cfg["xv"][0].x.copy()
    g = _abi.jp_cellgrid.from_buffer_copy(s._grid)
    g.xv[0] = _dev_axis(ctx, x, keep)
    Vp = _abi.ptr_array([v.ptr for v in Vd])

    def advect():
        return lib.jp_advection_split(h, C.byref(g), s.coord_ptrs(), s.index.ptr, Vp,
                                      C.byref(s._velgrid), _abi.JP_RK2, 0.5, cfg["dt"])

    assert advect() == 0
    xb = x.copy()
    xb[7] += 1e-6 * (xb[1] - xb[0])
    keep[0].copy_from(xb)  # in place: same pointer, other coordinates
    assert advect() == 0  # cached verdict
    ctx.set_option(_abi.JP_OPT_GEOMETRY_RECHECK, 1)
    try:
        before = _state(s)
        assert advect() == _abi.JP_ERR_ARG
        assert "uniform vertex" in lib.jp_last_error(h).decode()
        assert _same(_state(s), before)
    finally:
        ctx.set_option(_abi.JP_OPT_GEOMETRY_RECHECK, 0)
    keep[0].copy_from(x)
    assert advect() == 0
    keep[0].copy_from(xb)
    jp.DeviceArray(ctx, (4,)).free()  # any jp_free drops the cached verdicts
    assert advect() == _abi.JP_ERR_ARG
