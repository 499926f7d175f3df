"""Launch shapes that only large problems select, forced on small ones and
compared with the oracle (runs on an MI355X).

The library sizes some launches from the problem: the move's classify and
scatter and the split advection split each cell's slots into groups
(blockIdx.y) until the launch fills the chip, and the injection runs
persistent wavefronts over each colour's list.  At the test sizes every
launch takes the smallest shape (one 8-slot group, one wavefront per listed
cell); the large workloads take others (2D 512^2 split advection: 24-slot
groups; 3D 128^3 move with K = 54: whole cells in one multi-chunk group).
JP_OPT_SLOT_SPAN and JP_OPT_INJECT_WAVES force those shapes here.
"""
import contextlib

import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip import _abi
from justpic_hip.configs import gaussian_T
from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V
from test_gpu_parity import assert_field, assert_state_equal, dev_V, to_dev
from test_gpu_split import POS_TOL, compare_by_id, dyadic_state, split_state

pytestmark = pytest.mark.gpu

SPANS = [8, 16, 24, 64]  # 64 >= K: whole cells, several 8-slot chunks per thread


def say(*a):
    print("[gpu-shapes]", *a, flush=True)


@contextlib.contextmanager
def option(ctx, opt, value):
    ctx.set_option(opt, value)
    try:
        yield
    finally:
        ctx.set_option(opt, 0)


def move_case(D):
    if D == 2:  # the 2D workloads' K = 48, 24 particles per cell
        xv, xi = grid2d(40, 36)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 48, 6, seed=17)
    else:  # the 3D workload's K = 54, 32 particles per cell
        xv, xi = grid3d(10)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 54, 4, seed=17)
    return xv, xi, V, hp


@pytest.mark.parametrize("span", SPANS)
@pytest.mark.parametrize("D", [2, 3])
def test_move_slot_spans_fp64(ctx, D, span):
    """FP64 advection + move with forced slot groups: masks, positions, both
    fields and outcome counts bit for bit against the oracle (2 steps)."""
    say("fp64 move span", D, span)
    xv, xi, V, hp = move_case(D)
    rng = np.random.default_rng(3)
    T = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    p = to_dev(ctx, hp)
    Td, Id = jp.CellArray(ctx, hp.ni, hp.K), jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    Id.copy_from(ids)
    Vd = dev_V(ctx, V)
    dt = cfl_dt(xv, V)
    with option(ctx, _abi.JP_OPT_SLOT_SPAN, span):
        for it in range(2):
            jp.advection_(p, jp.RungeKutta2(), Vd, dt)
            O.advection(hp, V, O.JP_RK2, 0.5, dt)
            st_d = jp.move_particles_(p, (Td, Id), stats=True)
            st_o = O.move_particles(hp, [T, ids])
            say("stats", it, st_d, st_o)
            assert st_o["moved"] > 0
            assert st_d == st_o
            assert_state_equal(p, hp, exact=True)
            assert_field(Td.data.flat(), T, exact=True)
            assert_field(Id.data.flat(), ids, exact=True)


@pytest.mark.parametrize("span", SPANS)
@pytest.mark.parametrize("D", [2, 3])
def test_move_slot_spans_split(ctx, D, span):
    """Split move with forced slot groups on exactly representable states:
    bit for bit against the FP64 oracle (K = 48 in 2D, 54 in 3D)."""
    say("split move span", D, span)
    xv, xi, hp = dyadic_state(D, 16 if D == 2 else 8, 48 if D == 2 else 54, seed=31 + D)
    T = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    s = jp.to_split(to_dev(ctx, hp))
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    with option(ctx, _abi.JP_OPT_SLOT_SPAN, span):
        st_d = jp.move_particles_(s, (Td,), stats=True)
    st_o = O.move_particles(hp, [T])
    say("stats", st_d, st_o)
    assert st_o["moved"] > 0 and st_o["dropped"] > 0
    assert st_d == st_o
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    for d in range(D):
        assert np.array_equal(coords[d], hp.coords[d], equal_nan=True)
    assert np.array_equal(Td.data.flat(), T, equal_nan=True)


@pytest.mark.parametrize("D", [2, 3])
def test_split_advection_slot_spans(ctx, D):
    """Split advection with every slot-group size gives the same offsets bit
    for bit (a particle's arithmetic does not depend on its group), within
    1e-6 of the domain of the FP64 oracle after one RK2 step."""
    say("split advection spans", D)
    xv, xi, V, hp = move_case(D)
    Vd = dev_V(ctx, V)
    dt = cfl_dt(xv, V)
    base = jp.to_split(to_dev(ctx, hp))
    ref = None
    for span in [0] + SPANS:
        s = jp.to_split(jp.to_f64(base))
        with option(ctx, _abi.JP_OPT_SLOT_SPAN, span):
            jp.advection_(s, jp.RungeKutta2(), Vd, dt)
        u = [c.data.flat() for c in s.off]
        if ref is None:
            ref = u
        else:
            for a, b in zip(u, ref):
                assert np.array_equal(a, b, equal_nan=True), f"span {span}"
    O.advection(hp, V, O.JP_RK2, 0.5, dt)
    coords, index = split_state(s)
    m = index == 1
    # offsets stay with their original cell until the move: compare absolute
    err = max(float(np.max(np.abs(coords[d][m] - hp.coords[d][m]))) for d in range(D))
    say("one-step error", err)
    assert err <= POS_TOL


@pytest.mark.parametrize("waves", [1, 3, 7])
@pytest.mark.parametrize("D", [2, 3])
def test_inject_persistent_waves(ctx, D, waves):
    """Injection with a forced number of persistent wavefronts per colour (so
    that each wavefront walks several listed cells in turn): FP64 bit for bit
    against the oracle, and the split form with the same masks and copies."""
    say("inject waves", D, waves)
    xv, xi, hp = dyadic_state(D, 16 if D == 2 else 8, 16, seed=41 + D)
    O.move_particles(hp, [])
    rng = np.random.default_rng(6)
    T = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    minx = 10 if D == 2 else 12
    p = to_dev(ctx, hp, min_xcell=minx, seed=5)
    s = jp.to_split(p)
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(T)
    Ts = Td.copy()
    with option(ctx, _abi.JP_OPT_INJECT_WAVES, waves):
        st_d = jp.inject_particles_(p, (Td,), stats=True)
        st_s = jp.inject_particles_(s, (Ts,), stats=True)
    st_o = O.inject_particles(hp, [T], minx, seed=5, counter=0)
    say("stats", st_d, st_s, st_o)
    assert st_o["injected"] > waves
    assert st_d == st_o and st_s == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(Td.data.flat(), T, exact=True)
    coords, index = split_state(s)
    assert np.array_equal(index, hp.index)
    assert np.array_equal(Ts.data.flat(), T, equal_nan=True)


def cell_sets_differ(a, b, C, K, tol):
    """Order-free comparison of two states (coords, index) cell by cell: the
    particles of each cell sorted by x; a cell differs when its particle counts
    differ or a sorted position is more than tol apart.  Returns (number of
    differing cells, max position error over the agreeing cells)."""
    def sorted_cells(coords, index):
        act = index.reshape(K, C).T == 1
        xs = [np.where(act, c.reshape(K, C).T, np.inf) for c in coords]
        order = np.argsort(xs[0], axis=1, kind="stable")
        return act.sum(axis=1), [np.take_along_axis(x, order, axis=1) for x in xs]
    na, xa = sorted_cells(*a)
    nb, xb = sorted_cells(*b)
    dev = np.zeros((C, K))
    for u, v in zip(xa, xb):
        fin = np.isfinite(u) & np.isfinite(v)
        dev = np.maximum(dev, np.abs(np.subtract(u, v, out=np.zeros((C, K)), where=fin)))
        dev = np.where(np.isfinite(u) != np.isfinite(v), np.inf, dev)
    worst = dev.max(axis=1)
    bad = (na != nb) | (worst > tol)
    return int(np.count_nonzero(bad)), float(worst[~bad].max(initial=0.0))


def test_split_loop_with_inject_and_p2g(ctx):
    """The 512^2 workload's per-cell shape on 128^2 cells (nxcell 24, max 48,
    min 12, T from grid2particle!), 10 steps of RK2 advection + move + inject
    + particle2grid in the split form against the FP64 oracle.  Injection
    copies fields, so particles have no unique id; cells are compared as
    sets of positions instead.  A particle that the float32 offsets place on
    the other side of a cell edge than the FP64 path does (expected rate
    ~1e-7 per particle step) changes its cell's count and with it that
    cell's injections, slot order and the injected positions; those cells
    are counted, and must stay below 1e-3 of the cells (measured: a handful
    of 16384).  Every other cell agrees within 1e-6 of the domain."""
    n, steps = 128, 10
    xv, xi = grid2d(n)
    V = stream_V(xi)
    hp = host_particles(xv, xi, 48, 6, seed=23)
    dt = cfl_dt(xv, V)
    Tn = gaussian_T(xv)
    Tp = O.grid2particle(hp, Tn, np.full(hp.S, np.nan))
    Tp = np.where(hp.index == 1, Tp, np.nan)
    s = jp.to_split(to_dev(ctx, hp, nxcell=24, min_xcell=12, seed=9))
    Td = jp.CellArray(ctx, hp.ni, hp.K)
    Td.copy_from(Tp)
    Vd = dev_V(ctx, V)
    F = jp.nodal_field(s)
    for it in range(steps):
        jp.advection_(s, jp.RungeKutta2(), Vd, dt)
        jp.move_particles_(s, (Td,))
        jp.inject_particles_(s, (Td,))
        jp.particle2grid_(F, Td, s)
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        O.move_particles(hp, [Tp])
        O.inject_particles(hp, [Tp], 12, seed=9, counter=it)
        Fo = O.particle2grid(hp, Tp)
    coords, index = split_state(s)
    n_act = int(np.count_nonzero(hp.index))
    cells, err = cell_sets_differ((coords, index), (hp.coords, hp.index), hp.C, hp.K, POS_TOL)
    dF = np.abs(F.flat() - Fo)
    say(f"split loop 128^2 x {steps}: {n_act} particles, max err {err:.3e}, differing cells "
        f"{cells} of {hp.C}, |dF| max {np.nanmax(dF):.3e} p99.9 {np.nanquantile(dF, 0.999):.3e}")
    assert err <= POS_TOL
    assert cells <= 1e-3 * hp.C
    assert np.array_equal(np.isnan(F.flat()), np.isnan(Fo))
    # float32 weights (relative error ~1e-7 of each weight) on values in [0, 1]
    assert np.nanquantile(dF, 0.999) <= 1e-5

