"""GPU checks of the batched ensemble (runs on an MI355X).

The ensemble advects many grids in one persistent launch with the same
per-particle arithmetic as the single-grid FP32-split advection, so every
launch variant must reproduce per-grid advection bit for bit; a captured
time step must reproduce the same step run eagerly, on every replay.
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.configs import gaussian_T

pytestmark = pytest.mark.gpu

SIZES = [(24, 20), (40, 40), (65, 33), (17, 70)]


def say(*a):
    print("[gpu-ensemble]", *a, flush=True)


def member(ctx, nx, ny, amp, seed):
    from justpic_hip.grids import Axis, staggered_grid
    from justpic_hip.configs import sample
    xv = (Axis.lin(0.0, 1.0, nx + 1), Axis.lin(0.0, 1.0, ny + 1))
    xi = staggered_grid(xv)
    Vx = sample(lambda x, y: amp * np.sin(np.pi * x) * np.cos(np.pi * y), xi[0])
    Vy = sample(lambda x, y: -amp * np.cos(np.pi * x) * np.sin(np.pi * y), xi[1])
    dt = min(1.0 / nx / np.max(np.abs(Vx)), 1.0 / ny / np.max(np.abs(Vy)))
    p = jp.init_particles(ctx, 24, 48, 12, *xi, seed=seed, form="split")
    V = (jp.DeviceArray.from_host(ctx, Vx), jp.DeviceArray.from_host(ctx, Vy))
    T = jp.DeviceArray.from_host(ctx, gaussian_T(xv))
    f = jp.CellArray(ctx, p.ni, p.max_xcell, fill=float("nan"))
    jp.grid2particle_(f, T, p)
    return p, V, dt, f


def build(ctx, seed=0):
    ms = [member(ctx, nx, ny, 100.0 + 60 * i, seed + i) for i, (nx, ny) in enumerate(SIZES)]
    return ms


def state(p):
    return [o.data.flat() for o in p.off], p.index.data.flat()


@pytest.mark.parametrize("variant", range(8))
def test_ensemble_advect_matches_single(ctx, variant):
    say("variant", variant)
    a = build(ctx)
    b = build(ctx)
    ens = jp.Ensemble(ctx, [(p, V, dt) for p, V, dt, _ in a])
    ens.set_variant(variant)
    for _ in range(2):
        ens.advect()
        for p, V, dt, _ in b:
            jp.advection_(p, jp.RungeKutta2(), V, dt)
    for (pa, *_), (pb, *_) in zip(a, b):
        oa, ia = state(pa)
        ob, ib = state(pb)
        assert np.array_equal(ia, ib)
        for d in range(2):
            assert np.array_equal(oa[d], ob[d], equal_nan=True)
    ens.close()


def test_ensemble_tune_keeps_positions(ctx):
    a = build(ctx, seed=5)
    ens = jp.Ensemble(ctx, [(p, V, dt) for p, V, dt, _ in a])
    before = [state(p) for p, *_ in a]
    best, ms = ens.tune(reps=2)
    say("tune", best, [round(t, 4) for t in ms])
    assert 0 <= best < 8
    assert all(t > 0 for t in ms)
    for (p, *_), (o, i) in zip(a, before):
        o2, i2 = state(p)
        assert np.array_equal(i, i2)
        for d in range(2):
            assert np.array_equal(o[d], o2[d], equal_nan=True)
    ens.close()


def test_ensemble_graph_replays_equal_eager_steps(ctx):
    a = build(ctx, seed=9)
    b = build(ctx, seed=9)
    ea = jp.Ensemble(ctx, [(p, V, dt) for p, V, dt, _ in a])
    eb = jp.Ensemble(ctx, [(p, V, dt) for p, V, dt, _ in b])
    fa = [f for *_, f in a]
    fb = [f for *_, f in b]
    Fa = [jp.nodal_field(p) for p, *_ in a]
    Fb = [jp.nodal_field(p) for p, *_ in b]
    # one eager step each (sizes the context scratch), then capture a's step
    ea.step(fa, Fa)
    eb.step(fb, Fb)
    g = ea.capture_step(fa, Fa)
    for r in range(3):
        ctx.graph_launch(g)
        ctx.sync()
        eb.step(fb, Fb)
        ctx.sync()
        for (pa, *_), (pb, *_), x, y, X, Y in zip(a, b, fa, fb, Fa, Fb):
            oa, ia = state(pa)
            ob, ib = state(pb)
            assert np.array_equal(ia, ib), f"replay {r}: mask"
            for d in range(2):
                assert np.array_equal(oa[d], ob[d], equal_nan=True), f"replay {r}: offsets"
            assert np.array_equal(x.data.flat(), y.data.flat(), equal_nan=True), f"replay {r}"
            assert np.array_equal(X.flat(), Y.flat(), equal_nan=True), f"replay {r}: p2g"
    ctx.graph_destroy(g)
    ea.close()
    eb.close()


def test_ensemble_rejects_fp64_members(ctx):
    p, V, dt, _ = member(ctx, 16, 16, 100.0, 1)
    q = jp.to_f64(p)
    with pytest.raises(ValueError):
        jp.Ensemble(ctx, [(q, V, dt)])


def test_ensemble_step_matches_per_grid_calls(ctx):
    """The step fans the members' calls out over sub-context streams; every
    member must end exactly as with plain per-grid calls on one context."""
    say("fan-out vs per-grid calls")
    a = build(ctx, seed=13)
    b = build(ctx, seed=13)
    ea = jp.Ensemble(ctx, [(p, V, dt) for p, V, dt, _ in a], streams=3)
    eb = jp.Ensemble(ctx, [(p, V, dt) for p, V, dt, _ in b])
    Fa = [jp.nodal_field(p) for p, *_ in a]
    Fb = [jp.nodal_field(p) for p, *_ in b]
    for _ in range(3):
        ea.step([f for *_, f in a], Fa)
        eb.advect()
        for (p, _, _, f), F in zip(b, Fb):
            jp.move_particles_(p, (f,))
            jp.particle2grid_(F, f, p)
    ctx.sync()
    for (pa, *_, x), (pb, *_, y), X, Y in zip(a, b, Fa, Fb):
        oa, ia = state(pa)
        ob, ib = state(pb)
        assert np.array_equal(ia, ib)
        for d in range(2):
            assert np.array_equal(oa[d], ob[d], equal_nan=True)
        assert np.array_equal(x.data.flat(), y.data.flat(), equal_nan=True)
        assert np.array_equal(X.flat(), Y.flat(), equal_nan=True)
    ea.close()
    eb.close()
