"""advection_MQS!, advection_LinP! and semilagrangian_advection! (linear,
MQS, LinP) on the GPU against the oracle, bit for bit (runs on an MI355X).

Cases: 2D uniform and refined (vector) grids, the README ghost axes, 3D;
Euler, RK2 (alpha 0.5 and 2/3) and RK4; scalar and tuple fields (the
tuple linear method's spacing at the node); the reference's refined-grid
test values; a backtrack leaving the grid reported as an error, at once on
a synchronous context and at the next sync on an asynchronous one.
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.configs import stream2d
from justpic_hip.grids import Axis, compute_dx, expand_range, staggered_grid
from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, sample, stream_V
from test_gpu_parity import assert_state_equal, dev_V, to_dev

pytestmark = pytest.mark.gpu

SCHEMES = ("MQS", "LinP")
ADVECT = {"MQS": jp.advection_MQS_, "LinP": jp.advection_LinP_}
SML = {"linear": jp.semilagrangian_advection_, "MQS": jp.semilagrangian_advection_MQS_,
       "LinP": jp.semilagrangian_advection_LinP_}
METHODS = [(jp.Euler(), O.JP_EULER, 0.5), (jp.RungeKutta2(), O.JP_RK2, 0.5),
           (jp.RungeKutta2(2 / 3), O.JP_RK2, 2 / 3), (jp.RungeKutta4(), O.JP_RK4, 0.5)]


def say(*a):
    print("[gpu-schemes]", *a, flush=True)


def adv_case(kind):
    if kind == "3d":
        xv, xi = grid3d(8)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 40, 3, seed=12)
    elif kind == "readme":
        cfg = stream2d(48, readme=True)
        xv, xi, V = cfg["xv"], cfg["xi_vel"], cfg["V"]
        hp = host_particles(xv, xi, 24, 4, seed=12)
    else:
        xv, xi = grid2d(26, 21, kind=kind)
        V = stream_V(xi, 3.0 if kind == "vector" else 250.0)
        hp = host_particles(xv, xi, 24, 4, seed=12)
    return xv, xi, V, hp


@pytest.mark.parametrize("mi", range(len(METHODS)))
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("kind", ["range", "vector", "readme", "3d"])
def test_advection_scheme_matches_oracle(ctx, kind, scheme, mi):
    method, code, alpha = METHODS[mi]
    say("advect", kind, scheme, mi)
    xv, xi, V, hp = adv_case(kind)
    O.move_particles(hp, [])  # holes in the slot planes
    p = to_dev(ctx, hp)
    Vd = dev_V(ctx, V)
    dt = 0.8 * cfl_dt(xv, V)
    for _ in range(2):
        ADVECT[scheme](p, method, Vd, dt)
        O.advection_scheme(hp, V, code, alpha, dt, scheme)
        assert_state_equal(p, hp, exact=True)
        jp.move_particles_(p)
        O.move_particles(hp, [])


def test_advection_scheme_rejects_split(ctx):
    xv, xi, V, hp = adv_case("range")
    s = jp.to_split(to_dev(ctx, hp))
    with pytest.raises(ValueError):
        jp.advection_MQS_(s, jp.RungeKutta2(), dev_V(ctx, V), 1e-3)


def sml_case(D, kind="range", n=12):
    if D == 2 and kind == "vector":
        xv, _ = grid2d(n, n + 3, kind="vector")
    elif D == 2:
        xv = (Axis.lin(0.0, 1.0, n + 1), Axis.lin(0.0, 2.0, n + 4))
    else:
        xv = tuple(Axis.lin(0.0, 1.0, n + 1) for _ in range(3))
    xi = staggered_grid(xv)
    V = stream_V(xi, 1.0) if D == 2 else cellular_V(xi, 1.0)
    h = min(np.min(a.dx) for a in xv)
    dt = 0.7 * h / max(np.abs(v).max() for v in V)
    return xv, xi, V, dt


def smooth(axes, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(sample(lambda *X: np.sin(3 * X[0] + seed) * np.cos(2 * X[-1]), axes)
                             + 0.01 * rng.random(tuple(a.n for a in axes)))


@pytest.mark.parametrize("nf", [1, 2])
@pytest.mark.parametrize("method", ["rk2", "rk2a", "rk4"])
@pytest.mark.parametrize("scheme", ["linear", "MQS", "LinP"])
@pytest.mark.parametrize("D,kind", [(2, "range"), (2, "vector"), (3, "range")])
def test_semilagrangian_matches_oracle(ctx, D, kind, scheme, method, nf):
    say("sml", D, kind, scheme, method, nf)
    xv, xi, V, dt = sml_case(D, kind)
    meth, code, alpha = {"rk2": (jp.RungeKutta2(), O.JP_RK2, 0.5),
                         "rk2a": (jp.RungeKutta2(0.3), O.JP_RK2, 0.3),
                         "rk4": (jp.RungeKutta4(), O.JP_RK4, 0.5)}[method]
    F0 = [smooth(xv, 1 + f) for f in range(nf)]
    Finit = [np.full_like(F0[0], -5.0) for _ in range(nf)]
    F0d = [jp.DeviceArray.from_host(ctx, f) for f in F0]
    Fd = [jp.DeviceArray.from_host(ctx, f) for f in Finit]
    Vd = dev_V(ctx, V)
    if nf == 1:
        SML[scheme](Fd[0], F0d[0], meth, Vd, xi, xv, dt)
    else:
        SML[scheme](tuple(Fd), tuple(F0d), meth, Vd, xi, xv, dt)
    grid = tuple(compute_dx(a) for a in xv)
    gvi = tuple(tuple(compute_dx(a) for a in g) for g in xi)
    want = [f.copy() for f in Finit]
    O.semilagrangian(grid, gvi, V, want, F0, code, alpha, dt, scheme,
                     di_at_node=(nf > 1 and scheme == "linear"))
    for a, b in zip(Fd, want):
        assert np.array_equal(a.to_host(), b)


def test_semilagrangian_reference_refined_grid(ctx):
    """test/test_2D.jl:169-228 on the GPU: F[3, 3] = 1.18 by MQS, LinP and
    the scalar linear method, 1.14 by the NTuple linear method (its spacing
    taken at the node on this refined grid)."""
    xv = (Axis(np.array([0.0, 0.1, 0.3, 0.6, 1.0]), "vector"),
          Axis(np.linspace(0.0, 1.0, 5), "vector"))
    xc = [Axis((a.x[:-1] + a.x[1:]) / 2, "vector") for a in xv]
    xi = ((xv[0], expand_range(xc[1])), (expand_range(xc[0]), xv[1]))
    F0 = sample(lambda x, y: x + 2 * y, xv)
    V = tuple(jp.DeviceArray.from_host(ctx, np.full(tuple(a.n for a in g), u))
              for g, u in zip(xi, (0.08, 0.02)))
    F0d = jp.DeviceArray.from_host(ctx, F0)
    for scheme, tup, want in [("MQS", True, 1.18), ("LinP", True, 1.18), ("linear", False, 1.18),
                              ("linear", True, 1.14)]:
        Fd = F0d.copy()
        if tup:
            SML[scheme]((Fd,), (F0d,), jp.RungeKutta2(), V, xi, xv, 1.0)
        else:
            SML[scheme](Fd, F0d, jp.RungeKutta2(), V, xi, xv, 1.0)
        got = Fd.to_host()[2, 2]
        say("refined", scheme, tup, got)
        assert abs(got - want) <= 1e-12


def test_semilagrangian_out_of_grid(ctx):
    """Backtracks that leave the grid: the call raises; nodes whose backtrack
    stays inside hold the oracle's values, the others are unchanged."""
    xv, xi, _, _ = sml_case(2)
    V = tuple(np.full(tuple(a.n for a in g), 3.0) for g in xi)
    F0 = smooth(xv, 2)
    init = np.full_like(F0, -1.0)
    Fd, F0d = jp.DeviceArray.from_host(ctx, init), jp.DeviceArray.from_host(ctx, F0)
    Vd = dev_V(ctx, V)
    with pytest.raises(ValueError, match="left the velocity or vertex grid"):
        jp.semilagrangian_advection_(Fd, F0d, jp.RungeKutta2(), Vd, xi, xv, 0.2)
    want = init.copy()
    grid = tuple(compute_dx(a) for a in xv)
    gvi = tuple(tuple(compute_dx(a) for a in g) for g in xi)
    rc = O.semilagrangian(grid, gvi, V, [want], [F0], O.JP_RK2, 0.5, 0.2, check=False)
    assert rc != 0
    got = Fd.to_host()
    assert np.array_equal(got, want)
    assert (got == -1.0).any() and (got != -1.0).any()
    # the next call on the context succeeds: the flag was reported and cleared
    V2 = dev_V(ctx, tuple(np.zeros_like(v) for v in V))
    jp.semilagrangian_advection_(Fd, F0d, jp.RungeKutta2(), V2, xi, xv, 0.2)


def test_semilagrangian_async_reports_at_sync(ctx):
    xv, xi, _, _ = sml_case(2)
    V = dev_V(ctx, tuple(np.full(tuple(a.n for a in g), 3.0) for g in xi))
    F0d = jp.DeviceArray.from_host(ctx, smooth(xv, 2))
    Fd = F0d.copy()
    a = jp.Context(0, synchronous=False)
    try:
        Va = tuple(jp.DeviceArray.from_host(a, v.to_host()) for v in V)
        Fa, F0a = jp.DeviceArray.from_host(a, Fd.to_host()), jp.DeviceArray.from_host(a, F0d.to_host())
        jp.semilagrangian_advection_(Fa, F0a, jp.RungeKutta2(), Va, xi, xv, 0.2)  # no raise
        with pytest.raises(ValueError, match="left the velocity or vertex grid"):
            a.sync()
        a.sync()  # reported once
    finally:
        a.close()


def test_semilagrangian_rejects_aliasing(ctx):
    xv, xi, V, dt = sml_case(2)
    Fd = jp.DeviceArray.from_host(ctx, smooth(xv, 1))
    with pytest.raises(ValueError, match="different arrays"):
        jp.semilagrangian_advection_(Fd, Fd, jp.RungeKutta2(), dev_V(ctx, V), xi, xv, dt)
    with pytest.raises(ValueError):
        jp.semilagrangian_advection_(Fd, Fd.copy(), jp.Euler(), dev_V(ctx, V), xi, xv, dt)
