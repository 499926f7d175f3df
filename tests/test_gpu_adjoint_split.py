"""The adjoint's forward pass in the FP32-split form (runs on an MI355X).

- The split trajectories are the product split path's: the same particles
  advanced by advection_ + move_particles_ on a split container (a particle
  id carried as a field) end at the same positions bit for bit, for every
  particle the container keeps (the flat list has no slot capacity).
- Against the FP64 oracle's flat model: final positions within 1e-6 of the
  domain, the same survivors except near the walls, and F close to the
  oracle's (float32 positions move the 1/d^2 weights).
- The gradient from the split forward (the FP64 pullback differentiates
  the FP64 map along the split trajectory, recomputing each block from the
  split checkpoints) on the 256^2 inverse problem, against the FP64
  forward's gradient (itself within 1e-5 of central differences,
  tests/test_gpu_adjoint.py) and against central differences of the FP64
  loss on 8 random Vx/Vy entries.  The split trajectory is ~1e-8 of the
  domain from the FP64 one, and the 1/d^2 weights of particle2grid amplify
  that near nodes, so the bars are 1e-5 for the median entry and 1e-3 for
  the worst (measured: printed by the test).
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.configs import inverse2d
from oracle import oracle as O

from helpers import adjoint_case, blob, flat_active

pytestmark = pytest.mark.gpu

EULER, RK2, RK4 = O.JP_EULER, O.JP_RK2, O.JP_RK4


def say(*a):
    print("[gpu-adjoint-split]", *a, flush=True)


def integrator(method, alpha):
    if method == EULER:
        return jp.Euler()
    if method == RK2:
        return jp.RungeKutta2(alpha)
    return jp.RungeKutta4()


def container(ctx, hp):
    return jp.particles_from_host(ctx, hp.xi_vel, hp.coords, hp.index, hp.K, hp.K, 1)


def test_split_forward_is_the_split_path(ctx):
    say("split path")
    cfg = inverse2d(64)
    p = jp.init_particles(ctx, 16, 64, 8, *cfg["xi_vel"], seed=11)
    coords, index = p.host_state()
    m = index.astype(bool)
    n = int(m.sum())
    x0 = [c[m] for c in coords]
    N = 60
    op = jp.AdvectionP2G(p, jp.RungeKutta2(), cfg["dt"], N, 10, form="split")
    op.forward(cfg["V"], x0, np.ones(n))
    s = jp.to_split(p)
    ids = np.full(index.size, np.nan)
    ids[m] = np.arange(n)
    f = jp.CellArray(ctx, p.ni, p.max_xcell, fill=float("nan"))
    f.data.copy_from(ids)
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in cfg["V"])
    dropped = 0
    for _ in range(N):
        jp.advection_(s, jp.RungeKutta2(), Vd, cfg["dt"])
        dropped += jp.move_particles_(s, (f,), stats=True)["dropped"]
    q = jp.to_f64(s)
    coords, index = q.host_state()
    m2 = index.astype(bool)
    got = f.data.flat()[m2].astype(np.int64)
    alive = np.nonzero(op.alive.flat()[:n])[0]
    say("  dropped by the container:", dropped, "of", n)
    assert np.all(np.isin(got, alive))
    assert alive.size - got.size == dropped
    for d in range(2):
        assert np.array_equal(coords[d][m2], op.xN[d].flat()[:n][got])


@pytest.mark.parametrize("method", [RK2, RK4])
def test_split_forward_is_the_split_path_3d(ctx, method):
    """3D: the flat split forward and the split container path, bit for bit
    for the particles the container keeps."""
    say("split path 3d", method)
    hp, V, dt = adjoint_case(3, False)
    x0, m = flat_active(hp)
    n = x0[0].size
    N = 12
    p = container(ctx, hp)
    op = jp.AdvectionP2G(p, integrator(method, 0.5), dt, N, 4, form="split")
    op.forward(V, x0, np.ones(n))
    s = jp.to_split(p)
    ids = np.full(hp.S, np.nan)
    ids[m] = np.arange(n)
    f = jp.CellArray(ctx, hp.ni, hp.K, fill=float("nan"))
    f.data.copy_from(ids)
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in V)
    for _ in range(N):
        jp.advection_(s, integrator(method, 0.5), Vd, dt)
        jp.move_particles_(s, (f,))
    coords, index = jp.to_f64(s).host_state()
    m2 = index.astype(bool)
    got = f.data.flat()[m2].astype(np.int64)
    assert got.size > 0.9 * n
    for d in range(3):
        assert np.array_equal(coords[d][m2], op.xN[d].flat()[:n][got])


# position bars: 1e-6 of the domain in 2D; in the 3D cellular flow (strain
# rate pi*100 at the stagnation points) 20 RK4 steps at 0.7 of the CFL limit
# on a 5^3 grid amplify float32 rounding by about e^(pi*100*0.014) ~ 80, so
# the bar there is 5e-5
CASES = [(2, RK2, 0.5, False, 1e-6), (2, RK2, 0.3, True, 1e-6), (2, EULER, 0.5, True, 1e-6),
         (2, RK4, 0.5, False, 1e-6), (3, RK4, 0.5, False, 5e-5), (3, RK2, 0.5, True, 1e-6)]


@pytest.mark.parametrize("D,method,alpha,outflow,bar", CASES)
def test_split_forward_against_oracle(ctx, D, method, alpha, outflow, bar):
    say("oracle", D, method, alpha, outflow)
    hp, V, dt = adjoint_case(D, outflow)
    N, every = 20, 3
    x0, _ = flat_active(hp)
    Tp = blob(x0)
    n = Tp.size
    r = O.adjoint_forward(hp, V, method, alpha, dt, N, x0, Tp, every=every)
    op = jp.AdvectionP2G(container(ctx, hp), integrator(method, alpha), dt, N, every,
                         form="split")
    F = op.forward(V, x0, Tp).flat()
    al = op.alive.flat()[:n].astype(bool)
    both = al & r["alive"].astype(bool)
    err = max(float(np.max(np.abs(op.xN[d].flat()[:n][both] - r["xN"][d][both])))
              for d in range(D))
    # a particle within a few float32 ulps of a wall can be on the other side
    flips = int(np.count_nonzero(al != r["alive"].astype(bool)))
    ok = np.isfinite(F) & np.isfinite(r["F"])
    dF = np.abs(F[ok] - r["F"][ok])
    say("  max position error", err, "survivor flips", flips, "of", n,
        "|dF| max", dF.max(), "p99", np.quantile(dF, 0.99))
    assert err <= bar
    assert flips <= max(2, n // 1000)
    assert np.array_equal(np.isfinite(F), np.isfinite(r["F"])) or flips > 0
    assert np.quantile(dF, 0.99) <= 1e-4


def test_split_forward_gradient(ctx):
    say("inverse problem gradient")
    cfg = inverse2d(256)
    p = jp.init_particles(ctx, 16, 32, 8, *cfg["xi_vel"], seed=7)
    x0d, _ = jp.flatten(p)
    x0 = [x.to_host() for x in x0d]
    Tp = cfg["T"](x0[0], x0[1])
    N, dt = 100, cfg["dt"]
    Ft = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, 10).forward(cfg["V_true"], x0d, Tp).flat()
    V = cfg["V"]
    grads = {}
    for form in ("f64", "split"):
        op = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, 10, form=form)
        op.forward(V, x0d, Tp)
        L, Fb = jp.l2_loss(op._fw["F"], Ft)
        Vb, _, _ = op.pullback(Fb)
        grads[form] = ([v.to_host().ravel(order="F") for v in Vb], L)
        say("  loss", form, L)
    rels = []
    for c in range(2):
        g64, gs = grads["f64"][0][c], grads["split"][0][c]
        big = np.abs(g64) >= 1e-3 * np.max(np.abs(g64))
        rels.append(np.abs(gs[big] - g64[big]) / np.abs(g64[big]))
    rels = np.concatenate(rels)
    say("  split vs FP64 gradient relative difference: median", np.median(rels), "p99",
        np.quantile(rels, 0.99), "worst", rels.max(), "over", rels.size, "entries")
    assert np.median(rels) <= 1e-5
    assert rels.max() <= 1e-3

    def dloss(Fp, Fm):
        ok = np.isfinite(Fp) & np.isfinite(Fm) & np.isfinite(Ft)
        a, b = Fp[ok], Fm[ok]
        return np.sum((a - b) * (0.5 * (a + b) - Ft[ok]))

    # a difference whose perturbed runs leave a particle in another cell than
    # the unperturbed run straddles a jump of particle2grid! and is retaken
    # with a step 100 times smaller (as in tests/test_gpu_adjoint.py)
    cell_ax = [np.asarray(a.x) for a in cfg["xv"]]

    def branch(o):
        xs = [x.flat() for x in o.xN]
        return o.alive.flat(), [np.searchsorted(ax, x, side="right")
                                for ax, x in zip(cell_ax, xs)]

    ref = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, 10)
    ref.forward(V, x0d, Tp)
    base = branch(ref)
    rng = np.random.default_rng(99)
    for c in range(2):
        g = grads["split"][0][c]
        cand = np.nonzero(np.abs(g) >= 1e-2 * np.max(np.abs(g)))[0]
        for e in rng.choice(cand, 4, replace=False):
            v = V[c].ravel(order="F")
            h = 1e-6 * max(1.0, abs(v[e]))
            for _ in range(3):
                Fs = []
                same = True
                for sgn in (1, -1):
                    Vp = [a.ravel(order="F").copy() for a in V]
                    Vp[c][e] += sgn * h
                    Vp = [a.reshape(b.shape, order="F") for a, b in zip(Vp, V)]
                    o = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, 10)
                    Fs.append(o.forward(Vp, x0d, Tp).flat())
                    al, cells = branch(o)
                    same = same and np.array_equal(al, base[0]) and all(
                        np.array_equal(a, b) for a, b in zip(cells, base[1]))
                if same:
                    break
                h /= 100.0
            fd = dloss(Fs[0], Fs[1]) / (2 * h)
            rel = abs(fd - g[e]) / abs(fd)
            say("  FD", c, int(e), fd, g[e], rel)
            assert rel <= 1e-3
