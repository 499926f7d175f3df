"""GPU checks of the adjoint of advection + particle2grid (runs on an MI355X).

- Forward against the oracle's flat model: final positions, survivor flags,
  checkpoints, F and W bit for bit (the nodal gather adds in particle order,
  as the oracle does).
- Backward against the oracle's adjoint for the same cotangent: x0bar and
  Tpbar bit for bit; Vbar within 1e-12 of its largest magnitude (FP64 atomic
  adds, whose order varies).
- Checkpoint intervals 1, 3, 10 and 16 give the same cotangents.
- The forward is the drop-in path: advection_ + move_particles_ on the
  container (a particle id carried as a field) give the same positions bit
  for bit for every particle the container keeps; the flat list keeps the
  particles a full cell drops.
- The inverse problem (256^2 cells, nxcell 16, 100 RK2 steps, checkpoints
  every 10): central finite differences on 32 random Vx/Vy entries, relative
  error <= 1e-5.  Entries are drawn among those whose gradient exceeds 1e-3 of
  the largest (the loss depends on them); a difference whose two runs leave a
  particle in another cell than the unperturbed run straddles a discontinuity
  of particle2grid! and is retaken with a step 100 times smaller.  The
  forward is deterministic, so the differences carry no run-to-run noise.
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
    print("[gpu-adjoint]", *a, flush=True)


def integrator(method, alpha):
    if method == EULER:
        return jp.Euler()
    if method == RK2:
        return jp.RungeKutta2(alpha)
    return jp.RungeKutta4()


def container(ctx, hp):
    return jp.particles_from_host(ctx, hp.xi_vel, hp.coords, hp.index, hp.K, hp.K, 1)


def close_to(a, b, rel):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) <= rel * scale


CASES = [(2, RK2, 0.5, False), (2, RK2, 0.3, True), (2, EULER, 0.5, True), (2, RK4, 0.5, False),
         (3, RK4, 0.5, False), (3, RK2, 0.5, True)]


@pytest.mark.parametrize("D,method,alpha,outflow", CASES)
def test_adjoint_matches_oracle(ctx, D, method, alpha, outflow):
    say("oracle parity", D, method, alpha, outflow)
    hp, V, dt = adjoint_case(D, outflow)
    N, every = 20, 3
    x0, _ = flat_active(hp)
    Tp = blob(x0)
    n = Tp.size
    r = O.adjoint_forward(hp, V, method, alpha, dt, N, x0, Tp, every=every)
    p = container(ctx, hp)
    op = jp.AdvectionP2G(p, integrator(method, alpha), dt, N, every)
    F = op.forward(V, x0, Tp)
    assert np.array_equal(op.alive.flat()[:n], r["alive"])
    for d in range(D):
        assert np.array_equal(op.xN[d].flat()[:n], r["xN"][d], equal_nan=True)
    ck = op._fw["ckpt"].flat()[:r["ckpt"].size].reshape(r["ckpt"].shape)
    assert np.array_equal(ck, r["ckpt"], equal_nan=True)
    assert np.array_equal(F.flat(), r["F"], equal_nan=True)
    assert np.array_equal(op.W.flat(), r["W"])
    ok = np.isfinite(r["F"])
    assert ok.any()

    rng = np.random.default_rng(4)
    Fb = rng.standard_normal(r["F"].size)
    Fb[~ok] = 0.0
    Vb, xb, Tb = O.adjoint_backward(hp, V, method, alpha, dt, N, x0, Tp, r["F"], r["W"], Fb)
    gVb, gxb, gTb = op.pullback(Fb)
    for d in range(D):
        assert np.array_equal(gxb[d].flat()[:n], xb[d])
    assert np.array_equal(gTb.flat()[:n], Tb)
    for c in range(D):
        assert close_to(gVb[c].to_host(), Vb[c], 1e-12)


def test_checkpoint_intervals_agree(ctx):
    say("checkpoint intervals")
    hp, V, dt = adjoint_case(2, outflow=False)
    N = 23
    x0, _ = flat_active(hp)
    Tp = blob(x0)
    n = Tp.size
    r = O.adjoint_forward(hp, V, RK2, 0.5, dt, N, x0, Tp)
    Fb = np.cos(np.arange(r["F"].size))
    Fb[~np.isfinite(r["F"])] = 0.0
    p = container(ctx, hp)
    out = []
    for every in (1, 3, 10, 16):
        op = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, every)
        F = op.forward(V, x0, Tp)
        assert np.array_equal(F.flat(), r["F"], equal_nan=True)
        Vb, xb, Tb = op.pullback(Fb)
        out.append(([v.to_host() for v in Vb], [x.flat()[:n] for x in xb], Tb.flat()[:n]))
    for Vb, xb, Tb in out[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(xb, out[0][1]))
        assert np.array_equal(Tb, out[0][2])
        assert all(close_to(a, b, 1e-13) for a, b in zip(Vb, out[0][0]))


def test_forward_is_the_dropin_path(ctx):
    say("drop-in path")
    cfg = inverse2d(64)
    p = jp.init_particles(ctx, 16, 64, 8, *cfg["xi_vel"], seed=11)
    coords, index = p.host_state()
    m = index.astype(bool)
    n = int(m.sum())
    x0 = [c[m] for c in coords]
    ids = np.full(index.size, np.nan)
    ids[m] = np.arange(n)
    f = jp.CellArray(ctx, p.ni, p.max_xcell, fill=float("nan"))
    f.data.copy_from(ids)
    N = 100
    op = jp.AdvectionP2G(p, jp.RungeKutta2(), cfg["dt"], N, 10)
    op.forward(cfg["V"], x0, np.ones(n))
    Vd = tuple(jp.DeviceArray.from_host(ctx, v) for v in cfg["V"])
    dropped = 0
    for _ in range(N):
        jp.advection_(p, jp.RungeKutta2(), Vd, cfg["dt"])
        dropped += jp.move_particles_(p, (f,), stats=True)["dropped"]
    coords, index = p.host_state()
    m2 = index.astype(bool)
    got = f.data.flat()[m2].astype(np.int64)
    alive = np.nonzero(op.alive.flat()[:n])[0]
    # the container loses the particles a full cell drops; the flat list has
    # no capacity and keeps them (the rotating cell deletes none)
    say("  dropped by the container:", dropped, "of", n)
    assert np.all(np.isin(got, alive))
    assert alive.size - got.size == dropped
    for d in range(2):
        assert np.array_equal(coords[d][m2], op.xN[d].flat()[:n][got])


def test_inverse_problem_gradient_matches_central_differences(ctx):
    say("inverse problem FD")
    cfg = inverse2d(256)
    p = jp.init_particles(ctx, 16, 32, 8, *cfg["xi_vel"], seed=7)
    x0d, _ = jp.flatten(p)
    x0 = [x.to_host() for x in x0d]
    Tp = cfg["T"](x0[0], x0[1])
    N, dt = 100, cfg["dt"]
    op_t = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, 10)
    Ft = op_t.forward(cfg["V_true"], x0d, Tp).flat()
    V = cfg["V"]
    op = jp.AdvectionP2G(p, jp.RungeKutta2(), dt, N, 10)
    F0 = op.forward(V, x0d, Tp).flat()
    L, Fb = jp.l2_loss(op._fw["F"], Ft)
    Vb, _, _ = op.pullback(Fb)
    Vb = [v.to_host().ravel(order="F") for v in Vb]
    cell_ax = [np.asarray(a.x) for a in cfg["xv"]]

    def branch(o):
        xs = [x.flat() for x in o.xN]
        cells = [np.searchsorted(ax, x, side="right") for ax, x in zip(cell_ax, xs)]
        return o.alive.flat(), cells

    base = branch(op)
    say("  loss", L, "particles", Tp.size)

    def dloss(Fp, Fm):
        ok = np.isfinite(Fp) & np.isfinite(Fm) & np.isfinite(Ft)
        a, b = Fp[ok], Fm[ok]
        return np.sum((a - b) * (0.5 * (a + b) - Ft[ok]))

    rng = np.random.default_rng(2024)
    worst, retakes = 0.0, 0
    for c in range(2):
        g = Vb[c]
        cand = np.nonzero(np.abs(g) >= 1e-3 * np.max(np.abs(g)))[0]
        for e in rng.choice(cand, 16, replace=False):
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
                retakes += 1
                h /= 100.0
            fd = dloss(Fs[0], Fs[1]) / (2 * h)
            rel = abs(fd - g[e]) / abs(fd)
            worst = max(worst, rel)
            assert rel <= 1e-5, (c, int(e), fd, g[e], rel)
    say("  worst relative error", worst, "retaken", retakes)
    assert np.isfinite(F0).sum() > 0.99 * F0.size
