"""Pins of the oracle's adjoint of advection + particle2grid (CPU only).

- The flat-list forward model is the reference path: N steps of
  jpo_advection + jpo_move_particles (with a particle id carried as a field)
  followed by jpo_particle2grid give the same positions bit for bit, the same
  deletions, and the same nodal field up to summation order.
- The backward is the derivative of that forward: central finite differences
  on velocity entries, initial positions and tracer values, relative error
  <= 1e-5 for Euler, RK2 (alpha 0.5 and 0.3) and RK4, 2D and 3D.
- Uniform velocity: the interpolation has zero slope and its weights sum to
  one, so sum(Vbar_c) = nsteps * dt * sum(x0bar_c) for every integrator.
- Checkpoint block b holds the positions after b*every steps.
"""
import numpy as np
import pytest

from helpers import adjoint_case, blob, flat_active, sample
from oracle import oracle as O

EULER, RK2, RK4 = O.JP_EULER, O.JP_RK2, O.JP_RK4


@pytest.mark.parametrize("D,method,alpha,outflow", [
    (2, RK2, 0.5, False), (2, RK2, 0.3, False), (2, EULER, 0.5, True), (2, RK4, 0.5, True),
    (3, RK4, 0.5, False), (3, RK2, 0.5, True)])
def test_flat_forward_is_the_reference_path(D, method, alpha, outflow):
    hp, V, dt = adjoint_case(D, outflow)
    N = 20
    x0, m = flat_active(hp)
    n = x0[0].size
    Tp = blob(x0)
    r = O.adjoint_forward(hp, V, method, alpha, dt, N, x0, Tp, every=10)
    ids = np.full(hp.S, np.nan)
    ids[m] = np.arange(n)
    Tf = np.full(hp.S, np.nan)
    Tf[m] = Tp
    for _ in range(N):
        O.advection(hp, V, method, alpha, dt)
        st = O.move_particles(hp, (ids, Tf))
        assert st["dropped"] == 0
    m2 = hp.index.astype(bool)
    got = ids[m2].astype(np.int64)
    alive = np.nonzero(r["alive"])[0]
    assert np.array_equal(np.sort(got), alive)
    if outflow:
        assert 0 < alive.size < n
    for d in range(D):
        assert np.array_equal(hp.coords[d][m2], r["xN"][d][got])
    F = O.particle2grid(hp, Tf)
    assert np.array_equal(np.isnan(F), np.isnan(r["F"]))
    ok = np.isfinite(F)
    np.testing.assert_allclose(r["F"][ok], F[ok], rtol=1e-13, atol=0)


def _loss_and_grads(D, method, alpha, nsteps=25):
    hp, V, dt = adjoint_case(D)
    x0, _ = flat_active(hp)
    Tp = blob(x0)
    rng = np.random.default_rng(0)
    Vt = tuple(v * (1 + 0.1 * rng.standard_normal(v.shape)) for v in V)
    Ft = O.adjoint_forward(hp, Vt, method, alpha, dt, nsteps, x0, Tp)["F"]

    def loss(V, x0=x0, Tp=Tp):
        r = O.adjoint_forward(hp, V, method, alpha, dt, nsteps, x0, Tp)
        d = r["F"] - Ft
        d[~np.isfinite(d)] = 0
        return 0.5 * np.sum(d * d), r, d

    L, r, Fb = loss(V)
    Vb, xb, Tb = O.adjoint_backward(hp, V, method, alpha, dt, nsteps, x0, Tp, r["F"], r["W"], Fb)
    return loss, V, x0, Tp, Vb, xb, Tb, rng


def _perturbed(V, c, e, h):
    out = [v.ravel(order="F").copy() for v in V]
    out[c][e] += h
    return tuple(a.reshape(v.shape, order="F") for a, v in zip(out, V))


@pytest.mark.parametrize("D,method,alpha", [
    (2, RK2, 0.5), (2, RK2, 0.3), (2, EULER, 0.5), (2, RK4, 0.5), (3, RK2, 0.5), (3, RK4, 0.5)])
def test_adjoint_matches_central_differences(D, method, alpha):
    loss, V, x0, Tp, Vb, xb, Tb, rng = _loss_and_grads(D, method, alpha)
    tol = 1e-5
    for c in range(D):
        g = Vb[c].ravel(order="F")
        nz = np.nonzero(g)[0]
        assert nz.size > 8
        picks = list(np.argsort(-np.abs(g))[:4]) + list(rng.choice(nz, 4, replace=False))
        for e in picks:
            h = 1e-6 * max(1.0, abs(V[c].ravel(order="F")[e]))
            fd = (loss(_perturbed(V, c, e, h))[0] - loss(_perturbed(V, c, e, -h))[0]) / (2 * h)
            assert abs(fd - g[e]) <= tol * abs(fd), (c, e, fd, g[e])
    hx = 1e-9
    for d in range(D):
        for i in np.argsort(-np.abs(xb[d]))[:3]:
            xp = [x.copy() for x in x0]
            xm = [x.copy() for x in x0]
            xp[d][i] += hx
            xm[d][i] -= hx
            fd = (loss(V, xp)[0] - loss(V, xm)[0]) / (2 * hx)
            assert abs(fd - xb[d][i]) <= tol * abs(fd), (d, i, fd, xb[d][i])
    for i in np.argsort(-np.abs(Tb))[:3]:
        tp = Tp.copy()
        tm = Tp.copy()
        tp[i] += 1e-6
        tm[i] -= 1e-6
        fd = (loss(V, x0, tp)[0] - loss(V, x0, tm)[0]) / 2e-6
        assert abs(fd - Tb[i]) <= tol * abs(fd)


@pytest.mark.parametrize("D,method,alpha", [
    (2, EULER, 0.5), (2, RK2, 0.5), (2, RK2, 0.3), (2, RK4, 0.5), (3, RK2, 0.5), (3, RK4, 0.5)])
def test_uniform_velocity_identity(D, method, alpha):
    hp, _, _ = adjoint_case(D)
    xi = hp.xi_vel
    u = (0.4, -0.3, 0.2)
    V = tuple(sample(lambda *X, c=c: np.full_like(X[0], u[c]), xi[c]) for c in range(D))
    dt = 0.02
    N = 15
    x0, _ = flat_active(hp)
    Tp = blob(x0)
    r = O.adjoint_forward(hp, V, method, alpha, dt, N, x0, Tp)
    assert 0 < r["alive"].sum() < Tp.size  # some particles leave the domain
    rng = np.random.default_rng(1)
    Fb = rng.standard_normal(r["F"].size)
    Fb[~np.isfinite(r["F"])] = 0
    Vb, xb, Tb = O.adjoint_backward(hp, V, method, alpha, dt, N, x0, Tp, r["F"], r["W"], Fb)
    for c in range(D):
        lhs = Vb[c].sum()
        rhs = N * dt * xb[c].sum()
        scale = N * dt * np.abs(xb[c]).sum()
        assert abs(lhs - rhs) <= 1e-12 * scale, (c, lhs, rhs)
    # deleted particles carry no cotangent
    dead = r["alive"] == 0
    assert np.all(Tb[dead] == 0) and all(np.all(x[dead] == 0) for x in xb)


def test_checkpoints_hold_block_starts():
    hp, V, dt = adjoint_case(2, outflow=True)
    x0, _ = flat_active(hp)
    Tp = blob(x0)
    every, N = 3, 10
    r = O.adjoint_forward(hp, V, RK2, 0.5, dt, N, x0, Tp, every=every)
    assert r["ckpt"].shape == (4, 2, Tp.size)
    for b in range(4):
        q = O.adjoint_forward(hp, V, RK2, 0.5, dt, b * every, x0, Tp, every=every)
        for d in range(2):
            assert np.array_equal(r["ckpt"][b, d], q["xN"][d], equal_nan=True)
    z = O.adjoint_forward(hp, V, RK2, 0.5, dt, 0, x0, Tp, every=every)
    assert z["ckpt"].shape == (1, 2, Tp.size)
    assert all(np.array_equal(z["ckpt"][0, d], x0[d]) for d in range(2))
