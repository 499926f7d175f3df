"""The batched ensemble against the FP64 oracle, and the 64-member ensemble
through properties (runs on an MI355X).

* A few members of uneven size (including a 192^2 member with K = 48 and a
  member on the README grid) take ensemble steps (persistent split
  advection, per-member move on the sub-streams); every member is compared
  with the oracle's advection! + move_particles! by particle identity:
  positions within 1e-6 of the domain, cell disagreements within 1e-4.
* The 64-member workload: every member's particle count changes only by
  the deletions and drops its move reports; a captured hipGraph step
  equals the same step run eagerly, member by member, bit for bit.
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.configs import stream2d
from oracle import oracle as O

from helpers import host_particles
from test_gpu_split import POS_TOL, compare_by_id, split_state, to_dev

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-ensemble-oracle]", *a, flush=True)


MEMBERS = [  # (cells per side, K, particles per quadrant, amplitude, README grid)
    (64, 48, 6, 120.0, False),
    (192, 48, 6, 250.0, False),
    (96, 32, 4, 300.0, True),
]


def test_ensemble_members_vs_fp64_oracle(ctx):
    hosts, devs, ids_h, ids_d, Vs = [], [], [], [], []
    for i, (n, K, npq, amp, readme) in enumerate(MEMBERS):
        cfg = stream2d(n, amp=amp, readme=readme)
        hp = host_particles(cfg["xv"], cfg["xi_vel"], K, npq, seed=30 + i)
        ids = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
        s = jp.to_split(to_dev(ctx, hp))
        idd = jp.CellArray(ctx, hp.ni, hp.K)
        idd.copy_from(ids)
        V = tuple(jp.DeviceArray.from_host(ctx, v) for v in cfg["V"])
        hosts.append((hp, cfg))
        devs.append((s, V, cfg["dt"]))
        ids_h.append(ids)
        ids_d.append(idd)
        Vs.append(V)
    ens = jp.Ensemble(ctx, devs)
    nodal = [jp.nodal_field(s) for s, _, _ in devs]
    steps = 8
    for it in range(steps):
        ens.step(ids_d, nodal)
        for (hp, cfg), ids in zip(hosts, ids_h):
            O.advection(hp, cfg["V"], O.JP_RK2, 0.5, cfg["dt"])
            O.move_particles(hp, [ids])
    ctx.sync()
    for (hp, cfg), (s, _, _), ids, idd in zip(hosts, devs, ids_h, ids_d):
        coords, index = split_state(s)
        err, dis, n = compare_by_id((coords, index, idd.data.flat()), (hp.coords, hp.index, ids),
                                    hp.C)
        say(f"member {hp.ni} K {hp.K}: {n} particles, {steps} steps, err {err:.3e}, "
            f"disagreements {dis}")
        assert err <= POS_TOL
        assert dis <= 1e-4 * n
    ens.close()


def _member_state(p, f, F):
    return ([o.data.flat() for o in p.off], p.index.data.flat(), f.data.flat(), F.flat())


def test_ensemble_64_members_properties(ctx):
    """The 64-member workload (cells drawn from {64..512}^2, amplitude
    U[50, 400], own CFL dt, nxcell 24, K 48): particle counts conserved up to
    reported deletions and drops; a graph step equals an eager step."""
    ctx_a = jp.Context(0, synchronous=False)
    ens_a, f_a, F_a = jp.random_ensemble(ctx_a, 64, seed=11)
    say("built ensemble A", sum(p.C for p, _, _ in ens_a.members), "cells")
    # conservation over one step with per-member outcome counts
    before = [p.active() for p, _, _ in ens_a.members]
    ens_a.advect()
    stats = [jp.move_particles_(p, (f,), stats=True) for (p, _, _), f in zip(ens_a.members, f_a)]
    after = [p.active() for p, _, _ in ens_a.members]
    lost = sum(s["deleted"] + s["dropped"] for s in stats)
    say("conservation", sum(before), sum(after), lost)
    for b, a, s in zip(before, after, stats):
        assert a == b - s["deleted"] - s["dropped"]
    assert sum(s["moved"] for s in stats) > 0
    ens_a.close()
    del ens_a, f_a, F_a
    ctx_a.close()
    # graph replays against eager steps on an identical ensemble
    ctx_g = jp.Context(0, synchronous=False)
    ens_g, f_g, F_g = jp.random_ensemble(ctx_g, 64, seed=12)
    ens_e, f_e, F_e = jp.random_ensemble(ctx_g, 64, seed=12)
    ens_g.step(f_g, F_g)  # sizes the scratch; the eager ensemble takes the same step
    ens_e.step(f_e, F_e)
    g = ens_g.capture_step(f_g, F_g)
    for r in range(2):
        ctx_g.graph_launch(g)
        ctx_g.sync()
        ens_e.step(f_e, F_e)
        ctx_g.sync()
        for (pg, _, _), fg, Fg, (pe, _, _), fe, Fe in zip(ens_g.members, f_g, F_g, ens_e.members,
                                                          f_e, F_e):
            a = _member_state(pg, fg, Fg)
            b = _member_state(pe, fe, Fe)
            for d in range(2):
                assert np.array_equal(a[0][d], b[0][d], equal_nan=True), (r, pg.ni)
            assert np.array_equal(a[1], b[1]), (r, pg.ni)
            assert np.array_equal(a[2], b[2], equal_nan=True), (r, pg.ni)
            assert np.array_equal(a[3], b[3], equal_nan=True), (r, pg.ni)
        say("replay", r, "equal")
    ctx_g.graph_destroy(g)
    ens_g.close()
    ens_e.close()
    ctx_g.close()
