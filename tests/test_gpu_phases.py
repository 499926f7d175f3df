"""inject_particles_phase! and update_phase_ratios! on the GPU against the
oracle (runs on an MI355X).  Bar: bit-exact (masks, positions, phases,
interpolated fields, every phase-ratio value; NaN where the oracle has NaN).
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.grids import Axis, staggered_grid
from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V
from test_gpu_parity import assert_field, assert_state_equal, dev_V, to_dev

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-phases]", *a, flush=True)


def dev_cells(ctx, hp, a):
    ca = jp.CellArray(ctx, hp.ni, hp.K)
    ca.copy_from(a)
    return ca


def check_ratios(pr, want):
    for k, arr in want.items():
        got = pr.fields[k].to_host()
        assert np.array_equal(np.isnan(got), np.isnan(arr)), k
        m = ~np.isnan(arr)
        assert np.array_equal(got[m], arr[m]), (k, np.max(np.abs(got[m] - arr[m])))


@pytest.mark.parametrize("nph", [2, 5, 16])
@pytest.mark.parametrize("D", [2, 3])
def test_phase_ratios_match_oracle(ctx, D, nph):
    say("phase ratios", D, nph)
    if D == 2:
        xv, xi = grid2d(40, 36)
        hp = host_particles(xv, xi, 48, 6, seed=12)
    else:
        xv, xi = grid3d(7)
        hp = host_particles(xv, xi, 64, 6, seed=12)
    rng = np.random.default_rng(nph)
    phases = np.where(hp.index == 1, rng.integers(1, nph + 1, hp.S), np.nan).astype(float)
    # a few empty cells: NaN centres, zero faces
    if D == 2:
        for e in range(0, hp.S, 97):
            hp.index[e] = 0
            for c in hp.coords:
                c[e] = np.nan
    want = O.update_phase_ratios(hp, phases, nph)
    p = to_dev(ctx, hp)
    pr = jp.PhaseRatios(ctx, nph, hp.ni)
    jp.update_phase_ratios_(pr, p, dev_cells(ctx, hp, phases))
    check_ratios(pr, want)


@pytest.mark.parametrize("D", [2, 3])
def test_inject_phase_matches_oracle(ctx, D):
    """A moved state with thinned cells: positions, masks, phases, a nodal
    and a cell-centred field interpolated into the new particles, stats."""
    say("inject phase", D)
    if D == 2:
        xv, xi = grid2d(24, 20)
        V = stream_V(xi)
        hp = host_particles(xv, xi, 32, 5, seed=21)
        minx = 16
    else:
        xv, xi = grid3d(6)
        V = cellular_V(xi)
        hp = host_particles(xv, xi, 40, 3, seed=21)
        minx = 20
    rng = np.random.default_rng(3)
    phases = np.where(hp.index == 1, rng.integers(1, 4, hp.S), np.nan).astype(float)
    O.advection(hp, V, O.JP_RK2, 0.5, cfl_dt(xv, V))
    O.move_particles(hp, [phases])
    # thin out: drop a third of the particles
    drop = (hp.index == 1) & (rng.random(hp.S) < 0.35)
    hp.index[drop] = 0
    for c in hp.coords:
        c[drop] = np.nan
    phases[drop] = np.nan
    xc = O.centre_axes(hp)
    mesh_v = np.meshgrid(*[a.x for a in xv], indexing="ij")
    mesh_c = np.meshgrid(*[a.x for a in xc], indexing="ij")
    nodal = np.asfortranarray(np.cos(3 * sum(mesh_v)) + mesh_v[0] ** 2)
    centred = np.asfortranarray(np.sin(5 * mesh_c[0]) * np.cos(4 * mesh_c[-1]))
    a1 = np.where(hp.index == 1, 0.5, np.nan)
    a2 = np.where(hp.index == 1, -0.5, np.nan)
    p = to_dev(ctx, hp, min_xcell=minx, seed=17)
    ph_d = dev_cells(ctx, hp, phases)
    a1d, a2d = dev_cells(ctx, hp, a1), dev_cells(ctx, hp, a2)
    fields = (jp.DeviceArray.from_host(ctx, nodal), jp.DeviceArray.from_host(ctx, centred))
    st_d = jp.inject_particles_phase_(p, ph_d, (a1d, a2d), fields, stats=True)
    st_o = O.inject_particles_phase(hp, phases, [a1, a2], [nodal, centred], minx, seed=17,
                                    counter=0)
    say("stats", st_d, st_o)
    assert st_o["injected"] > 0
    assert st_d == st_o
    assert_state_equal(p, hp, exact=True)
    assert_field(ph_d.data.flat(), phases, exact=True)
    assert_field(a1d.data.flat(), a1, exact=True)
    assert_field(a2d.data.flat(), a2, exact=True)


def test_pure_shear_loop(ctx):
    """test/test_2D.jl:386-480 on the GPU: pure shear on [-1/2, 1/2]^2
    (64^2 cells; nxcell 48, max 64, min 40 for K <= 64), phase 1 where x < y,
    else 2; per step advection! (RK2), move_particles! with the phases,
    inject_particles_phase!, update_phase_ratios!.  Every step equals the
    oracle bit for bit, and every ratio field sums to one."""
    n = 64
    ax = Axis.lin(-0.5, 0.5, n + 1)
    xv = (ax, ax)
    xi = staggered_grid(xv)
    Vx = np.asfortranarray(np.repeat((xi[0][0].x * -1.0)[:, None], xi[0][1].n, axis=1))
    Vy = np.asfortranarray(np.repeat((-xi[1][1].x * -1.0)[None, :], xi[1][0].n, axis=0))
    V = (Vx, Vy)
    dt = 0.25 * (1.0 / n) / max(np.abs(Vx).max(), np.abs(Vy).max())
    hp = host_particles(xv, xi, 64, 12, seed=5)
    x, y = hp.coords
    phases = np.where(hp.index == 1, np.where(x < y, 1.0, 2.0), np.nan)
    p = to_dev(ctx, hp, nxcell=48, min_xcell=40, seed=2)
    ph_d = dev_cells(ctx, hp, phases)
    Vd = dev_V(ctx, V)
    pr = jp.PhaseRatios(ctx, 2, hp.ni)
    for it in range(6):
        jp.advection_(p, jp.RungeKutta2(), Vd, dt)
        jp.move_particles_(p, (ph_d,))
        st = jp.inject_particles_phase_(p, ph_d, stats=True)
        jp.update_phase_ratios_(pr, p, ph_d)
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        O.move_particles(hp, [phases])
        st_o = O.inject_particles_phase(hp, phases, [], [], 40, seed=2, counter=it)
        want = O.update_phase_ratios(hp, phases, 2)
        assert st == st_o
        assert_state_equal(p, hp, exact=True)
        assert_field(ph_d.data.flat(), phases, exact=True)
        check_ratios(pr, want)
    say("injected on the last step", st)
    dims = jp.phases.ratio_dims(hp.ni)
    for k in ("vertex", "center", "Vx", "Vy"):
        r = pr.fields[k].to_host().reshape(2, -1)
        assert np.all(np.abs(r.sum(axis=0) - 1.0) < 1e-12), k
        assert r.shape[1] == int(np.prod(dims[k]))
