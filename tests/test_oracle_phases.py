"""Pins of the oracle's inject_particles_phase! and update_phase_ratios! (CPU).

* Partition of unity at every location, on the reference's own setups
  (test/test_CellArrays.jl:80-183: 5 random phases, 2D and 3D, all eight
  fields; test/test_2D.jl:386-480: the pure-shear two-phase split).
* Weights computed by hand for particles at dyadic offsets (a centre, a
  vertex, a face), so a dropped term of the bilinear weight, a wrong
  half-cell test or a wrong location fails.
* Relabelling two phases swaps their ratio planes exactly.
* Injection: every short quadrant is filled (the plain injection stops at
  the first full quadrant), the new particle takes the phase of its nearest
  neighbour found by brute force, and fields are interpolated (a linear
  field reproduced) and clamped to the corner extrema.
"""
import numpy as np
import pytest

from justpic_hip.grids import Axis, staggered_grid
from oracle import oracle as O

from helpers import clear, grid2d, grid3d, host_particles, place


def ratios_by_location(out, dims, nph):
    L = int(np.prod(dims))
    return out.reshape(nph, L).T  # (L, nph)


@pytest.mark.parametrize("D", [2, 3])
def test_partition_of_unity_random_phases(D):
    """test/test_CellArrays.jl:80-112 (2D) and :144-183 (3D), with K <= 64."""
    if D == 2:
        xv, xi = grid2d(40, 36)
        hp = host_particles(xv, xi, 48, 12, seed=3)
    else:
        xv, xi = grid3d(8)
        hp = host_particles(xv, xi, 64, 8, seed=3)
    nph = 5
    rng = np.random.default_rng(1)
    phases = np.where(hp.index == 1, rng.integers(1, nph + 1, hp.S), np.nan).astype(float)
    out = O.update_phase_ratios(hp, phases, nph)
    dims = O.phase_ratio_shapes(hp.ni)
    assert set(out) == set(dims)
    for k, v in out.items():
        r = ratios_by_location(v, dims[k], nph)
        s = r.sum(axis=1)
        assert np.all(np.abs(s - 1.0) < 1e-12), (k, s.min(), s.max())
        assert np.all((r >= 0) & (r <= 1))


def test_partition_of_unity_pure_shear_split():
    """test/test_2D.jl:386-457: vertices on [-1/2, 1/2]^2, phase 1 where
    x < y, else 2; after update_phase_ratios! every field sums to one."""
    n = 32
    ax = Axis.lin(-0.5, 0.5, n + 1)
    xv = (ax, ax)
    xi = staggered_grid(xv)
    hp = host_particles(xv, xi, 48, 12, seed=9)
    x, y = hp.coords
    phases = np.where(hp.index == 1, np.where(x < y, 1.0, 2.0), np.nan)
    out = O.update_phase_ratios(hp, phases, 2)
    dims = O.phase_ratio_shapes(hp.ni)
    for k in ("vertex", "center", "Vx", "Vy"):
        r = ratios_by_location(out[k], dims[k], 2)
        assert np.all(np.abs(r.sum(axis=1) - 1.0) < 1e-12), k
    # far from the diagonal a location is pure
    c = ratios_by_location(out["center"], dims["center"], 2).reshape(n, n, 2, order="F")
    assert c[0, n - 1, 0] == 1.0 and c[n - 1, 0, 1] == 1.0


def test_hand_computed_weights():
    """One cell of width 1/8 with two particles: centre, vertex and face
    ratios from weights worked out by hand.  The grid's vertices and centres
    are exact dyadics (node k of the ghosted centre axis at (k - 1/2)/8)."""
    ax = Axis.lin(0.0, 1.0, 9)
    ext = Axis(np.arange(10) / 8.0 - 1 / 16, "range")
    xv = (ax, ax)
    xi = ((ax, ext), (ext, ax))
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    h = 1.0 / 8
    # cell (2, 3): x in [1/4, 3/8], y in [3/8, 1/2]; centre (5/16, 7/16)
    cx, cy = 5 / 16, 7 / 16
    ea = place(hp, (2, 3), 0, (cx + 1 / 64, cy))                # phase 1
    eb = place(hp, (2, 3), 2, (cx - 1 / 32, cy + 1 / 32))       # phase 2
    phases = np.full(hp.S, np.nan)
    phases[ea], phases[eb] = 1.0, 2.0
    out = O.update_phase_ratios(hp, phases, 2)
    dims = O.phase_ratio_shapes(hp.ni)
    # centre: w_a = (1 - (1/64)/h) * 1 = 7/8; w_b = (1 - (1/32)/h)^2 = 9/16
    wa, wb = 7 / 8, 9 / 16
    inv = 1.0 / (wa + wb)
    c = ratios_by_location(out["center"], dims["center"], 2).reshape(8, 8, 2, order="F")
    assert c[2, 3, 0] == wa * inv and c[2, 3, 1] == wb * inv
    assert np.isnan(c[0, 0, 0])  # no particles: 0 * inv(0)
    # vertex (2, 4) at (1/4, 1/2): |dx| = 1/16 + 1/64 = 5/64 >= h/2? no (h/2 = 1/16 = 4/64):
    # a is 5/64 away in x, outside half a cell; b is 1/32 in x and 1/32 in y: inside
    v = ratios_by_location(out["vertex"], dims["vertex"], 2).reshape(9, 9, 2, order="F")
    assert v[2, 4, 0] == 0.0 and v[2, 4, 1] == 1.0
    # Vx face of cell (2, 3) at x = 3/8 (index (3, 3)): a is 3/64 away (inside),
    # b is 3/32 + ... outside; the neighbour cell (3, 3) is empty
    f = ratios_by_location(out["Vx"], dims["Vx"], 2).reshape(9, 8, 2, order="F")
    assert f[3, 3, 0] == 1.0 and f[3, 3, 1] == 0.0
    # faces without any particle within half a cell are 0, not NaN
    assert f[6, 6, 0] == 0.0 and f[6, 6, 1] == 0.0


def test_relabelling_swaps_planes():
    xv, xi = grid2d(20, 16)
    hp = host_particles(xv, xi, 32, 5, seed=4)
    rng = np.random.default_rng(2)
    ph = np.where(hp.index == 1, rng.integers(1, 3, hp.S), np.nan).astype(float)
    sw = np.where(np.isnan(ph), np.nan, 3.0 - ph)
    a = O.update_phase_ratios(hp, ph, 2)
    b = O.update_phase_ratios(hp, sw, 2)
    dims = O.phase_ratio_shapes(hp.ni)
    for k in a:
        ra = ratios_by_location(a[k], dims[k], 2)
        rb = ratios_by_location(b[k], dims[k], 2)
        assert np.array_equal(ra[:, 0], rb[:, 1], equal_nan=True), k
        assert np.array_equal(ra[:, 1], rb[:, 0], equal_nan=True), k


def _nearest(hp, e_new, I):
    """brute-force nearest active particle among the 3^D neighbour cells"""
    best, be = np.inf, -1
    pn = np.array([c[e_new] for c in hp.coords])
    for off in np.ndindex(*([3] * hp.D)):
        J = tuple(i + o - 1 for i, o in zip(I, off))
        if any(j < 0 or j >= n for j, n in zip(J, hp.ni)):
            continue
        c = int(np.ravel_multi_index(J, hp.ni, order="F"))
        for k in range(hp.K):
            e = c + hp.C * k
            if e == e_new or not hp.index[e]:
                continue
            p = np.array([q[e] for q in hp.coords])
            d = np.sqrt(np.sum((p - pn) ** 2))
            if d < best:
                best, be = d, e
    return be


def test_inject_phase_fills_every_short_quadrant_and_copies_nearest_phase():
    xv, xi = grid2d(8)
    hp = O.HostParticles(xv, 16, xi)
    clear(hp)
    h = 1.0 / 8
    rng = np.random.default_rng(5)
    phases = np.full(hp.S, np.nan)
    # cell (3, 3): its first quadrant holds 4 particles, the others none
    slot = 0
    for _ in range(4):
        e = place(hp, (3, 3), slot, (3 * h + rng.uniform(0.05, 0.45) * h,
                                      3 * h + rng.uniform(0.05, 0.45) * h))
        phases[e] = 1.0
        slot += 1
    # neighbours with other phases
    for cell, ph in (((4, 3), 2.0), ((3, 4), 3.0), ((2, 2), 4.0)):
        for k in range(3):
            e = place(hp, cell, k, ((cell[0] + rng.uniform(0.1, 0.9)) * h,
                                    (cell[1] + rng.uniform(0.1, 0.9)) * h))
            phases[e] = ph
    c = int(np.ravel_multi_index((3, 3), hp.ni, order="F"))
    before = hp.index.copy()
    plain = hp.copy()
    O.inject_particles(plain, [], 8, seed=3, counter=0)
    new_plain = np.nonzero((plain.index == 1) & (before == 0))[0]
    # the plain injection breaks at the full first quadrant of cell (3, 3)
    assert not any(e % hp.C == c for e in new_plain)
    st = O.inject_particles_phase(hp, phases, [], [], 8, seed=3, counter=0)
    new_all = np.nonzero((hp.index == 1) & (before == 0))[0]
    assert st["injected"] == len(new_all)
    new = [e for e in new_all if e % hp.C == c]
    assert len(new) == 6  # quadrants 2, 3, 4 topped up to cld(8, 4) = 2
    # each quadrant now holds at least 2 particles
    x, y = hp.coords[0], hp.coords[1]
    idx = [c + hp.C * k for k in range(hp.K) if hp.index[c + hp.C * k]]
    for q in range(4):
        vx, vy = 3 * h + (q & 1) * h / 2, 3 * h + (q >> 1) * h / 2
        n = sum(1 for e in idx if vx < x[e] < vx + h / 2 and vy < y[e] < vy + h / 2)
        assert n >= 2, q
    # phase of the nearest neighbour at the time of injection (brute force on
    # a replay: particles injected earlier in the same cell count as neighbours)
    replay = O.HostParticles(xv, 16, xi)
    for d in range(2):
        replay.coords[d][:] = hp.coords[d]
    # cell (3, 3) (1-based (4, 4)) is in the last of the four colours: the
    # particles injected into its neighbours came earlier
    replay.index[:] = before
    for e in new_all:
        if e % hp.C != c:
            replay.index[e] = 1
    for e in sorted(new, key=lambda e: (e // hp.C)):  # slots ascending
        replay.index[e] = 1
        src = _nearest(replay, e, (3, 3))
        assert phases[e] == phases[src]


def test_inject_phase_fields_interpolated_and_clamped():
    """Nodal field linear in x and y: a new particle gets the bilinear value
    (exact up to rounding) inside the corner extrema; a cell-centred field is
    read on the centre grid at the shifted index and clamped to its corners."""
    n = 8
    xv, xi = grid2d(n)
    hp = O.HostParticles(xv, 16, xi)
    clear(hp)
    h = 1.0 / n
    phases = np.full(hp.S, np.nan)
    for cell in ((2, 2), (2, 3), (3, 2), (6, 6), (0, 0)):
        for k in range(2):
            e = place(hp, cell, k, ((cell[0] + 0.2 + 0.5 * k) * h, (cell[1] + 0.3 + 0.4 * k) * h))
            phases[e] = 1.0
    X, Y = np.meshgrid(xv[0].x, xv[1].x, indexing="ij")
    nodal = np.asfortranarray(1.0 + 2.0 * X + 3.0 * Y)
    xc = O.centre_axes(hp)
    XC, YC = np.meshgrid(xc[0].x, xc[1].x, indexing="ij")
    centred = np.asfortranarray(np.sin(7 * XC) * np.cos(5 * YC))
    a_n = np.full(hp.S, np.nan)
    a_c = np.full(hp.S, np.nan)
    before = hp.index.copy()
    O.inject_particles_phase(hp, phases, [a_n, a_c], [nodal, centred], 12, seed=8, counter=1)
    new = np.nonzero((hp.index == 1) & (before == 0))[0]
    assert new.size > 0
    for e in new:
        x, y = hp.coords[0][e], hp.coords[1][e]
        assert abs(a_n[e] - (1 + 2 * x + 3 * y)) < 1e-13
        c = e % hp.C
        i, j = np.unravel_index(c, hp.ni, order="F")
        # centre-grid cell: shifted by the cell centre, clamped to 0..n-2
        ic = min(max(i - 1 if x < xv[0].x[i] + h / 2 else i, 0), n - 2)
        jc = min(max(j - 1 if y < xv[1].x[j] + h / 2 else j, 0), n - 2)
        corners = centred[ic:ic + 2, jc:jc + 2]
        assert corners.min() <= a_c[e] <= corners.max()
