"""Pins of the oracle's grid2particle_flip!, particle2centroid! and
centroid2particle! (CPU).

* The reference's known answers (test/test_interpolation_kernels.jl:62-180):
  a linear field y is reproduced on the particles by grid2particle_flip! with
  F0 = F and by centroid2particle!, and particle2centroid! of it stays within
  its error bound.
* Closed forms: FLIP with α = 0 and F0 = F leaves Fp unchanged bit for bit;
  α = 1 is pure PIC (the nodal interpolation); a constant shift c of F over F0
  adds c; particle2centroid! of two particles at dyadic offsets equals the
  weighted mean computed by hand, with bilinear weights in the scalar method
  and inverse squared distances in the 2D tuple method.
"""
import numpy as np
import pytest

from justpic_hip.grids import Axis
from oracle import oracle as O

from helpers import clear, grid2d, grid3d, host_particles, place


def case(D, npq=2):
    if D == 2:
        xv, xi = grid2d(4)
        hp = host_particles(xv, xi, 4 * npq + 2, npq, seed=11)
    else:
        xv, xi = grid3d(4)
        hp = host_particles(xv, xi, 8 * npq + 2, npq, seed=11)
    return xv, xi, hp


def nodal(xv, f):
    mesh = np.meshgrid(*[a.x for a in xv], indexing="ij")
    return np.asfortranarray(f(*mesh))


@pytest.mark.parametrize("D", [2, 3])
def test_flip_reproduces_linear_field(D):
    """test/test_interpolation_kernels.jl:86-99 (2D) / :150-166 (3D)."""
    xv, xi, hp = case(D)
    T = nodal(xv, lambda *X: X[1])
    pT = O.grid2particle(hp, T, np.full(hp.S, np.nan))
    m = hp.index == 1
    assert np.max(np.abs(pT[m] - hp.coords[1][m])) < 1e-14
    O.grid2particle_flip(hp, [pT], [T], [T], 0.0)
    assert np.max(np.abs(pT[m] - hp.coords[1][m])) < 1e-14


@pytest.mark.parametrize("D", [2, 3])
def test_flip_closed_forms(D):
    xv, xi, hp = case(D, npq=3)
    rng = np.random.default_rng(4)
    F0 = nodal(xv, lambda *X: np.cos(3 * X[0]) + X[-1] ** 2)
    Fp = np.where(hp.index == 1, rng.random(hp.S), np.nan)
    m = hp.index == 1
    # alpha = 0, F = F0: unchanged bit for bit
    a = Fp.copy()
    O.grid2particle_flip(hp, [a], [F0], [F0], 0.0)
    assert np.array_equal(a, Fp, equal_nan=True)
    # alpha = 1: pure PIC
    b = Fp.copy()
    O.grid2particle_flip(hp, [b], [F0], [np.zeros_like(F0)], 1.0)
    pic = O.grid2particle(hp, F0, np.full(hp.S, np.nan))
    assert np.max(np.abs(b[m] - pic[m])) < 1e-14
    # constant shift of F over F0 with alpha = 0: Fp + c
    c = 0.625
    d = Fp.copy()
    O.grid2particle_flip(hp, [d], [F0 + c], [F0], 0.0)
    assert np.max(np.abs(d[m] - (Fp[m] + c))) < 1e-14
    # inactive slots are skipped
    assert np.all(np.isnan(d[~m]))
    # blend: alpha * PIC + (1 - alpha) * FLIP
    e = Fp.copy()
    O.grid2particle_flip(hp, [e], [F0 + c], [F0], 0.25)
    pic_c = O.grid2particle(hp, F0 + c, np.full(hp.S, np.nan))
    assert np.max(np.abs(e[m] - (0.25 * pic_c[m] + 0.75 * (Fp[m] + c)))) < 1e-13


@pytest.mark.parametrize("D", [2, 3])
def test_centroid2particle_reproduces_linear_field(D):
    """test/test_interpolation_kernels.jl:107-110: a linear centroid field y
    interpolated back to every particle (also where the centre cell is
    clamped at the boundary: linear extrapolation)."""
    xv, xi, hp = case(D)
    xc = O.centre_axes(hp)
    Tc = nodal(xc, lambda *X: X[1])
    pT = np.full(hp.S, np.nan)
    O.centroid2particle(hp, [pT], [Tc])
    m = hp.index == 1
    assert np.max(np.abs(pT[m] - hp.coords[1][m])) < 1e-14


@pytest.mark.parametrize("D", [2, 3])
def test_particle2centroid_error_bound(D):
    """test/test_interpolation_kernels.jl:112-116."""
    xv, xi, hp = case(D)
    xc = O.centre_axes(hp)
    Tc = nodal(xc, lambda *X: X[1]).ravel(order="F")
    pT = np.where(hp.index == 1, hp.coords[1], np.nan)
    for tup in (False, True):
        Tc2 = O.particle2centroid(hp, [pT], tuple_method=tup)[0]
        assert np.linalg.norm(Tc2 - Tc) / Tc.size < 1e-1


def test_particle2centroid_hand_computed():
    """Cell (1, 1) of a 1/4 grid with exact centres: two particles with
    values 3 and 11; scalar method: bilinear weights (7/8)(1) and (3/4)(1/2);
    2D tuple method: inverse squared distances 1/(1/32)^2 and 1/(sqrt(2)/16)^2."""
    ax = Axis.lin(0.0, 1.0, 5)
    ext = Axis(np.arange(6) / 4.0 - 1 / 8, "range")
    xv = (ax, ax)
    xi = ((ax, ext), (ext, ax))
    hp = O.HostParticles(xv, 4, xi)
    clear(hp)
    cx = cy = 3 / 8
    ea = place(hp, (1, 1), 0, (cx + 1 / 32, cy))
    eb = place(hp, (1, 1), 1, (cx - 1 / 16, cy + 1 / 16))
    Fp = np.full(hp.S, np.nan)
    Fp[ea], Fp[eb] = 3.0, 11.0
    F = O.particle2centroid(hp, [Fp])[0].reshape(4, 4, order="F")
    wa = (1 - (1 / 32) * 4) * 1.0                  # 7/8
    wb = (1 - (1 / 16) * 4) * (1 - (1 / 16) * 4)   # 9/16
    assert F[1, 1] == (wb * 11.0 + wa * 3.0) / (wa + wb)
    Ft = O.particle2centroid(hp, [Fp], tuple_method=True)[0].reshape(4, 4, order="F")
    da, db = 1 / 32, np.sqrt(2 * (1 / 16) ** 2)
    ia, ib = 1.0 / (da * da), 1.0 / (db * db)
    assert Ft[1, 1] == (ib * 11.0 + ia * 3.0) * (1.0 / (ia + ib))
    assert np.isnan(F[0, 0]) and np.isnan(Ft[3, 3])  # empty cells: 0/0
