"""Pins of the oracle's primitives against the reference's own known-answer
tests and published vectors (no GPU)."""
import numpy as np
import pytest

from oracle import oracle as O


# test/test_integrators.jl:3-80 -------------------------------------------
def test_rk2_first_stage_exact():
    assert O.first_stage(O.JP_RK2, 0.5, 0.1, (3.0, 4.0), (1.0, 2.0)) == (1.15, 2.2)
    assert O.first_stage(O.JP_RK2, 0.5, 0.1, (3.0, 4.0, 5.0), (1.0, 2.0, 3.0)) == (1.15, 2.2, 3.25)


def test_euler_first_stage_exact():
    assert O.first_stage(O.JP_EULER, 0.5, 0.1, (3.0, 4.0), (1.0, 2.0)) == (1.3, 2.4)
    assert O.first_stage(O.JP_EULER, 0.5, 0.1, (3.0, 4.0, 5.0), (1.0, 2.0, 3.0)) == (1.3, 2.4, 3.5)


def test_rk2_second_stage_approx():
    p1 = O.first_stage(O.JP_RK2, 0.5, 0.1, (3.0, 4.0), (1.0, 2.0))
    assert np.allclose(O.second_stage(0.5, 0.1, (3.0, 4.0), (3.1, 4.1), p1), (1.46, 2.61))
    p1 = O.first_stage(O.JP_RK2, 0.5, 0.1, (3.0, 4.0, 5.0), (1.0, 2.0, 3.0))
    assert np.allclose(O.second_stage(0.5, 0.1, (3.0, 4.0, 5.0), (3.1, 4.1, 5.1), p1),
                       (1.46, 2.61, 3.76))


def test_rk2_alpha_two_thirds():
    a = 2 / 3
    p1 = O.first_stage(O.JP_RK2, a, 0.1, (3.0, 4.0), (1.0, 2.0))
    assert np.allclose(p1, (1.2, 2.266666666))
    assert np.allclose(O.second_stage(a, 0.1, (3.0, 4.0), (3.1, 4.1), p1), (1.5075, 2.67416666))
    p1 = O.first_stage(O.JP_RK2, a, 0.1, (3.0, 4.0, 5.0), (1.0, 2.0, 3.0))
    assert np.allclose(p1, (1.2, 2.266666666, 3.3333333333333335))
    assert np.allclose(O.second_stage(a, 0.1, (3.0, 4.0, 5.0), (3.1, 4.1, 5.1), p1),
                       (1.5075, 2.67416666, 3.8408333333333333))


def test_rk2_second_stage_weights_closed_form():
    """second_stage = p + dt*((1 - 1/(2a)) v0 + 1/(2a) v1): checked for a
    spread of alpha against the exact rational value (a swapped weight or a
    missing 1/(2a) fails)."""
    from fractions import Fraction as Fr
    for a in (0.5, 2 / 3, 0.25, 0.9):
        dt, v0, v1, p = 0.125, 3.0, -5.0, 1.0
        out = O.second_stage(a, dt, (v0,), (v1,), (p,))[0]
        w1 = Fr(1, 2) / Fr(a)
        exact = Fr(p) + Fr(dt) * ((1 - w1) * Fr(v0) + w1 * Fr(v1))
        assert abs(out - float(exact)) <= 4 * abs(np.spacing(float(exact)))


def _fma(a, b, c):
    """correctly rounded a*b + c (exact rationals, then one rounding)"""
    from fractions import Fraction as Fr
    return float(Fr(a) * Fr(b) + Fr(c))


def _mul(a, b):
    from fractions import Fraction as Fr
    return float(Fr(a) * Fr(b))


def test_rk2_blended_fold_follows_muladdmacro():
    """Pins the fold of the blended second stage (src/Advection/RK2.jl:32-35),
    p + s*dt*((1 - half*inv(a))*v0 + half*inv(a)*v1) under @muladd.

    MuladdMacro (compat 0.2.4) folds a sum of products from the left: the
    first product stays a plain product and each later one is fused into the
    running sum; a product of more than two factors splits into (all but the
    last factor, the last factor); summands without a product are added
    first.  Its README documents exactly this:
      a031*k1 + a032*k2            -> muladd(a032, k2, a031*k1)
      bhat1*k1 + bhat4*k4 + ...    -> muladd(bhat10, k10, ... muladd(bhat4, k4, bhat1*k1))
      uprev + dt*(...)             -> muladd(dt, (...), uprev)
      update - dt*(...)            -> muladd(-dt, (...), update)
    So the stage is muladd(s*dt, muladd(half*inv(a), v1, (1 - half*inv(a))*v0), p)
    with 1 - half*inv(a) = muladd(-half, inv(a), 1) (the same value: half*inv(a)
    is exact).  On inputs where fusing the first product instead rounds
    differently, the oracle must give the documented value and not the other."""
    rng = np.random.default_rng(2024)
    differ = 0
    trials = 400
    for _ in range(trials):
        a = float(rng.uniform(0.05, 0.95))
        dt = float(rng.uniform(1e-3, 1.0))
        v0, v1, p = (float(x) for x in rng.uniform(-300, 300, 3))
        ia = 1.0 / a
        w0 = _fma(-0.5, ia, 1.0)
        w1 = _mul(0.5, ia)
        doc = _fma(dt, _fma(w1, v1, _mul(w0, v0)), p)        # MuladdMacro's fold
        other = _fma(dt, _fma(w0, v0, _mul(w1, v1)), p)      # first product fused
        got = O.second_stage(a, dt, (v0,), (v1,), (p,))[0]
        assert got == doc, (a, dt, v0, v1, p)
        differ += got != other
    # the two orders really differ on these inputs, so the check discriminates
    assert differ > trials // 20, differ


def test_first_stage_and_euler_folds():
    """p + s*a*dt*v -> muladd(s*a*dt, v, p) (RK2.jl:13) and p + s*dt*v ->
    muladd(s*dt, v, p) (Euler.jl:5), s = +1: one rounding of (a*dt) then one
    fused multiply-add; the unfused p + fl(a*dt*v) differs on these inputs."""
    rng = np.random.default_rng(7)
    differ = 0
    for _ in range(200):
        a = float(rng.uniform(0.05, 0.95))
        dt = float(rng.uniform(1e-3, 1.0))
        v, p = (float(x) for x in rng.uniform(-300, 300, 2))
        got = O.first_stage(O.JP_RK2, a, dt, (v,), (p,))[0]
        assert got == _fma(_mul(a, dt), v, p)
        differ += got != p + _mul(_mul(a, dt), v)
        got_e = O.first_stage(O.JP_EULER, a, dt, (v,), (p,))[0]
        assert got_e == _fma(dt, v, p)
    assert differ > 10


# test/test_interpolation_kernels.jl:47-60 --------------------------------
def test_lerp_exact():
    assert O.lerp([1.0, 2.0], [0.5]) == 1.5
    assert O.lerp([1.0, 2.0, 1.0, 2.0], [0.5, 0.5]) == 1.5
    assert O.lerp([1.0, 2.0] * 4, [0.5, 0.5, 0.5]) == 1.5


def test_lerp_corner_order_and_multilinearity():
    """n-linear interpolation of f(x,y,z) = 1 + 2x + 3y + 5z + 7xy + 11xz + 13yz
    + 17xyz from corner values in (x fastest) order is exact at dyadic t."""
    f = lambda x, y, z: 1 + 2 * x + 3 * y + 5 * z + 7 * x * y + 11 * x * z + 13 * y * z + 17 * x * y * z
    v3 = [f(i & 1, (i >> 1) & 1, (i >> 2) & 1) for i in range(8)]
    v2 = [f(i & 1, (i >> 1) & 1, 0) for i in range(4)]
    for t in [(0.25, 0.5, 0.75), (0.125, 0.875, 0.375), (0.0, 1.0, 0.5)]:
        assert O.lerp(v3, t) == f(*t)
        assert O.lerp(v2, t[:2]) == f(t[0], t[1], 0)


# Philox4x32-10 known-answer vectors (Random123 distribution, kat_vectors) --
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_kat(ctr, key, out):
    assert list(O.philox(ctr, key)) == out


def test_uniform_range_and_spread():
    u = np.array([O.uniform(7, c, s, d, 1, 0) for c in range(40) for s in range(8) for d in range(3)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.05
    assert len(np.unique(u)) == u.size


# find_parent_cell_bisection (src/Utils.jl:108-130) -----------------------
def test_bisection_brackets_every_seed():
    rng = np.random.default_rng(3)
    for x in (np.linspace(0, 1, 11), np.sort(np.concatenate([[0, 1], rng.random(30)]))):
        n = x.size
        for p in rng.uniform(x[0], x[-1], 200):
            for seed in range(1, n):
                k = O.bisect(p, x, seed)
                assert x[k - 1] <= p <= x[k]
                # strictly interior points have one bracket, whatever the seed
                if x[k - 1] < p < x[k]:
                    assert k == np.searchsorted(x, p, side="right")


def test_bisection_ties_are_valid_brackets():
    x = np.linspace(0, 1, 9)
    for j in range(x.size):
        for seed in range(1, x.size):
            k = O.bisect(x[j], x, seed)
            assert x[k - 1] <= x[j] <= x[k]
            assert k in (j, j + 1)
    # a tie resolves to the seed's own bracket when the seed brackets it
    assert O.bisect(x[3], x, 3) == 3 and O.bisect(x[3], x, 4) == 4


# cell_index (test/test_2D.jl:143-167) -----------------------------------
def test_cell_index_brackets():
    rng = np.random.default_rng(5)
    x = np.linspace(0, 1, 11)
    for px in rng.random(50):
        i = O.cell_index(px, x)
        assert x[i - 1] <= px < x[i]
    y = np.linspace(-1, 0, 11)
    for py in -rng.random(50):
        j = O.cell_index(py, y)
        assert y[j - 1] <= py < y[j]
