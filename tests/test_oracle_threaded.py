"""The threaded oracle build (the benchmark's CPU baseline) computes what the
sequential oracle computes: the same cells run in parallel only where the
reference runs them in parallel (one colour phase of move and inject, the
independent advection, particle2grid and grid2particle loops)."""
import numpy as np
import pytest

from oracle import oracle as O

from helpers import cellular_V, cfl_dt, grid2d, grid3d, host_particles, stream_V


@pytest.fixture
def threaded():
    O.set_threaded(True)
    try:
        yield O.threads()
    finally:
        O.set_threaded(False)


def run_steps(hp, V, dt, steps, minx, D):
    T = np.where(hp.index == 1, np.arange(hp.S, dtype=np.float64), np.nan)
    stats, F = [], None
    for it in range(steps):
        O.advection(hp, V, O.JP_RK2, 0.5, dt)
        stats.append(O.move_particles(hp, [T]))
        stats.append(O.inject_particles(hp, [T], minx, seed=3, counter=it))
        F = O.particle2grid(hp, T)
    return T, F, stats


@pytest.mark.parametrize("D", [2, 3])
def test_threaded_oracle_equals_sequential(threaded, D):
    assert threaded >= 1
    if D == 2:
        xv, xi = grid2d(48, 40)
        V = stream_V(xi)
        K, npq, minx = 48, 6, 12
    else:
        xv, xi = grid3d(10)
        V = cellular_V(xi)
        K, npq, minx = 54, 4, 14
    dt = cfl_dt(xv, V)
    hp_t = host_particles(xv, xi, K, npq, seed=8)
    out_t = run_steps(hp_t, V, dt, 3, minx, D)
    O.set_threaded(False)
    hp_s = host_particles(xv, xi, K, npq, seed=8)
    out_s = run_steps(hp_s, V, dt, 3, minx, D)
    assert out_t[2] == out_s[2]
    assert sum(s.get("moved", 0) for s in out_s[2]) > 0
    assert np.array_equal(hp_t.index, hp_s.index)
    for d in range(D):
        assert np.array_equal(hp_t.coords[d], hp_s.coords[d], equal_nan=True)
    assert np.array_equal(out_t[0], out_s[0], equal_nan=True)
    assert np.array_equal(out_t[1], out_s[1], equal_nan=True)
