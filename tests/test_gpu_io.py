"""Host layout and checkpoints (runs on an MI355X).

The device CellArray (C, K, 1) and the CPU CellArray (1, K, C) hold the same
elements permuted by (3, 2, 1) (src/CellArrays/conversion.jl:32-43; the
reference's own check: test/test_save_load.jl:154-159 compares the sizes of
the GPU data with permutedims of the CPU data).  The GPU transpose must put
slot k of cell c of the device layout at [1, k, c] of the CPU layout, and a
checkpoint must restore a state that steps on exactly like the original.
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.configs import stream2d

from helpers import grid3d

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.uint8, np.float32, np.float64])
@pytest.mark.parametrize("C,K", [(1, 1), (63, 48), (1000, 54), (4097, 64)])
def test_cpu_layout_is_the_permutation(ctx, dtype, C, K):
    rng = np.random.default_rng(C * K)
    dev = (rng.integers(0, 255, C * K) if dtype == np.uint8 else rng.random(C * K)).astype(dtype)
    ca = jp.CellArray(ctx, (C,), K, dtype=dtype)
    ca.copy_from(dev)
    h = jp.cellarray_to_host(ca)
    assert h.shape == (1, K, C)
    # GPU data (C, K, 1) permuted by (3, 2, 1) is the CPU data (1, K, C)
    want = np.transpose(dev.reshape((C, K, 1), order="F"), (2, 1, 0))
    assert np.array_equal(h, want)
    back = jp.cellarray_from_host(ctx, h, (C,))
    assert np.array_equal(back.data.flat(), dev)


def test_particles_host_copy_round_trip(ctx):
    xv, xi = grid3d(6)
    p = jp.init_particles(ctx, 16, 24, 4, *xi, seed=2)
    h = jp.particles_to_host(p)
    coords, index = p.host_state()
    K, C = p.max_xcell, p.C
    assert h["index"].dtype == np.bool_ and h["index"].shape == (1, K, C)
    for d in range(3):
        assert h["coords"][d].shape == (1, K, C)
        # slot k of cell c (0-based) sits at [0, k, c]
        assert np.array_equal(h["coords"][d][0].ravel(order="F"), coords[d].reshape(K, C).T.ravel(),
                              equal_nan=True)
    q = jp.particles_from_host_layout(ctx, h)
    c2, i2 = q.host_state()
    assert np.array_equal(i2, index)
    for d in range(3):
        assert np.array_equal(c2[d], coords[d], equal_nan=True)


def test_checkpoint_restores_a_state_that_steps_identically(ctx, tmp_path):
    cfg = stream2d(48, readme=True)
    p = jp.init_particles(ctx, 24, 48, 12, *cfg["xi_vel"], seed=5)
    V = tuple(jp.DeviceArray.from_host(ctx, v) for v in cfg["V"])
    T = jp.CellArray(ctx, p.ni, p.max_xcell)
    _, idx = p.host_state()
    T.copy_from(np.where(idx == 1, np.random.default_rng(1).random(idx.size), np.nan))
    phases = jp.CellArray(ctx, p.ni, p.max_xcell)
    phases.copy_from(np.where(idx == 1, np.random.default_rng(2).integers(1, 4, idx.size),
                              np.nan))
    for _ in range(3):
        jp.advection_(p, jp.RungeKutta2(), V, cfg["dt"])
        jp.move_particles_(p, (T, phases))
        jp.inject_particles_(p, (T, phases))
    f = jp.checkpointing_particles(str(tmp_path), p, phases=phases, t=0.25, dt=cfg["dt"],
                                   particle_args=(T,))
    assert f.endswith("particles_checkpoint.npz")
    r = jp.load_checkpoint(f, ctx)
    q, Tq, ph_q = r["particles"], r["particle_args"][0], r["phases"]
    assert r["time"] == 0.25 and r["timestep"] == cfg["dt"]
    assert q.inject_counter == p.inject_counter and q.seed == p.seed
    for _ in range(2):
        for s, Ts, phs in ((p, T, phases), (q, Tq, ph_q)):
            jp.advection_(s, jp.RungeKutta2(), V, cfg["dt"])
            jp.move_particles_(s, (Ts, phs))
            jp.inject_particles_(s, (Ts, phs))
    ca, ia = p.host_state()
    cb, ib = q.host_state()
    assert np.array_equal(ia, ib)
    for d in range(2):
        assert np.array_equal(ca[d], cb[d], equal_nan=True)
    assert np.array_equal(T.data.flat(), Tq.data.flat(), equal_nan=True)
    assert np.array_equal(phases.data.flat(), ph_q.data.flat(), equal_nan=True)
