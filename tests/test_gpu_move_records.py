"""The move's staged records with 0, 1 and several fields, FP64 and
FP32-split, 2D and 3D, against the oracle bit for bit (runs on an MI355X).

Classify stages each mover as one record (coordinates, then the fields),
packed per cell; the scatter gathers the records by origin.  The record
size follows the form and the field count (the 2D split record with one
field is the 16-byte case read and written whole), so each count is a
layout of its own.  The states are dyadic (exact in float32) with crowded
cells (drops) and moves in every direction.
"""
import numpy as np
import pytest

import justpic_hip as jp
from oracle import oracle as O

from test_gpu_move_far import jumped_state
from test_gpu_move_tiles import shifted_state
from test_gpu_parity import assert_field, assert_state_equal, to_dev
from test_gpu_split import split_state

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-move-records]", *a, flush=True)


def fields_for(hp, nf, seed):
    """nf fields: the particle id, then random values (NaN in empty slots)."""
    rng = np.random.default_rng(seed)
    act = hp.index == 1
    out = []
    for f in range(nf):
        v = np.arange(hp.S, dtype=np.float64) if f == 0 else rng.random(hp.S) - 0.5
        out.append(np.where(act, v, np.nan))
    return out


def state(D, seed):
    if D == 2:
        return shifted_state(64, 32, 16, 0.75, seed=seed, drift_y=0.4)
    return jumped_state((8, 8, 8), 12, 0.7, 0.0, 1, seed=seed)


@pytest.mark.parametrize("nf", [0, 1, 3])
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("form", ["fp64", "split"])
def test_move_records(ctx, D, nf, form):
    say(form, D, "fields", nf)
    hp = state(D, seed=10 * D + nf)
    vals = fields_for(hp, nf, seed=nf)
    dev = [jp.CellArray(ctx, hp.ni, hp.K) for _ in range(nf)]
    for d, v in zip(dev, vals):
        d.copy_from(v)
    p = to_dev(ctx, hp)
    if form == "split":
        p = jp.to_split(p)
    st_d = jp.move_particles_(p, tuple(dev), stats=True)
    st_o = O.move_particles(hp, vals)
    say("stats", st_d, st_o)
    assert st_o["moved"] > 0 and st_o["dropped"] > 0
    assert st_d == st_o
    if form == "fp64":
        assert_state_equal(p, hp, exact=True)
        for d, v in zip(dev, vals):
            assert_field(d.data.flat(), v, exact=True)
    else:
        coords, index = split_state(p)
        assert np.array_equal(index, hp.index)
        for d in range(D):
            assert np.array_equal(coords[d], hp.coords[d], equal_nan=True)
        for d, v in zip(dev, vals):
            assert np.array_equal(d.data.flat(), v, equal_nan=True)
