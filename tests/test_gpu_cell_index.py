"""cell_index on the GPU against the oracle (runs on an MI355X).

cell_index(x, xv) = abs(trunc((x - xv[1]) / |xv[2] - xv[1]|)) + 1
(src/particles.jl:205-224); the oracle's version is pinned by the
reference's bracket tests (test/test_2D.jl:143-167, tests/test_oracle_kats.py).
"""
import numpy as np
import pytest

import justpic_hip as jp
from justpic_hip.grids import Axis
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("origin", [0.0, -1.25, 3.0])
def test_cell_index_matches_oracle(ctx, origin):
    rng = np.random.default_rng(4)
    axes = (Axis.lin(origin, origin + 2.0, 41), Axis.lin(origin, origin + 1.0, 17),
            Axis.lin(0.0, 3.0, 13))
    n = 20000
    x = []
    for a in axes:
        v = rng.uniform(a.first, a.last, n)
        v[:50] = a.x[rng.integers(0, a.n, 50)]          # exactly on vertices
        v[50:60] = a.first                                # the first vertex
        v[60:70] = np.nextafter(a.x[5], -np.inf)          # just below a vertex
        x.append(v)
    xd = tuple(jp.DeviceArray.from_host(ctx, v) for v in x)
    out = jp.cell_index(xd, axes)
    for d, a in enumerate(axes):
        got = out[d].to_host()
        want = np.array([O.cell_index(v, a.x) for v in x[d]], dtype=np.int64)
        assert np.array_equal(got, want), d
        assert got.min() >= 1 and got.max() <= a.n  # n - 1 cells; the last vertex gives n
    # one dimension alone
    (o1,) = jp.cell_index(xd[:1], axes[:1])
    assert np.array_equal(o1.to_host(), out[0].to_host())
