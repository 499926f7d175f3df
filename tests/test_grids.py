"""Host grid helpers (CPU): the README's expand_range and the affine maps the
FP32-split advection derives from the velocity axes.

README.md:48-55 rounds the ghost ends of the expanded centre axes to 5
significant digits.  The resulting axes are uniform, but their node spacing
and origin differ from the standard staggering (node k at (k - 1/2) h), by
up to 0.012 cell at 256^2.  The split kernels map a position at cell I plus
offset u to the node coordinate (I + u) r + o with r = h / h_node and
o = (x_vertex[0] - x_node[0]) / h_node; these tests pin what those maps are.
"""
import numpy as np
import pytest

from justpic_hip.configs import stream2d
from justpic_hip.grids import Axis, expand_range, round_sigdigits, steprange


def affine(vertex, node):
    h = vertex.x[1] - vertex.x[0]
    hn = (node.x[-1] - node.x[0]) / (node.n - 1)
    return h / hn, (vertex.x[0] - node.x[0]) / hn


def test_round_sigdigits():
    # Julia: round(-0.001953125; sigdigits=5) == -0.0019531, round(1.001953125; sigdigits=5) == 1.002
    assert round_sigdigits(-0.001953125, 5) == -0.0019531
    assert round_sigdigits(1.001953125, 5) == 1.002
    assert round_sigdigits(0.0, 5) == 0.0


def test_readme_expand_range_endpoints():
    n = 256
    dx = 1.0 / n
    xc = Axis.range(dx / 2, 1 - dx / 2, n)
    e = expand_range(xc, 5, lin=False)
    assert e.n == n + 2
    assert e.x[0] == -0.0019531 and e.x[-1] == 1.002


@pytest.mark.parametrize("n,shift", [(128, 0.0008), (256, 0.012)])
def test_readme_axes_are_off_standard_staggering(n, shift):
    """The README ghost axes are uniform but not the standard staggering; the
    largest node displacement against (k - 1/2) h is about `shift` cells."""
    cfg = stream2d(n, readme=True)
    xv = cfg["xv"]
    h = xv[0].x[1] - xv[0].x[0]
    for c, d in ((0, 1), (1, 0)):
        ax = cfg["xi_vel"][c][d]
        std = xv[d].x[0] + (np.arange(ax.n) - 0.5) * h
        dev = np.max(np.abs(ax.x - std)) / h
        assert shift * 0.8 < dev < shift * 1.2, dev
        # uniform to far better than the 1e-9 of a cell the split form accepts
        hn = (ax.x[-1] - ax.x[0]) / (ax.n - 1)
        assert np.max(np.abs(ax.x - (ax.x[0] + np.arange(ax.n) * hn))) <= 1e-12 * hn
        r, o = affine(xv[d], ax)
        assert abs(r - 1) > 1e-6 and abs(o - 0.5) > 1e-7
        # the affine map reproduces every node: node k sits at (I + u) = (k - o)/r
        k = np.arange(ax.n)
        X = (k - o) / r
        assert np.max(np.abs(xv[d].x[0] + X * h - ax.x)) <= 1e-12


def test_standard_staggering_maps_are_exact():
    """Test-suite grids (LinRange ghost axes): r = 1 and o = 0 or 1/2 up to
    rounding of the coordinates."""
    cfg = stream2d(64, readme=False)
    xv = cfg["xv"]
    for c in range(2):
        for d in range(2):
            r, o = affine(xv[d], cfg["xi_vel"][c][d])
            assert abs(r - 1) < 1e-13
            assert abs(o - (0.0 if c == d else 0.5)) < 1e-12


def test_steprange_is_correctly_rounded():
    x = steprange(-0.0019531, 1.002, 258)
    assert x[0] == -0.0019531 and x[-1] == 1.002
    from fractions import Fraction
    a, b = Fraction(-0.0019531), Fraction(1.002)
    for j in (1, 77, 200, 256):
        assert x[j] == float(a + (b - a) * j / 257)
