"""Host-side grid axes.

JustPIC keeps, for every axis of the vertex, center and staggered velocity
grids, the coordinates (``particles.xvi``, ``xci``, ``xi_vel``), the spacing
(``particles.di``) and the inverse spacing (``particles._di``).  The spacing
rule depends on how the user built the axis
(src/Particles/particles_utils.jl:72-85 for vectors, :132-146 for ranges):

* a range axis stores the scalar ``x[2] - x[1]`` and uses it for every cell;
* a vector axis stores ``diff(x)``.

:class:`Axis` materialises both rules as arrays of ``n-1`` spacings so the HIP
library has one code path.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


def linrange(a: float, b: float, n: int) -> np.ndarray:
    """Values of Julia's ``LinRange(a, b, n)``: ``(1-t)*a + t*b`` with
    ``t = j/(n-1)`` in Float64 (Julia Base ``lerpi``)."""
    a = float(a)
    b = float(b)
    if n == 1:
        return np.array([a])
    j = np.arange(n, dtype=np.float64)
    t = j / float(n - 1)
    return (1.0 - t) * a + t * b


def steprange(a: float, b: float, n: int) -> np.ndarray:
    """Correctly rounded values of ``a + j*(b-a)/(n-1)`` in exact rational
    arithmetic: the values Julia's TwicePrecision ``range(a, b, length=n)``
    produces for these inputs up to its internal rounding (the exact Julia
    values cannot be checked without Julia; both oracle and GPU consume the
    same materialised array, so parity does not depend on it)."""
    fa, fb = Fraction(float(a)), Fraction(float(b))
    if n == 1:
        return np.array([float(a)])
    step = (fb - fa) / (n - 1)
    return np.array([float(fa + step * j) for j in range(n)], dtype=np.float64)


def round_sigdigits(x: float, sig: int) -> float:
    """Julia ``round(x; sigdigits=sig)`` (round-half-even on the decimal digits)."""
    if x == 0.0 or not math.isfinite(x):
        return x
    return float(f"{x:.{sig - 1}e}")


class Axis:
    """One grid axis with the reference's spacing rule.

    ``kind='range'``: scalar spacing ``x[1]-x[0]`` replicated (range grids);
    ``kind='vector'``: ``diff(x)`` (vector grids).
    """

    def __init__(self, x, kind: str = "range"):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        if x.ndim != 1 or x.size < 2:
            raise ValueError("an axis needs at least two points")
        if not np.all(np.diff(x) > 0):
            raise ValueError("axis coordinates must be strictly increasing")
        if kind not in ("range", "vector"):
            raise ValueError(f"unknown axis kind {kind!r}")
        self.kind = kind
        self.x = x
        if kind == "range":
            self.dx = np.full(x.size - 1, x[1] - x[0], dtype=np.float64)
        else:
            self.dx = np.ascontiguousarray(np.diff(x))
        self.inv_dx = 1.0 / self.dx
        self.n = int(x.size)

    def __len__(self):
        return self.n

    @property
    def first(self):
        return float(self.x[0])

    @property
    def last(self):
        return float(self.x[-1])

    def inner(self):
        """x[2:end-1] keeping the kind (the reference slices velocity axes
        to obtain cell centers, src/Particles/particles_utils.jl:56-70)."""
        return Axis(self.x[1:-1], self.kind)

    @classmethod
    def lin(cls, a, b, n):
        return cls(linrange(a, b, n), "range")

    @classmethod
    def range(cls, a, b, n):
        return cls(steprange(a, b, n), "range")

    def __repr__(self):
        return f"Axis(n={self.n}, kind={self.kind}, [{self.first}, {self.last}])"


def compute_dx(ax: Axis) -> Axis:
    """The axis with the spacing rule of ``compute_dx``
    (src/Particles/utils.jl:46-49), which ``semilagrangian_advection!`` uses
    for its vertex and velocity grids: ``x[3] - x[2]`` (1-based) for a range,
    ``diff(x)`` for a vector."""
    o = Axis(ax.x, ax.kind)
    if ax.kind == "range":
        if ax.n < 3:
            raise ValueError("compute_dx of a range needs at least three points")
        o.dx = np.full(ax.n - 1, ax.x[2] - ax.x[1], dtype=np.float64)
        o.inv_dx = 1.0 / o.dx
    return o


def expand_range(ax: Axis, sigdigits: int | None = None, lin: bool = True) -> Axis:
    """One ghost point on each side of a center axis.

    ``sigdigits=None, lin=True`` is the test-suite helper
    (test/test_2D.jl:35-42: ``LinRange(x1-dx, x2+dx, n+2)``);
    ``sigdigits=5, lin=False`` is the README helper (README.md:48-55:
    ``range(round(x1-dx; sigdigits=5), round(x2+dx; sigdigits=5), length=n+2)``).
    Vector axes use ``vcat(x1-dx_left, x, xend+dx_right)`` (test/test_2D.jl:44-51).
    """
    x = ax.x
    if ax.kind == "vector":
        xl = x[0] - (x[1] - x[0])
        xr = x[-1] + (x[-1] - x[-2])
        return Axis(np.concatenate([[xl], x, [xr]]), "vector")
    dx = x[1] - x[0]
    xI = x[0] - dx
    xF = x[-1] + dx
    if sigdigits is not None:
        xI = round_sigdigits(xI, sigdigits)
        xF = round_sigdigits(xF, sigdigits)
    n = ax.n + 2
    return Axis.lin(xI, xF, n) if lin else Axis.range(xI, xF, n)


def staggered_grid(xv_axes, center_axes=None, readme=False):
    """Velocity grids of the reference examples: component c lives on the
    vertex axis along c and on the expanded center axes elsewhere
    (README.md:42-44; test/test_3D.jl grid_vx/vy/vz).

    Returns ``xi_vel`` as a tuple of D tuples of :class:`Axis`.
    """
    D = len(xv_axes)
    if center_axes is None:
        center_axes = []
        for ax in xv_axes:
            if ax.kind == "range":
                dx = ax.x[1] - ax.x[0]
                mk = Axis.range if readme else Axis.lin
                center_axes.append(mk(ax.first + dx / 2, ax.last - dx / 2, ax.n - 1))
            else:
                center_axes.append(Axis((ax.x[:-1] + ax.x[1:]) / 2, "vector"))
    ext = [expand_range(a, 5, lin=False) if readme else expand_range(a) for a in center_axes]
    return tuple(
        tuple(xv_axes[d] if d == c else ext[d] for d in range(D)) for c in range(D)
    )
