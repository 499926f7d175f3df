"""Pins of the oracle's MQS / LinP velocity interpolation
(advection_MQS!, advection_LinP!) and of semilagrangian_advection! (CPU).

* The reference's own test (test/test_2D.jl:169-228, refined grid): the
  velocity at p = (0.22, 0.48) of a linear field (values worked by hand:
  exact for the linear scheme, not for MQS and LinP on this grid), and one RK2 backtrack of F0 = x + 2y under a constant velocity (0.08, 0.02)
  with dt = 1 gives F[3, 3] = F0(0.22, 0.48) = 1.18 by MQS and LinP.  The
  NTuple linear method takes its spacing at the node (backtracking.jl:79),
  which on this refined grid gives 0.18 + 2 * 0.48 = 1.14 instead.
* Closed forms, computed here from the analytic fields (the velocity is read
  back from one Euler step, v = (p' - p) / dt):
  - linear fields are reproduced by every scheme on uniform grids;
  - MQS of x^2 along its own axis is x^2 + h^2/4 where the stencil is
    interior (the lerp error h^2 t(1-t) plus the correction h^2 (t-1/2)^2),
    the plain lerp elsewhere, and a field varying across the direction gets
    no correction;
  - LinP of x^2 is 2/3 (x^2 + h^2 t(1-t)) + 1/3 (x^2 + h^2 s(1-s) + h^2/4),
    s measured from the pressure point below p;
  - the reference's 3D index patterns: the z component of MQS corrects along
    x, not z (MQS.jl:67-71, 126-152), x and y take the extra nodes of both
    planes from plane k (MQS.jl:58-64), and the z component of LinP lerps its
    two pressure values with the y coordinate (augment_offset rows and
    swap_F, advection_LinP.jl:326-391).
* semilagrangian_advection!: zero velocity leaves F0; a dyadic uniform
  velocity moving one cell per step shifts F0 by one node bit for bit (RK2
  and RK4, every scheme); boundary nodes are never written; a point that
  leaves the grid is an error and leaves its node unchanged; a Gaussian
  carried by a cellular flow converges at second order to the exact
  backtrack.
"""
import numpy as np
import pytest

from justpic_hip._abi import JP_ERR_ARG
from justpic_hip.grids import Axis, compute_dx, expand_range, staggered_grid
from oracle import oracle as O

from helpers import clear, grid2d, grid3d, host_particles, place, sample

SCHEMES = ("linear", "MQS", "LinP")


def euler_velocity(hp, V, scheme, dt=2.0 ** -12):
    """v at every active particle from one Euler step of advection_scheme"""
    m = hp.index == 1
    p0 = [c.copy() for c in hp.coords]
    q = hp.copy()
    O.advection_scheme(q, V, O.JP_EULER, 0.5, dt, scheme)
    return m, p0, [(q.coords[d] - p0[d]) / dt for d in range(hp.D)]


def bracket(ax, p):
    """0-based bracket k and t = (p - x[k]) / h on a uniform axis"""
    k = np.clip(np.searchsorted(ax.x, p, side="right") - 1, 0, ax.n - 2)
    return k, (p - ax.x[k]) / (ax.x[1] - ax.x[0])


def interior(axes, p):
    ok = np.ones(p[0].shape, dtype=bool)
    for d, ax in enumerate(axes):
        k, _ = bracket(ax, p[d])
        ok &= (k >= 1) & (k <= ax.n - 3)
    return ok


# ---------------------------------------------------------------- reference test
def refined_grid():
    """test/test_2D.jl:169-177: vector axes, refined along x"""
    xv = Axis(np.array([0.0, 0.1, 0.3, 0.6, 1.0]), "vector")
    yv = Axis(np.linspace(0.0, 1.0, 5), "vector")
    xc = Axis((xv.x[:-1] + xv.x[1:]) / 2, "vector")
    yc = Axis((yv.x[:-1] + yv.x[1:]) / 2, "vector")
    grid_vi = ((xv, expand_range(yc)), (expand_range(xc), yv))
    return (xv, yv), grid_vi


# Vx at (0.22, 0.48) by hand.  Vx's x axis is the refined xv: bracket
# [0.1, 0.3], t = 0.6.  MQS (t >= 1/2) adds 0.5 (t - 0.5)^2 (v0 - 2 v1 + v2)
# with the nodes 0.1, 0.3, 0.6, where 2x has second difference 0.2: +0.001.
# LinP (p past the cell middle 0.2): pressure values 2 * 0.2 and
# 2 * (0.3 + 0.6)/2 = 0.9 at xP = 0.2, 0.4, tP = 0.1: 0.45 for x's part, so
# VP = 0.93 and 2/3 * 0.92 + 1/3 * 0.93.  Vy's y axis is uniform: exact.
REFINED_VX = {"linear": 0.92, "MQS": 0.921, "LinP": 2 / 3 * 0.92 + 1 / 3 * 0.93}


@pytest.mark.parametrize("scheme", SCHEMES)
def test_reference_refined_grid_velocity(scheme):
    """test/test_2D.jl:179-196: Vx = 2x + y, Vy = x - 3y interpolated at
    (0.22, 0.48) (the reference checks that the seed does not matter; the
    values are computed by hand above)."""
    xv, grid_vi = refined_grid()
    V = (sample(lambda x, y: 2 * x + y, grid_vi[0]), sample(lambda x, y: x - 3 * y, grid_vi[1]))
    hp = O.HostParticles(xv, 4, grid_vi)
    clear(hp)
    place(hp, (1, 1), 0, (0.22, 0.48))
    _, _, v = euler_velocity(hp, V, scheme)
    e = np.flatnonzero(hp.index)[0]
    assert abs(v[0][e] - REFINED_VX[scheme]) < 1e-10
    assert abs(v[1][e] - (-1.22)) < 1e-10


@pytest.mark.parametrize("scheme,tup,want", [("MQS", False, 1.18), ("MQS", True, 1.18),
                                             ("LinP", True, 1.18), ("LinP", False, 1.18),
                                             ("linear", False, 1.18), ("linear", True, 1.14)])
def test_reference_refined_grid_backtrack(scheme, tup, want):
    """test/test_2D.jl:198-227 (MQS, LinP); the linear methods by hand."""
    xv, grid_vi = refined_grid()
    grid = tuple(compute_dx(a) for a in xv)
    gvi = tuple(tuple(compute_dx(a) for a in g) for g in grid_vi)
    F0 = sample(lambda x, y: x + 2 * y, xv)
    V = tuple(np.full(tuple(a.n for a in g), u) for g, u in zip(grid_vi, (0.08, 0.02)))
    F = F0.copy()
    O.semilagrangian(grid, gvi, V, [F], [F0], O.JP_RK2, 0.5, 1.0, scheme,
                     di_at_node=(tup and scheme == "linear"))
    assert abs(F[2, 2] - want) <= 1e-12
    assert F[2, 2] != F0[2, 2]


# ---------------------------------------------------------------- closed forms
def linear_field(c, D):
    a = [0.3, -1.7, 2.2][:D]
    return lambda *X: (1.0 + c) + sum(ai * (c + 1) * Xi for ai, Xi in zip(a, X))


@pytest.mark.parametrize("case,scheme", [("2d_range", s) for s in SCHEMES] +
                         [("2d_vector", "linear"), ("3d", "linear")])
def test_linear_fields_reproduced(case, scheme):
    """every scheme on uniform 2D grids; the linear scheme also on refined
    and 3D grids (MQS's second differences and LinP's pressure points do not
    vanish on refined grids, as the refined-grid values above show; the 3D
    stencils are the next test's)"""
    if case == "3d":
        xv, xi = grid3d(6)
        hp = host_particles(xv, xi, 24, 2, seed=4)
    else:
        xv, xi = grid2d(9, 7, kind=case[3:])
        hp = host_particles(xv, xi, 16, 3, seed=4)
    D = hp.D
    V = tuple(sample(linear_field(c, D), xi[c]) for c in range(D))
    m, p0, v = euler_velocity(hp, V, scheme)
    for c in range(D):
        want = linear_field(c, D)(*[p[m] for p in p0])
        assert np.max(np.abs(v[c][m] - want)) < 1e-9, c


def test_3d_linear_field_stencils():
    """3D, V_c = a_c + g_c . x on uniform grids.  The reference's 3D MQS
    takes extra nodes from other planes or rows than the corners it corrects
    (MQS.jl:58-71, 126-152), so a linear field gets a second difference of
    -g.h across them, worked out here per component:
      x: lin + t_z/2 (t_x - 1/2)^2 (-g_x[z] h_z)
      y: lin + t_z/2 (t_y - 1/2)^2 (-g_y[z] h_z)
      z: lin + t_y/2 (t_x - 1/2)^2 (-g_z[y] h_y)
    on interior stencils.  LinP's x and y components are exact; its z
    component lerps values that sit one y step apart with t_x, the two
    pressure points with t_y and one x step with t_z (augment_offset rows and
    swap_F), so VP = F(x0, y0, z_lo) + t_x g[y] h_y + t_y g[z] h_z +
    t_zP g[x] h_x and the result is 2/3 F(p) + 1/3 VP."""
    xv, xi = grid3d(6)
    hp = host_particles(xv, xi, 24, 2, seed=4)
    G = [(0.3, -1.7, 2.2), (-0.8, 0.6, 1.3), (1.1, 2.4, -0.9)]
    V = tuple(sample(lambda x, y, z, c=c: (1.0 + c) + G[c][0] * x + G[c][1] * y + G[c][2] * z,
                     xi[c]) for c in range(3))
    m = hp.index == 1
    p = [a[m] for a in hp.coords]
    lin = [(1.0 + c) + G[c][0] * p[0] + G[c][1] * p[1] + G[c][2] * p[2] for c in range(3)]

    def tk(c, d):
        ax = xi[c][d]
        k, t = bracket(ax, p[d])
        return k, t, ax.x[1] - ax.x[0]

    _, _, v = euler_velocity(hp, V, "MQS")
    for c, (a, b, zz) in enumerate([(0, 2, 2), (1, 2, 2), (0, 1, 1)]):
        # component c: correction direction a, weight t along b, gradient along zz
        _, ta, _ = tk(c, a)
        _, tb, _ = tk(c, b)
        _, _, hz = tk(c, zz)
        inn = interior(xi[c], p)
        want = np.where(inn, lin[c] + tb * 0.5 * (ta - 0.5) ** 2 * (-G[c][zz] * hz), lin[c])
        assert inn.any()
        assert np.max(np.abs(v[c][m] - want)) < 1e-9, c

    _, _, v = euler_velocity(hp, V, "LinP")
    for c in (0, 1):
        assert np.max(np.abs(v[c][m] - lin[c])) < 1e-9, c
    kx, tx, hx = tk(2, 0)
    ky, ty, hy = tk(2, 1)
    kz, tz, hz = tk(2, 2)
    x0, y0 = xi[2][0].x[kx], xi[2][1].x[ky]
    zc = xi[2][2].x[kz]
    up = p[2] > zc + hz / 2
    z_lo = np.where(up, zc + hz / 2, zc - hz / 2)
    tzp = (p[2] - z_lo) / hz
    g = G[2]
    vp = (3.0 + g[0] * x0 + g[1] * y0 + g[2] * z_lo) + tx * g[1] * hy + ty * g[2] * hz + \
        tzp * g[0] * hx
    inn = interior(xi[2], p)
    want = np.where(inn, 2 / 3 * lin[2] + 1 / 3 * vp, lin[2])
    assert np.max(np.abs(v[2][m] - want)) < 1e-9


def quad_case(D, n=8):
    if D == 2:
        xv, xi = grid2d(n)
        hp = host_particles(xv, xi, 16, 3, seed=6)
    else:
        xv, xi = grid3d(n)
        hp = host_particles(xv, xi, 24, 2, seed=6)
    return xv, xi, hp


def fields(xi, D, f):
    """component c sampled from f(c, X)"""
    return tuple(sample(lambda *X, c=c: f(c, X), xi[c]) for c in range(D))


@pytest.mark.parametrize("D", [2, 3])
def test_mqs_quadratic_along_own_axis(D):
    """V_c = x_c^2: x_c^2 + h^2/4 on interior stencils, the lerp elsewhere."""
    xv, xi, hp = quad_case(D)
    V = fields(xi, D, lambda c, X: X[c] ** 2)
    m, p0, v = euler_velocity(hp, V, "MQS")
    p = [a[m] for a in p0]
    for c in range(D if D == 2 else 2):  # the 3D z component: see the quirk test
        h = xi[c][c].x[1] - xi[c][c].x[0]
        _, t = bracket(xi[c][c], p[c])
        inn = interior(xi[c], p)
        want = np.where(inn, p[c] ** 2 + h * h / 4, p[c] ** 2 + h * h * t * (1 - t))
        assert inn.any() and (~inn).any()
        assert np.max(np.abs(v[c][m] - want)) < 1e-9, c


def test_mqs_corrects_only_along_its_direction():
    """2D: V_y = x^2 varies across y's direction: no correction, the lerp."""
    xv, xi, hp = quad_case(2)
    V = fields(xi, 2, lambda c, X: X[0] ** 2)
    m, p0, v = euler_velocity(hp, V, "MQS")
    p = [a[m] for a in p0]
    ax = xi[1][0]
    h = ax.x[1] - ax.x[0]
    _, t = bracket(ax, p[0])
    want = p[0] ** 2 + h * h * t * (1 - t)
    assert np.max(np.abs(v[1][m] - want)) < 1e-9


@pytest.mark.parametrize("D", [2, 3])
def test_linp_quadratic_along_own_axis(D):
    """V_c = x_c^2: 2/3 (x^2 + h^2 t(1-t)) + 1/3 (x^2 + h^2 s(1-s) + h^2/4)
    on interior stencils, s from the pressure point below p (x and y
    components; the 3D z component is the quirk test's)."""
    xv, xi, hp = quad_case(D)
    V = fields(xi, D, lambda c, X: X[c] ** 2)
    m, p0, v = euler_velocity(hp, V, "LinP")
    p = [a[m] for a in p0]
    for c in range(D if D == 2 else 2):
        ax = xi[c][c]
        h = ax.x[1] - ax.x[0]
        k, t = bracket(ax, p[c])
        s = np.where(t < 0.5, t + 0.5, t - 0.5)
        vl = p[c] ** 2 + h * h * t * (1 - t)
        vp = p[c] ** 2 + h * h * s * (1 - s) + h * h / 4
        inn = interior(xi[c], p)
        want = np.where(inn, 2 / 3 * vl + 1 / 3 * vp, vl)
        assert np.max(np.abs(v[c][m] - want)) < 1e-9, c


def test_3d_index_patterns():
    """The reference's 3D stencils as written (see the module docstring)."""
    xv, xi, hp = quad_case(3)
    m = hp.index == 1
    p = [a[m] for a in hp.coords]
    hz = xi[2][2].x[1] - xi[2][2].x[0]
    kz, tz = bracket(xi[2][2], p[2])
    inn_z = interior(xi[2], p)
    assert inn_z.any()
    # MQS z component of z^2: no correction along z
    _, _, v = euler_velocity(hp, fields(xi, 3, lambda c, X: X[2] ** 2), "MQS")
    assert np.max(np.abs(v[2][m] - (p[2] ** 2 + hz * hz * tz * (1 - tz)))) < 1e-9
    # MQS z component of x^2: corrected along x on interior stencils
    _, _, v = euler_velocity(hp, fields(xi, 3, lambda c, X: X[0] ** 2), "MQS")
    ax = xi[2][0]
    hx = ax.x[1] - ax.x[0]
    _, tx = bracket(ax, p[0])
    want = np.where(inn_z, p[0] ** 2 + hx * hx / 4, p[0] ** 2 + hx * hx * tx * (1 - tx))
    assert np.max(np.abs(v[2][m] - want)) < 1e-9
    # MQS x component of x^2 (1 + z): the top plane's corners (factor
    # B = 1 + z1) corrected with the extra node of plane k (factor A):
    # second difference 2 h^2 B + x_e^2 (A - B), x_e the node left or right
    _, _, v = euler_velocity(hp, fields(xi, 3, lambda c, X: X[0] ** 2 * (1 + X[2])), "MQS")
    ax, az = xi[0][0], xi[0][2]
    h = ax.x[1] - ax.x[0]
    k, t = bracket(ax, p[0])
    k2, t2 = bracket(az, p[2])
    A, B = 1 + az.x[k2], 1 + az.x[np.minimum(k2 + 1, az.n - 1)]
    xe = np.where(t < 0.5, ax.x[np.maximum(k - 1, 0)], ax.x[np.minimum(k + 2, ax.n - 1)])
    lerp_x = p[0] ** 2 + h * h * t * (1 - t)
    bot = lerp_x * A + (t - 0.5) ** 2 * h * h * A
    top = lerp_x * B + 0.5 * (t - 0.5) ** 2 * (2 * h * h * B + xe ** 2 * (A - B))
    inn = interior(xi[0], p)
    want = np.where(inn, bot + t2 * (top - bot), lerp_x * (1 + p[2]))
    assert np.max(np.abs(v[0][m] - want)) < 1e-9
    # LinP z component of z^2: the pressure pair lerped with t_y
    _, _, v = euler_velocity(hp, fields(xi, 3, lambda c, X: X[2] ** 2), "LinP")
    zc = xi[2][2].x
    b = kz + (p[2] > zc[kz] + hz / 2)
    A = (zc[b - 1] ** 2 + zc[b] ** 2) / 2
    B = (zc[np.minimum(b + 1, zc.size - 1)] ** 2 + zc[b] ** 2) / 2
    ay = xi[2][1]
    _, ty = bracket(ay, p[1])
    vl = p[2] ** 2 + hz * hz * tz * (1 - tz)
    want = np.where(inn_z, 2 / 3 * vl + 1 / 3 * (A + ty * (B - A)), vl)
    assert np.max(np.abs(v[2][m] - want)) < 1e-9


# ---------------------------------------------------------------- semi-Lagrangian
def sml_grid(D, n=8):
    xv = tuple(Axis.lin(0.0, 1.0, n + 1) for _ in range(D))
    xi = staggered_grid(xv)
    return (tuple(compute_dx(a) for a in xv),
            tuple(tuple(compute_dx(a) for a in g) for g in xi), xv, xi)


def smooth(axes, seed=0):
    rng = np.random.default_rng(seed)
    return sample(lambda *X: np.sin(3 * X[0] + seed) * np.cos(2 * X[-1]) + 0.1 * X[1], axes) + \
        0.01 * rng.random(tuple(a.n for a in axes))


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("D", [2, 3])
def test_sml_zero_velocity_is_identity(D, scheme):
    grid, gvi, xv, xi = sml_grid(D)
    V = tuple(np.zeros(tuple(a.n for a in g)) for g in xi)
    F0 = smooth(xv)
    F = np.full_like(F0, -7.0)
    O.semilagrangian(grid, gvi, V, [F], [F0], O.JP_RK2, 0.5, 0.01, scheme)
    inner = tuple(slice(1, -1) for _ in range(D))
    assert np.array_equal(F[inner], F0[inner])
    edge = np.ones(F.shape, dtype=bool)
    edge[inner] = False
    assert np.all(F[edge] == -7.0)


@pytest.mark.parametrize("method", [O.JP_RK2, O.JP_RK4])
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("D", [2, 3])
def test_sml_dyadic_shift_is_exact(D, scheme, method):
    """h = 1/8, V = 2 along x, dt = 1/16: every interior node takes F0 of
    its left neighbour, bit for bit."""
    grid, gvi, xv, xi = sml_grid(D)
    V = tuple(np.full(tuple(a.n for a in g), 2.0 if c == 0 else 0.0) for c, g in enumerate(xi))
    F0 = smooth(xv, 3)
    G0 = smooth(xv, 5)
    F, G = np.zeros_like(F0), np.zeros_like(G0)
    O.semilagrangian(grid, gvi, V, [F, G], [F0, G0], method, 0.5, 1 / 16, scheme)
    inner = tuple(slice(1, -1) for _ in range(D))
    left = (slice(0, -2),) + inner[1:]
    assert np.array_equal(F[inner], F0[left])
    assert np.array_equal(G[inner], G0[left])


def test_sml_out_of_grid_is_an_error():
    grid, gvi, xv, xi = sml_grid(2)
    V = tuple(np.full(tuple(a.n for a in g), 3.0) for g in xi)
    F0 = smooth(xv)
    F = np.full_like(F0, -1.0)
    # one step of 3 * 0.2 moves every node further than 0.375 out of some row
    rc = O.semilagrangian(grid, gvi, V, [F], [F0], O.JP_RK2, 0.5, 0.2, "linear", check=False)
    assert rc == JP_ERR_ARG
    x = xv[0].x
    # nodes whose backtrack stays inside were written, the others not
    inside = np.zeros(F.shape, dtype=bool)
    inside[1:-1, 1:-1] = ((x[1:-1, None] - 0.6 >= 0) & (x[None, 1:-1] - 0.6 >= 0))
    assert np.all(F[~inside] == -1.0)
    assert np.all(F[inside] != -1.0)


def test_sml_converges_second_order():
    """A Gaussian carried by the cellular flow vx = sin(pi x) cos(pi y),
    vy = -cos(pi x) sin(pi y) (no normal velocity on the walls) for one
    step: against the exact backtrack (numerical RK4 on the analytic
    velocity), the error falls by about 4 when h halves (interpolation of F0
    and of V are both O(h^2)), for every scheme and both integrators."""
    g = lambda x, y: np.exp(-((x - 0.6) ** 2 + (y - 0.45) ** 2) / 0.02)

    def vel(x, y):
        return np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)

    dt = 0.05
    errs = {}
    for n in (32, 64):
        grid, gvi, xv, xi = sml_grid(2, n)
        V = (sample(lambda x, y: vel(x, y)[0], xi[0]), sample(lambda x, y: vel(x, y)[1], xi[1]))
        F0 = sample(g, xv)
        X, Y = np.meshgrid(xv[0].x, xv[1].x, indexing="ij")
        h = -dt / 200
        for _ in range(200):  # exact backtrack
            k1 = vel(X, Y)
            k2 = vel(X + h / 2 * k1[0], Y + h / 2 * k1[1])
            k3 = vel(X + h / 2 * k2[0], Y + h / 2 * k2[1])
            k4 = vel(X + h * k3[0], Y + h * k3[1])
            X = X + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            Y = Y + h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        exact = g(X, Y)
        for scheme in SCHEMES:
            for method in (O.JP_RK2, O.JP_RK4):
                F = F0.copy()
                O.semilagrangian(grid, gvi, V, [F], [F0], method, 0.5, dt, scheme)
                errs[n, scheme, method] = np.max(np.abs(F - exact)[1:-1, 1:-1])
    for scheme in SCHEMES:
        for method in (O.JP_RK2, O.JP_RK4):
            e32, e64 = errs[32, scheme, method], errs[64, scheme, method]
            assert e64 < 0.02 and e32 / e64 > 3.0, (scheme, method, e32, e64)
