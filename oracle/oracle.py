"""ctypes wrapper of the C oracle (oracle/jp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the reported CPU
baseline, never as the product path.  Arrays are host numpy arrays in the
library's layout (element (cell c, slot k) at [c + C*k]).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from justpic_hip._abi import (JP_EULER, JP_RK2, JP_RK4, jp_axis, jp_cellgrid,
                              jp_stats, jp_velgrid, ptr_array)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjporacle.so")
LIB_OMP_PATH = os.path.join(HERE, "libjporacle_omp.so")

_lib = None
_path = LIB_PATH


def set_threaded(on: bool):
    """Run the oracle on all host threads (the OpenMP build: the cells of one
    colour phase in parallel, as the reference's threaded backend runs them)
    or sequentially (the default, and the only form the tests check with)."""
    global _lib, _path
    want = LIB_OMP_PATH if on else LIB_PATH
    if want != _path:
        _path, _lib = want, None


def threads() -> int:
    return int(lib().jpo_threads())


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_path):
            build()
        L = C.CDLL(_path)
        P = C.c_void_p
        sig = {
            "jpo_threads": (C.c_int, []),
            "jpo_lerp1": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "jpo_lerp2": (C.c_double, [P, P]),
            "jpo_lerp3": (C.c_double, [P, P]),
            "jpo_bisect": (C.c_int64, [C.c_double, P, C.c_int64, C.c_int64]),
            "jpo_first_stage": (None, [C.c_int, C.c_double, C.c_double, C.c_int, P, P, P]),
            "jpo_second_stage": (None, [C.c_double, C.c_double, C.c_int, P, P, P, P]),
            "jpo_distance": (C.c_double, [C.c_int, P, P]),
            "jpo_philox": (None, [P, P, P]),
            "jpo_uniform": (C.c_double, [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                         C.c_uint64]),
            "jpo_cell_index": (C.c_int64, [C.c_double, P, C.c_int64]),
            "jpo_init_particles": (C.c_int, [C.POINTER(jp_cellgrid), P, P, C.c_int, C.c_uint64]),
            "jpo_advection": (C.c_int, [C.POINTER(jp_cellgrid), P, P, P, C.POINTER(jp_velgrid),
                                        C.c_int, C.c_double, C.c_double]),
            "jpo_move_particles": (C.c_int, [C.POINTER(jp_cellgrid), P, P, C.c_int, P,
                                             C.POINTER(jp_stats)]),
            "jpo_particle2grid": (C.c_int, [C.POINTER(jp_cellgrid), P, P, P, P]),
            "jpo_particle2grid_multi": (C.c_int, [C.POINTER(jp_cellgrid), C.c_int, P, P, P, P]),
            "jpo_grid2particle": (C.c_int, [C.POINTER(jp_cellgrid), P, P, P, P]),
            "jpo_inject_particles": (C.c_int, [C.POINTER(jp_cellgrid), P, P, C.c_int, P, C.c_int,
                                               C.c_uint64, C.c_uint64, C.POINTER(jp_stats)]),
            "jpo_inject_particles_phase": (C.c_int, [C.POINTER(jp_cellgrid), P, P, P, P, C.c_int,
                                                     P, P, P, C.c_int, C.c_uint64, C.c_uint64,
                                                     C.POINTER(jp_stats)]),
            "jpo_update_phase_ratios": (C.c_int, [C.POINTER(jp_cellgrid), P, P, P, C.c_int, P, P,
                                                  P, P]),
            "jpo_grid2particle_flip": (C.c_int, [C.POINTER(jp_cellgrid), C.c_int, P, P, P, P, P,
                                                 C.c_double]),
            "jpo_particle2centroid": (C.c_int, [C.POINTER(jp_cellgrid), P, C.c_int, P, P, P,
                                                C.c_int]),
            "jpo_centroid2particle": (C.c_int, [C.POINTER(jp_cellgrid), P, C.c_int, P, P, P]),
            "jpo_advection_scheme": (C.c_int, [C.POINTER(jp_cellgrid), P, P, P,
                                               C.POINTER(jp_velgrid), C.c_int, C.c_double,
                                               C.c_double, C.c_int]),
            "jpo_semilagrangian": (C.c_int, [C.c_int, P, C.POINTER(jp_velgrid), P, C.c_int, P, P,
                                             C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]),
            "jpo_adjoint_forward": (C.c_int, [C.POINTER(jp_cellgrid), C.POINTER(jp_velgrid), P,
                                              C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                              C.c_int64, P, P, P, P, P, P, P]),
            "jpo_adjoint_backward": (C.c_int, [C.POINTER(jp_cellgrid), C.POINTER(jp_velgrid), P,
                                               C.c_int, C.c_double, C.c_double, C.c_int,
                                               C.c_int64, P, P, P, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p).value


def _chk(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} returned {rc}")


def host_axis(ax):
    return jp_axis(_p(ax.x), _p(ax.dx), _p(ax.inv_dx), ax.n)


# ---- primitives ---------------------------------------------------------

def lerp(v, t):
    v = np.ascontiguousarray(v, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    if t.size == 1:
        return lib().jpo_lerp1(float(t[0]), float(v[0]), float(v[1]))
    if t.size == 2:
        return lib().jpo_lerp2(_p(v), _p(t))
    return lib().jpo_lerp3(_p(v), _p(t))


def bisect(px, x, seed1):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().jpo_bisect(float(px), _p(x), x.size, int(seed1))


def first_stage(method, alpha, dt, v, p):
    v = np.ascontiguousarray(v, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty_like(p)
    lib().jpo_first_stage(method, float(alpha), float(dt), p.size, _p(v), _p(p), _p(out))
    return tuple(out)


def second_stage(alpha, dt, v0, v1, p):
    v0, v1, p = (np.ascontiguousarray(a, dtype=np.float64) for a in (v0, v1, p))
    out = np.empty_like(p)
    lib().jpo_second_stage(float(alpha), float(dt), p.size, _p(v0), _p(v1), _p(p), _p(out))
    return tuple(out)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().jpo_philox(_p(c), _p(k), _p(out))
    return out


def uniform(seed, cell, slot, dim, purpose, counter=0):
    return lib().jpo_uniform(seed, cell, slot, dim, purpose, counter)


def cell_index(x, xv):
    xv = np.ascontiguousarray(xv, dtype=np.float64)
    return lib().jpo_cell_index(float(x), _p(xv), xv.size)


# ---- particle state -----------------------------------------------------

class HostParticles:
    """Host copy of a particle container in the library layout."""

    def __init__(self, xv_axes, K, xi_vel=None):
        self.D = len(xv_axes)
        self.xv = tuple(xv_axes)
        self.ni = tuple(a.n - 1 for a in xv_axes)
        self.K = int(K)
        self.C = int(np.prod(self.ni))
        self.S = self.C * self.K
        self.xi_vel = xi_vel
        self.coords = [np.full(self.S, np.nan) for _ in range(self.D)]
        self.index = np.zeros(self.S, dtype=np.uint8)
        self._grid = self.cellgrid()

    def cellgrid(self):
        g = jp_cellgrid()
        g.ndim = self.D
        g.max_xcell = self.K
        for d in range(self.D):
            g.ni[d] = self.ni[d]
            g.xv[d] = host_axis(self.xv[d])
        return g

    def velgrid(self):
        vg = jp_velgrid()
        for c in range(self.D):
            for d in range(self.D):
                vg.ax[c][d] = host_axis(self.xi_vel[c][d])
        return vg

    def copy(self):
        o = HostParticles.__new__(HostParticles)
        o.__dict__.update(self.__dict__)
        o.coords = [a.copy() for a in self.coords]
        o.index = self.index.copy()
        o._grid = o.cellgrid()
        return o

    def coord_ptrs(self):
        return ptr_array([_p(a) for a in self.coords])


def init_particles(hp: HostParticles, np_quadrant, seed):
    _chk(lib().jpo_init_particles(C.byref(hp._grid), hp.coord_ptrs(), _p(hp.index),
                                  int(np_quadrant), int(seed)), "init")


def advection(hp: HostParticles, V, method, alpha, dt):
    V = [np.ascontiguousarray(v.ravel(order="F"), dtype=np.float64) for v in V]
    vg = hp.velgrid()
    _chk(lib().jpo_advection(C.byref(hp._grid), hp.coord_ptrs(), _p(hp.index),
                             ptr_array([_p(v) for v in V]), C.byref(vg), int(method),
                             float(alpha), float(dt)), "advection")


SCHEMES = {"linear": 0, "MQS": 1, "LinP": 2}


def advection_scheme(hp: HostParticles, V, method, alpha, dt, scheme):
    """advection! (scheme "linear"), advection_MQS! or advection_LinP!"""
    V = [np.ascontiguousarray(v.ravel(order="F"), dtype=np.float64) for v in V]
    vg = hp.velgrid()
    _chk(lib().jpo_advection_scheme(C.byref(hp._grid), hp.coord_ptrs(), _p(hp.index),
                                    ptr_array([_p(v) for v in V]), C.byref(vg), int(method),
                                    float(alpha), float(dt), SCHEMES[scheme]), "advection_scheme")


def semilagrangian(grid, xi_vel, V, F, F0, method, alpha, dt, scheme="linear",
                   di_at_node=False, check=True):
    """semilagrangian_advection! and its MQS / LinP forms on nodal fields.
    grid: D vertex axes, xi_vel: velocity axes, both already carrying
    compute_dx spacings (justpic_hip.grids.compute_dx); F, F0: lists of
    arrays of the nodal shape (F updated in place).  Returns the status
    (JP_ERR_ARG when a backtracked point left the grid) unless check."""
    D = len(grid)
    g = (jp_axis * 3)()
    for d in range(D):
        g[d] = host_axis(grid[d])
    vg = jp_velgrid()
    for c in range(D):
        for d in range(D):
            vg.ax[c][d] = host_axis(xi_vel[c][d])
    V = [np.ascontiguousarray(v.ravel(order="F"), dtype=np.float64) for v in V]
    Ff = [np.asfortranarray(f, dtype=np.float64) for f in F]
    F0f = [np.ascontiguousarray(np.asarray(f, dtype=np.float64).ravel(order="F")) for f in F0]
    rc = lib().jpo_semilagrangian(D, C.cast(g, C.c_void_p), C.byref(vg),
                                  ptr_array([_p(v) for v in V]), len(F),
                                  ptr_array([_p(f) for f in Ff]),
                                  ptr_array([_p(f) for f in F0f]), int(method), float(alpha),
                                  float(dt), SCHEMES[scheme], int(bool(di_at_node)))
    for f, ff in zip(F, Ff):
        f[...] = ff
    if check:
        _chk(rc, "semilagrangian")
    return rc


def move_particles(hp: HostParticles, args=()):
    st = jp_stats()
    _chk(lib().jpo_move_particles(C.byref(hp._grid), hp.coord_ptrs(), _p(hp.index), len(args),
                                  ptr_array([_p(a) for a in args]), C.byref(st)), "move")
    return st.as_dict()


def particle2grid(hp: HostParticles, Fp):
    nn = tuple(n + 1 for n in hp.ni)
    F = np.zeros(int(np.prod(nn)))
    _chk(lib().jpo_particle2grid(C.byref(hp._grid), _p(F), _p(Fp), hp.coord_ptrs(),
                                 _p(hp.index)), "p2g")
    return F


def particle2grid_multi(hp: HostParticles, Fps):
    nn = tuple(n + 1 for n in hp.ni)
    Fs = [np.zeros(int(np.prod(nn))) for _ in Fps]
    _chk(lib().jpo_particle2grid_multi(C.byref(hp._grid), len(Fps),
                                       ptr_array([_p(f) for f in Fs]),
                                       ptr_array([_p(f) for f in Fps]), hp.coord_ptrs(),
                                       _p(hp.index)), "p2g multi")
    return Fs


def grid2particle(hp: HostParticles, F, Fp=None):
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).ravel(order="F"))
    if Fp is None:
        Fp = np.zeros(hp.S)
    _chk(lib().jpo_grid2particle(C.byref(hp._grid), _p(Fp), _p(F), hp.coord_ptrs(),
                                 _p(hp.index)), "g2p")
    return Fp


def inject_particles(hp: HostParticles, args, min_xcell, seed, counter):
    st = jp_stats()
    _chk(lib().jpo_inject_particles(C.byref(hp._grid), hp.coord_ptrs(), _p(hp.index), len(args),
                                    ptr_array([_p(a) for a in args]), int(min_xcell), int(seed),
                                    int(counter), C.byref(st)), "inject")
    return st.as_dict()


def _flat(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def centre_axes(hp: HostParticles):
    """particles.xci: the interior of the staggered velocity axes
    (src/Particles/particles_utils.jl:56-70)."""
    xi = hp.xi_vel
    if hp.D == 2:
        return (xi[1][0].inner(), xi[0][1].inner())
    return (xi[1][0].inner(), xi[0][1].inner(), xi[0][2].inner())


def inject_particles_phase(hp: HostParticles, phases, args, fields, min_xcell, seed, counter):
    """inject_particles_phase!(particles, phases, args, fields): `fields[a]`
    is a nodal (ni+1) or cell-centred (ni) array, told apart by its shape."""
    xc = centre_axes(hp)
    axes = (jp_axis * hp.D)(*[host_axis(a) for a in xc])
    nodal = tuple(n + 1 for n in hp.ni)
    centred = []
    for f in fields:
        shp = np.shape(f)
        if shp == tuple(hp.ni):
            centred.append(1)
        elif shp == nodal:
            centred.append(0)
        else:
            raise ValueError(f"field of shape {shp} is neither nodal {nodal} nor centred {hp.ni}")
    flds = [_flat(f) for f in fields]
    cen = (C.c_int * max(1, len(centred)))(*centred)
    st = jp_stats()
    _chk(lib().jpo_inject_particles_phase(C.byref(hp._grid), axes, hp.coord_ptrs(),
                                          _p(hp.index), _p(phases), len(args),
                                          ptr_array([_p(a) for a in args]),
                                          ptr_array([_p(f) for f in flds]), cen, int(min_xcell),
                                          int(seed), int(counter), C.byref(st)),
         "inject phase")
    return st.as_dict()


def grid2particle_flip(hp: HostParticles, Fp, F, F0, alpha):
    """grid2particle_flip!: Fp (list of slot arrays) updated in place."""
    F = [_flat(f) for f in F]
    F0 = [_flat(f) for f in F0]
    _chk(lib().jpo_grid2particle_flip(C.byref(hp._grid), len(Fp), ptr_array([_p(a) for a in Fp]),
                                      ptr_array([_p(f) for f in F]),
                                      ptr_array([_p(f) for f in F0]), hp.coord_ptrs(),
                                      _p(hp.index), float(alpha)), "flip")
    return Fp


def particle2centroid(hp: HostParticles, Fp, tuple_method=False):
    """particle2centroid!: flat cell-centred arrays, one per field."""
    xc = centre_axes(hp)
    axes = (jp_axis * hp.D)(*[host_axis(a) for a in xc])
    F = [np.zeros(hp.C) for _ in Fp]
    _chk(lib().jpo_particle2centroid(C.byref(hp._grid), axes, len(Fp),
                                     ptr_array([_p(f) for f in F]),
                                     ptr_array([_p(_flat(a)) for a in Fp]), hp.coord_ptrs(),
                                     1 if tuple_method else 0), "p2c")
    return F


def centroid2particle(hp: HostParticles, Fp, F):
    """centroid2particle!: Fp (list of slot arrays) updated in place."""
    xc = centre_axes(hp)
    axes = (jp_axis * hp.D)(*[host_axis(a) for a in xc])
    F = [_flat(f) for f in F]
    _chk(lib().jpo_centroid2particle(C.byref(hp._grid), axes, len(Fp),
                                     ptr_array([_p(a) for a in Fp]),
                                     ptr_array([_p(f) for f in F]), hp.coord_ptrs()), "c2p")
    return Fp


def phase_ratio_shapes(ni):
    """Locations of every PhaseRatios field (src/PhaseRatios/constructors.jl:16-43)."""
    D = len(ni)
    out = {"center": tuple(ni), "vertex": tuple(n + 1 for n in ni)}
    for d, name in enumerate(("Vx", "Vy", "Vz")[:D]):
        out[name] = tuple(n + (1 if e == d else 0) for e, n in enumerate(ni))
    if D == 3:
        for name, o in (("xy", (1, 1, 0)), ("yz", (0, 1, 1)), ("xz", (1, 0, 1))):
            out[name] = tuple(n + k for n, k in zip(ni, o))
    return out


def update_phase_ratios(hp: HostParticles, phases, nphases):
    """update_phase_ratios! (src/PhaseRatios/utils.jl:15-43): a dict of flat
    arrays in the device CellArray layout (location l, phase k at
    [l + L k])."""
    xc = centre_axes(hp)
    axes = (jp_axis * hp.D)(*[host_axis(a) for a in xc])
    shp = phase_ratio_shapes(hp.ni)
    out = {k: np.full(int(np.prod(v)) * nphases, -1.0) for k, v in shp.items()}
    faces = [out[k] for k in ("Vx", "Vy", "Vz")[:hp.D]]
    mids = [out[k] for k in ("xy", "yz", "xz")] if hp.D == 3 else []
    _chk(lib().jpo_update_phase_ratios(C.byref(hp._grid), axes, hp.coord_ptrs(), _p(phases),
                                       int(nphases), _p(out["center"]), _p(out["vertex"]),
                                       ptr_array([_p(f) for f in faces]),
                                       ptr_array([_p(m) for m in mids]) if mids else None),
         "phase ratios")
    return out


def adjoint_forward(hp: HostParticles, V, method, alpha, dt, nsteps, x0, Tp, every=10):
    """Flat-list model F = particle2grid(advect^nsteps(x0), Tp) (host arrays).
    Returns dict(xN, alive, F, W, ckpt); F and W are flat nodal arrays."""
    D = hp.D
    V = [_flat(v) for v in V]
    x0 = [_flat(x) for x in x0]
    Tp = _flat(Tp)
    n = Tp.size
    nn = int(np.prod([k + 1 for k in hp.ni]))
    nseg = max(1, -(-int(nsteps) // int(every)))
    ckpt = np.zeros(max(1, nseg * D * n))
    alive = np.zeros(max(1, n), dtype=np.uint8)
    xN = [np.zeros(max(1, n)) for _ in range(D)]
    F = np.zeros(nn)
    W = np.zeros(nn)
    vg = hp.velgrid()
    _chk(lib().jpo_adjoint_forward(C.byref(hp._grid), C.byref(vg), ptr_array([_p(v) for v in V]),
                                   int(method), float(alpha), float(dt), int(nsteps), int(every),
                                   n, ptr_array([_p(x) for x in x0]), _p(Tp), _p(ckpt),
                                   _p(alive), ptr_array([_p(x) for x in xN]), _p(F), _p(W)),
         "adjoint forward")
    return dict(xN=[x[:n] for x in xN], alive=alive[:n], F=F, W=W,
                ckpt=ckpt[:nseg * D * n].reshape(nseg, D, n))


def adjoint_backward(hp: HostParticles, V, method, alpha, dt, nsteps, x0, Tp, F, W, Fbar):
    """Cotangents (Vbar tuple shaped like V, x0bar tuple, Tpbar) of the
    flat-list model for the nodal cotangent Fbar."""
    D = hp.D
    shapes = [np.shape(v) for v in V]
    V = [_flat(v) for v in V]
    x0 = [_flat(x) for x in x0]
    Tp, F, W, Fbar = (_flat(a) for a in (Tp, F, W, Fbar))
    n = Tp.size
    Vbar = [np.zeros_like(v) for v in V]
    x0bar = [np.zeros(max(1, n)) for _ in range(D)]
    Tpbar = np.zeros(max(1, n))
    vg = hp.velgrid()
    _chk(lib().jpo_adjoint_backward(C.byref(hp._grid), C.byref(vg), ptr_array([_p(v) for v in V]),
                                    int(method), float(alpha), float(dt), int(nsteps), n,
                                    ptr_array([_p(x) for x in x0]), _p(Tp), _p(F), _p(W),
                                    _p(Fbar), ptr_array([_p(v) for v in Vbar]),
                                    ptr_array([_p(x) for x in x0bar]), _p(Tpbar)),
         "adjoint backward")
    Vbar = tuple(v.reshape(s, order="F") for v, s in zip(Vbar, shapes))
    return Vbar, [x[:n] for x in x0bar], Tpbar[:n]


__all__ = ["adjoint_forward", "adjoint_backward", "JP_EULER", "JP_RK2", "JP_RK4", "HostParticles", "init_particles", "advection",
           "move_particles", "particle2grid", "particle2grid_multi", "grid2particle",
           "inject_particles", "lerp", "bisect", "first_stage", "second_stage", "philox",
           "uniform", "cell_index", "build", "lib"]
