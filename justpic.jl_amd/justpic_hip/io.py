"""Host copies in the CPU CellArray layout, and checkpoints.

The reference copies device particles to the host with ``Array(particles)``
(src/CellArrays/conversion.jl:32-69): every CellArray's data goes from the
GPU layout (C, K, 1) to the CPU layout (1, K, C), a permutation by (3, 2, 1).
``checkpointing_particles`` (src/IO/JLD2.jl:8-129) saves those host arrays,
with the phases, the particle fields, the time and the time step, and the GPU
constructors (ext/JustPICAMDGPUExt.jl:53-75) bring them back.

Here the permutation runs on the GPU (``jp_cellarray_to_cpu_layout``) and only
the permuted buffer crosses PCIe.  JLD2 (HDF5) is not available in this
environment, so a checkpoint is an ``.npz`` file with one entry per JLD2
key, each array in the CPU layout with its Julia shape (1, K, C) stored
column-major; it loads without pickle.
"""
from __future__ import annotations

import os

import numpy as np

from . import _abi
from .device import Context, DeviceArray
from .grids import Axis
from .particles import CellArray, Particles


def cellarray_to_host(ca: CellArray) -> np.ndarray:
    """``Array(ca)``: the CPU CellArray data, a (1, K, C) array (Fortran order,
    element (1, k, c) is slot k of cell c, both 1-based in Julia)."""
    ctx = ca.ctx
    buf = DeviceArray(ctx, (ca.C * ca.K,), ca.dtype)
    ctx.check(ctx.lib.jp_cellarray_to_cpu_layout(ctx.h, ca.ptr, buf.ptr, ca.C, ca.K,
                                                 ca.dtype.itemsize), "to_cpu_layout")
    flat = buf.to_host()
    buf.free()
    return flat.reshape((1, ca.K, ca.C), order="F")


def cellarray_from_host(ctx: Context, data, ni) -> CellArray:
    """Device CellArray from CPU-layout data of Julia shape (1, K, C)."""
    data = np.asarray(data)
    if data.ndim != 3 or data.shape[0] != 1:
        raise ValueError(f"CPU CellArray data must have shape (1, K, C), got {data.shape}")
    K, C = data.shape[1], data.shape[2]
    if C != int(np.prod(ni)):
        raise ValueError(f"{C} cells in the data, {int(np.prod(ni))} in ni={tuple(ni)}")
    dtype = np.uint8 if data.dtype == np.bool_ else data.dtype
    flat = np.ascontiguousarray(data.astype(dtype).ravel(order="F"))
    buf = DeviceArray.from_host(ctx, flat)
    ca = CellArray(ctx, ni, K, dtype=dtype)
    ctx.check(ctx.lib.jp_cellarray_from_cpu_layout(ctx.h, buf.ptr, ca.ptr, C, K,
                                                   np.dtype(dtype).itemsize), "from_cpu_layout")
    buf.free()
    return ca


def particles_to_host(p: Particles) -> dict:
    """``Array(particles)`` for an FP64 container: coordinates and the index
    mask in the CPU layout, with the scalar fields and the grids."""
    if p.form != "f64":
        raise ValueError("convert a split container with to_f64 first")
    return {
        "coords": tuple(cellarray_to_host(c) for c in p.coords),
        "index": cellarray_to_host(p.index).astype(np.bool_),
        "nxcell": p.nxcell, "max_xcell": p.max_xcell, "min_xcell": p.min_xcell, "np": p.np,
        "xvi": tuple(a.x for a in p.xvi), "xci": tuple(a.x for a in p.xci),
        "xi_vel": tuple(tuple(a.x for a in g) for g in p.xi_vel),
        "kinds": tuple(tuple(a.kind for a in g) for g in p.xi_vel),
        "seed": p.seed, "inject_counter": p.inject_counter,
    }


def particles_from_host_layout(ctx: Context, h: dict) -> Particles:
    """The device container of a ``particles_to_host`` copy (the GPU
    constructor of ext/JustPICAMDGPUExt.jl:53-75)."""
    xi_vel = tuple(tuple(Axis(x, k) for x, k in zip(g, kg))
                   for g, kg in zip(h["xi_vel"], h["kinds"]))
    D = len(xi_vel)
    xvi = tuple(xi_vel[d][d] for d in range(D))
    xci = tuple(Axis(x, xvi[d].kind) for d, x in enumerate(h["xci"]))
    ni = tuple(a.n - 1 for a in xvi)
    coords = [cellarray_from_host(ctx, c, ni) for c in h["coords"]]
    index = cellarray_from_host(ctx, h["index"], ni)
    p = Particles(ctx, coords, index, h["nxcell"], h["max_xcell"], h["min_xcell"], xvi, xci,
                  xi_vel, seed=int(h.get("seed", 0)))
    p.inject_counter = int(h.get("inject_counter", 0))
    return p


def checkpoint_name(dst, me=None):
    """src/IO/JLD2.jl:3-6 (``.npz`` instead of ``.jld2``)."""
    if me is None:
        return os.path.join(dst, "particles_checkpoint.npz")
    return os.path.join(dst, f"particles_checkpoint{int(me):04d}.npz")


def checkpointing_particles(dst, particles: Particles, me=None, *, phases=None, t=None, dt=None,
                            particle_args=None, **kwargs):
    """``checkpointing_particles(dst, particles; phases, t, dt, particle_args,
    kwargs...)`` (src/IO/JLD2.jl:8-129): every CellArray in the CPU layout;
    written to a temporary name and moved into place.  Returns the path."""
    os.makedirs(dst, exist_ok=True)
    fname = checkpoint_name(dst, me)
    h = particles_to_host(particles)
    out = {}
    for d, c in enumerate(h["coords"]):
        out[f"particles.coords.{d + 1}"] = c
    out["particles.index"] = h["index"]
    for k in ("nxcell", "max_xcell", "min_xcell", "np", "seed", "inject_counter"):
        out[f"particles.{k}"] = np.asarray(h[k])
    for d in range(len(h["xvi"])):
        out[f"particles.xci.{d + 1}"] = np.asarray(h["xci"][d])
        for e in range(len(h["xvi"])):
            out[f"particles.xi_vel.{d + 1}.{e + 1}"] = np.asarray(h["xi_vel"][d][e])
            out[f"particles.kind.{d + 1}.{e + 1}"] = np.asarray(h["kinds"][d][e])
    if phases is not None:
        out["phases"] = cellarray_to_host(phases)
    if particle_args is not None:
        for i, a in enumerate(particle_args):
            out[f"particle_args.{i + 1}"] = cellarray_to_host(a)
    if t is not None:
        out["time"] = np.asarray(t)
    if dt is not None:
        out["timestep"] = np.asarray(dt)
    for k, v in kwargs.items():
        if v is None:
            continue
        out[k] = cellarray_to_host(v) if isinstance(v, CellArray) else np.asarray(v)
    tmp = fname + ".tmp.npz"
    np.savez(tmp, **out)
    os.replace(tmp, fname)
    return fname


def load_checkpoint(fname, ctx: Context) -> dict:
    """Device state of a checkpoint: ``particles``, and ``phases``,
    ``particle_args``, ``time``, ``timestep`` when they were saved."""
    z = np.load(fname, allow_pickle=False)
    D = sum(1 for k in z.files if k.startswith("particles.coords."))
    h = {
        "coords": tuple(z[f"particles.coords.{d + 1}"] for d in range(D)),
        "index": z["particles.index"],
        "xci": tuple(z[f"particles.xci.{d + 1}"] for d in range(D)),
        "xi_vel": tuple(tuple(z[f"particles.xi_vel.{d + 1}.{e + 1}"] for e in range(D))
                        for d in range(D)),
        "kinds": tuple(tuple(str(z[f"particles.kind.{d + 1}.{e + 1}"]) for e in range(D))
                       for d in range(D)),
    }
    for k in ("nxcell", "max_xcell", "min_xcell", "np", "seed", "inject_counter"):
        h[k] = int(z[f"particles.{k}"])
    p = particles_from_host_layout(ctx, h)
    out = {"particles": p}
    if "phases" in z.files:
        out["phases"] = cellarray_from_host(ctx, z["phases"], p.ni)
    n_args = sum(1 for k in z.files if k.startswith("particle_args."))
    if n_args:
        out["particle_args"] = tuple(cellarray_from_host(ctx, z[f"particle_args.{i + 1}"], p.ni)
                                     for i in range(n_args))
    for k in ("time", "timestep"):
        if k in z.files:
            out[k] = float(z[k])
    return out
