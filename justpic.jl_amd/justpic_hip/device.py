"""Device context and device buffers over the C-ABI (no torch, no AMDGPU.jl).

The library owns the HIP runtime calls; this module only keeps pointers,
shapes and dtypes.  Shapes follow Julia (column-major) conventions: a nodal
field of size (nx+1, ny+1) is stored x-fastest, and ``to_host()`` returns a
Fortran-ordered numpy array with the Julia shape.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class JustPICError(RuntimeError):
    pass


class Context:
    """One HIP device + stream (``jp_ctx``).  ``synchronous=True`` keeps the
    reference's synchronous-launch semantics (src/launch.jl:65-69)."""

    def __init__(self, device: int = 0, synchronous: bool = True):
        self.lib = _abi.load_library()
        h = C.c_void_p()
        rc = self.lib.jp_ctx_create(C.byref(h), int(device))
        if rc != 0 or not h.value:
            raise JustPICError(f"jp_ctx_create(device={device}) failed with {rc} "
                               "(no HIP device visible?)")
        self.h = h
        self.device = device
        self.set_sync(synchronous)

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.jp_last_error(self.h)
            msg = msg.decode() if msg else ""
            if rc == _abi.JP_ERR_ARG:  # invalid argument: Julia's ArgumentError
                raise ValueError(f"{what}: {msg} (status {rc})")
            raise JustPICError(f"{what}: {msg} (status {rc})")

    def set_sync(self, synchronous: bool):
        self.check(self.lib.jp_ctx_set_sync(self.h, 1 if synchronous else 0), "set_sync")
        self.synchronous = bool(synchronous)

    def sync(self):
        self.check(self.lib.jp_sync(self.h), "sync")

    def set_option(self, option: int, value: int):
        """Launch option (``_abi.JP_OPT_*``; 0 restores the size-based choice).
        Results do not depend on it: it lets small problems run the launch
        shapes that large ones select."""
        self.check(self.lib.jp_ctx_set_option(self.h, int(option), int(value)), "set_option")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.jp_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # events / graphs
    def event(self):
        e = C.c_void_p()
        self.check(self.lib.jp_event_create(self.h, C.byref(e)), "event_create")
        return e

    def record(self, ev):
        self.check(self.lib.jp_event_record(self.h, ev), "event_record")

    def wait(self, ev):
        """This context's stream waits for `ev` (recorded on any context)."""
        self.check(self.lib.jp_event_wait(self.h, ev), "event_wait")

    def elapsed_ms(self, a, b) -> float:
        ms = C.c_float()
        self.check(self.lib.jp_event_elapsed_ms(self.h, a, b, C.byref(ms)), "elapsed")
        return float(ms.value)

    def graph_begin(self):
        self.check(self.lib.jp_graph_begin(self.h), "graph_begin")

    def graph_end(self):
        g = C.c_void_p()
        self.check(self.lib.jp_graph_end(self.h, C.byref(g)), "graph_end")
        return g

    def graph_launch(self, g):
        self.check(self.lib.jp_graph_launch(self.h, g), "graph_launch")

    def graph_destroy(self, g):
        self.check(self.lib.jp_graph_destroy(self.h, g), "graph_destroy")


class DeviceArray:
    """A device buffer with a Julia (column-major) shape and a numpy dtype."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if np.iterable(shape) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.lib.jp_malloc(ctx.h, C.byref(p), max(1, self.nbytes)), "malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx: Context, a, dtype=None):
        a = np.asarray(a, dtype=dtype)
        d = cls(ctx, a.shape, a.dtype)
        d.copy_from(a)
        return d

    def copy_from(self, a):
        a = np.asarray(a, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"size mismatch: {a.size} vs {self.size}")
        flat = np.ascontiguousarray(a.ravel(order="F"))
        self.ctx.check(self.ctx.lib.jp_memcpy_h2d(self.ctx.h, self.ptr, flat.ctypes.data,
                                                   self.nbytes), "h2d")

    def to_host(self) -> np.ndarray:
        flat = np.empty(self.size, dtype=self.dtype)
        self.ctx.check(self.ctx.lib.jp_memcpy_d2h(self.ctx.h, flat.ctypes.data, self.ptr,
                                                   self.nbytes), "d2h")
        return flat.reshape(self.shape, order="F")

    def flat(self) -> np.ndarray:
        return self.to_host().ravel(order="F")

    def fill(self, value):
        lib, h = self.ctx.lib, self.ctx.h
        if self.dtype == np.float64:
            self.ctx.check(lib.jp_fill_f64(h, self.ptr, float(value), self.size), "fill")
        elif self.dtype == np.float32:
            self.ctx.check(lib.jp_fill_f32(h, self.ptr, float(value), self.size), "fill")
        elif self.dtype.itemsize == 1:
            self.ctx.check(lib.jp_fill_u8(h, self.ptr, int(value), self.size), "fill")
        else:
            self.copy_from(np.full(self.size, value, dtype=self.dtype))
        return self

    def copy(self):
        o = DeviceArray(self.ctx, self.shape, self.dtype)
        o.copy_from_device(self)
        return o

    def copy_from_device(self, src: "DeviceArray"):
        if src.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {src.nbytes} vs {self.nbytes} bytes")
        self.ctx.check(self.ctx.lib.jp_memcpy_d2d(self.ctx.h, self.ptr, src.ptr, self.nbytes), "d2d")

    def free(self):
        if self.ptr:
            self.ctx.lib.jp_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h.value:
                self.free()
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"
