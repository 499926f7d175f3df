"""ctypes mirror of include/justpic_hip.h and loader of the HIP library.

The library is loaded from this package's directory (``libjustpic_hip.so``,
built in-tree by ``justpic.jl_amd/csrc/Makefile``).  There is no fallback: if
the library is missing every entry point raises, so a run can never silently
compute on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

JP_OK = 0
JP_ERR_ARG, JP_ERR_HIP, JP_ERR_STATE = -1, -2, -3
JP_EULER, JP_RK2, JP_RK4 = 0, 1, 2
JP_MAX_ARGS = 16
JP_ENS_VARIANTS = 8
JP_ADJ_MAX_EVERY = 16
JP_OPT_SLOT_SPAN, JP_OPT_INJECT_WAVES, JP_OPT_GEOMETRY_RECHECK = 1, 2, 3
JP_OPT_MOVE_PHASES, JP_OPT_ADVECT_SLOTS, JP_OPT_INJECT_PHASES = 4, 5, 6
JP_NOALLOC_MAX_CELLS = 1 << 21
JP_MAX_PHASES = 16
JP_INTERP_LINEAR, JP_INTERP_MQS, JP_INTERP_LINP = 0, 1, 2

LIB_NAME = "libjustpic_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)


class jp_axis(C.Structure):
    _fields_ = [
        ("x", C.c_void_p),
        ("dx", C.c_void_p),
        ("inv_dx", C.c_void_p),
        ("n", C.c_int64),
    ]


class jp_cellgrid(C.Structure):
    _fields_ = [
        ("ndim", C.c_int32),
        ("max_xcell", C.c_int32),
        ("ni", C.c_int64 * 3),
        ("xv", jp_axis * 3),
    ]


class jp_velgrid(C.Structure):
    _fields_ = [("ax", (jp_axis * 3) * 3)]


class jp_stats(C.Structure):
    _fields_ = [
        ("moved", C.c_int64),
        ("deleted", C.c_int64),
        ("dropped", C.c_int64),
        ("far", C.c_int64),
        ("injected", C.c_int64),
        ("orphan", C.c_int64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


def ptr_array(ptrs):
    """Host array of (device or host) pointers, as `double* const*`."""
    arr = (C.c_void_p * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


class HipLibraryMissing(RuntimeError):
    pass


_lib = None

_P = C.c_void_p
_SIGS = {
    "jp_ctx_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "jp_ctx_destroy": (C.c_int, [_P]),
    "jp_last_error": (C.c_char_p, [_P]),
    "jp_ctx_set_sync": (C.c_int, [_P, C.c_int]),
    "jp_ctx_set_stream": (C.c_int, [_P, _P]),
    "jp_ctx_stream": (_P, [_P]),
    "jp_ctx_set_option": (C.c_int, [_P, C.c_int, C.c_int64]),
    "jp_sync": (C.c_int, [_P]),
    "jp_malloc": (C.c_int, [_P, C.POINTER(C.c_void_p), C.c_size_t]),
    "jp_free": (C.c_int, [_P, _P]),
    "jp_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "jp_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "jp_memcpy_d2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "jp_fill_u8": (C.c_int, [_P, _P, C.c_uint8, C.c_int64]),
    "jp_fill_f64": (C.c_int, [_P, _P, C.c_double, C.c_int64]),
    "jp_fill_f32": (C.c_int, [_P, _P, C.c_float, C.c_int64]),
    "jp_cellarray_to_cpu_layout": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int, C.c_int]),
    "jp_cellarray_from_cpu_layout": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int, C.c_int]),
    "jp_event_create": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "jp_event_destroy": (C.c_int, [_P, _P]),
    "jp_event_record": (C.c_int, [_P, _P]),
    "jp_event_wait": (C.c_int, [_P, _P]),
    "jp_event_elapsed_ms": (C.c_int, [_P, _P, _P, C.POINTER(C.c_float)]),
    "jp_graph_begin": (C.c_int, [_P]),
    "jp_graph_end": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "jp_graph_launch": (C.c_int, [_P, _P]),
    "jp_graph_destroy": (C.c_int, [_P, _P]),
    "jp_init_particles": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, C.c_int, C.c_uint64]),
    "jp_advection": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P,
                               C.POINTER(jp_velgrid), C.c_int, C.c_double, C.c_double]),
    "jp_advection_scheme": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P,
                                      C.POINTER(jp_velgrid), C.c_int, C.c_double, C.c_double,
                                      C.c_int]),
    "jp_semilagrangian": (C.c_int, [_P, C.c_int, _P, C.POINTER(jp_velgrid), _P, C.c_int, _P, _P,
                                    C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]),
    "jp_move_particles": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, C.c_int, _P,
                                    C.POINTER(jp_stats)]),
    "jp_particle2grid": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, _P]),
    "jp_particle2grid_noalloc": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, _P]),
    "jp_particle2grid_multi": (C.c_int, [_P, C.POINTER(jp_cellgrid), C.c_int, _P, _P, _P, _P]),
    "jp_grid2particle": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, _P]),
    "jp_inject_particles": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, C.c_int, _P,
                                      C.c_int, C.c_uint64, C.c_uint64, C.POINTER(jp_stats)]),
    "jp_philox4x32_10": (None, [_P, _P, _P]),
    "jp_inject_particles_phase": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, _P, C.c_int,
                                            _P, _P, _P, C.c_int, C.c_uint64, C.c_uint64,
                                            C.POINTER(jp_stats)]),
    "jp_update_phase_ratios": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, C.c_int, _P,
                                         _P, _P, _P]),
    "jp_cell_index": (C.c_int, [_P, C.c_int, C.POINTER(jp_axis), C.c_int64, _P, _P]),
    "jp_grid2particle_flip": (C.c_int, [_P, C.POINTER(jp_cellgrid), C.c_int, _P, _P, _P, _P, _P,
                                        C.c_double]),
    "jp_particle2centroid": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, C.c_int, _P, _P, _P,
                                       C.c_int]),
    "jp_centroid2particle": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, C.c_int, _P, _P, _P]),
    "jp_split_from_f64": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P]),
    "jp_split_to_f64": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P]),
    "jp_advection_split": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P,
                                     C.POINTER(jp_velgrid), C.c_int, C.c_double, C.c_double]),
    "jp_move_particles_split": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, C.c_int, _P,
                                          C.POINTER(jp_stats)]),
    "jp_particle2grid_split": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, _P]),
    "jp_grid2particle_split": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, _P, _P]),
    "jp_inject_particles_split": (C.c_int, [_P, C.POINTER(jp_cellgrid), _P, _P, C.c_int, _P,
                                            C.c_int, C.c_uint64, C.c_uint64,
                                            C.POINTER(jp_stats)]),
    "jp_ensemble_create": (C.c_int, [_P, C.c_int, _P, _P, _P, _P, _P, _P, C.c_int, C.c_double,
                                     C.POINTER(C.c_void_p)]),
    "jp_ensemble_advect": (C.c_int, [_P, _P]),
    "jp_ensemble_set_variant": (C.c_int, [_P, _P, C.c_int]),
    "jp_ensemble_tune": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_int), _P]),
    "jp_ensemble_destroy": (C.c_int, [_P, _P]),
    "jp_adjoint_forward": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double, C.c_int,
                                     C.c_int, C.c_int64, _P, _P, _P, _P, _P, _P, _P]),
    "jp_adjoint_forward_split": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double,
                                           C.c_int, C.c_int, C.c_int64, _P, _P, _P, _P, _P, _P,
                                           _P]),
    "jp_adjoint_backward": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_double, C.c_double, C.c_int,
                                      C.c_int, C.c_int64, _P, _P, _P, _P, _P, _P, _P, _P,
                                      _P]),
}


def exported_symbols():
    """Names every C-ABI entry point the package binds (checked against the header)."""
    return sorted(_SIGS)


def load_library(path: str | None = None):
    """Load the HIP library once; raise HipLibraryMissing if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise HipLibraryMissing(
            f"{p} is not built; run `make -C justpic.jl_amd/csrc` "
            "(or __graft_entry__.build()) first")
    lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def f64ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous or a.flags.f_contiguous
    return a.ctypes.data_as(C.c_void_p).value
