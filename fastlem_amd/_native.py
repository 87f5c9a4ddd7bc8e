"""Loader for the native library ``libfastlem_hip.so`` (HIP kernels + C ABI).

The library is built in-tree by ``fastlem_amd/build.py`` (``hipcc
--offload-arch=gfx950``).  There is no fallback: if the shared object is
missing, importing the solver raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfastlem_hip.so")

_lib = None

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)


class TerrainDesc(ctypes.Structure):
    """Mirror of ``fl_terrain_desc`` in include/fastlem_hip.h."""

    _fields_ = [
        ("n", ctypes.c_uint64),
        ("nnz", ctypes.c_uint64),
        ("row_ptr", c_u32p),
        ("col", c_u32p),
        ("length", c_f64p),
        ("area", c_f64p),
        ("erodibility", c_f64p),
        ("uplift_rate", c_f64p),
        ("has_max_slope", c_u8p),
        ("max_slope", c_f64p),
        ("init_elevation", c_f64p),
        ("outlets", c_u32p),
        ("n_outlets", ctypes.c_uint64),
        ("max_iteration", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class SolveStats(ctypes.Structure):
    """Mirror of ``fl_solve_stats``."""

    _fields_ = [("iterations", ctypes.c_uint32), ("lake_iterations", ctypes.c_uint32),
                ("host_syncs", ctypes.c_uint32), ("order_from_forest", ctypes.c_uint32),
                ("visit_order_ms", ctypes.c_double), ("solve_ms", ctypes.c_double),
                ("outlet_phase_ms", ctypes.c_double), ("forest_ms", ctypes.c_double),
                ("ranks_ms", ctypes.c_double), ("compactions", ctypes.c_uint32), ("redone", ctypes.c_uint32),
                ("create_validate_ms", ctypes.c_double), ("create_relabel_ms", ctypes.c_double),
                ("create_upload_ms", ctypes.c_double), ("create_on_device", ctypes.c_uint32),
                ("create_clusters", ctypes.c_uint32), ("recomputed_sites", ctypes.c_uint64),
                ("first_sweep_rounds", ctypes.c_uint32), ("max_sweep_rounds", ctypes.c_uint32),
                ("rehinted", ctypes.c_uint32)]


class ServerCounts(ctypes.Structure):
    """Mirror of ``fl_server_counts``."""

    _fields_ = [("joins_idle", ctypes.c_uint64), ("joins_busy", ctypes.c_uint64), ("steps", ctypes.c_uint64),
                ("max_lane_terrains", ctypes.c_uint32), ("lanes", ctypes.c_uint32),
                ("prep_workers", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("prep_ms", ctypes.c_double),
                ("step_ms", ctypes.c_double), ("join_ms", ctypes.c_double), ("idle_ms", ctypes.c_double)]


class Tuning(ctypes.Structure):
    """Mirror of ``fl_tuning``."""

    _fields_ = [("site_block", ctypes.c_uint32), ("long_chain", ctypes.c_uint32), ("tail_units", ctypes.c_uint32),
                ("pipelining", ctypes.c_uint32), ("compact_min_idle", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class NativeError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"fastlem_hip error {code}: {message}")
        self.code = code


def _declare(lib):
    vp = ctypes.c_void_p
    i32 = ctypes.c_int32
    sz = ctypes.c_size_t
    f64 = ctypes.c_double
    u32 = ctypes.c_uint32
    u64 = ctypes.c_uint64
    sig = {
        "fl_last_error": (ctypes.c_char_p, []),
        "fl_version": (ctypes.c_char_p, []),
        "fl_ctx_create": (i32, [i32, ctypes.POINTER(vp)]),
        "fl_ctx_destroy": (i32, [vp]),
        "fl_generate_f64": (i32, [vp, ctypes.POINTER(TerrainDesc), c_f64p, c_u32p]),
        "fl_generate_f64_noalloc": (i32, [vp, ctypes.POINTER(TerrainDesc), c_f64p, c_u32p]),
        "fl_terrain_create": (i32, [vp, ctypes.POINTER(TerrainDesc), ctypes.POINTER(vp)]),
        "fl_terrain_destroy": (i32, [vp]),
        "fl_terrain_solve": (i32, [vp, vp, u32, c_u32p]),
        "fl_terrain_download": (i32, [vp, vp, c_f64p]),
        "fl_terrain_device_elevations": (i32, [vp, ctypes.POINTER(vp)]),
        "fl_terrain_device_bytes": (i32, [vp, ctypes.POINTER(u64)]),
        "fl_terrain_invalidate_cache": (i32, [vp]),
        "fl_batch_create": (i32, [vp, ctypes.POINTER(TerrainDesc), u32, ctypes.POINTER(vp)]),
        "fl_batch_solve": (i32, [vp, vp, c_u32p]),
        "fl_batch_download": (i32, [vp, vp, u32, c_f64p]),
        "fl_batch_count": (i32, [vp, c_u32p]),
        "fl_server_create": (i32, [i32, u32, u64, ctypes.POINTER(vp)]),
        "fl_server_submit": (i32, [vp, ctypes.POINTER(TerrainDesc), c_f64p, ctypes.POINTER(u64)]),
        "fl_server_wait": (i32, [vp, u64, c_u32p, ctypes.POINTER(f64), ctypes.POINTER(f64)]),
        "fl_server_stats": (i32, [vp, ctypes.POINTER(u64), c_u32p]),
        "fl_server_destroy": (i32, [vp]),
        "fl_server_counters": (i32, [vp, ctypes.POINTER(ServerCounts)]),
        "fl_terrain_last_stats": (i32, [vp, ctypes.POINTER(SolveStats)]),
        "fl_ctx_set_profiling": (i32, [vp, i32]),
        "fl_ctx_set_batch_compaction": (i32, [vp, u32]),
        "fl_ctx_set_pipelining": (i32, [vp, i32]),
        "fl_ctx_set_recompute_all": (i32, [vp, i32]),
        "fl_ctx_set_rehint_rounds": (i32, [vp, u32]),
        "fl_ctx_kernel_time": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(u64), ctypes.POINTER(f64),
                                     ctypes.POINTER(u64)]),
        "fl_ctx_reset_profile": (i32, [vp]),
        "fl_ctx_set_tuning": (i32, [vp, ctypes.POINTER(Tuning)]),
        "fl_ctx_tuning_for": (i32, [vp, u64, u32, ctypes.POINTER(Tuning)]),
        "fl_autotune": (i32, [vp, vp, u32, ctypes.POINTER(Tuning), ctypes.POINTER(f64), ctypes.POINTER(f64), c_u32p]),
        "fl_stage_receivers": (i32, [vp, vp, c_f64p, c_u32p]),
        "fl_stage_stream_tree": (i32, [vp, vp, c_f64p, c_u32p, c_u32p, ctypes.POINTER(i32)]),
        "fl_stage_iteration": (i32, [vp, vp, c_f64p, c_u32p, c_f64p, c_f64p, ctypes.POINTER(i32)]),
        "fl_stage_basin_order": (i32, [vp, vp, c_f64p, u32, c_u32p, c_u32p]),
        "fl_stage_visit_order": (i32, [vp, vp, c_u32p]),
        "fl_request_device_bytes": (i32, [ctypes.POINTER(TerrainDesc), ctypes.POINTER(u64)]),
        "fl_interp_create": (i32, [vp, u64, c_f64p, c_f64p, u64, c_u32p, ctypes.POINTER(vp)]),
        "fl_interp_query": (i32, [vp, vp, c_f64p, u64, c_f64p, c_f64p, c_f64p, c_u8p]),
        "fl_interp_destroy": (i32, [vp]),
        "fl_jitter_elevations": (None, [u64, sz, c_f64p, c_f64p]),
        "fl_random_sites": (None, [sz, f64, f64, f64, f64, c_f64p, c_f64p]),
        "fl_chacha_block": (None, [ctypes.c_int, c_u32p, c_u32p]),
        "fl_voronoi_cells": (ctypes.c_int, [sz, c_f64p, c_f64p, sz, c_i64p, f64, f64, f64, f64,
                                            c_f64p, c_f64p, c_f64p]),
        "fl_build_adjacency": (ctypes.c_int, [sz, c_f64p, c_f64p, sz, c_i64p, c_u32p, c_u32p, c_f64p]),
        "fl_lake_visit_order": (ctypes.c_int, [sz, c_u32p, c_u32p, c_f64p, c_u32p, sz, c_u32p,
                                               ctypes.POINTER(u64)]),
        "fl_outlet_pop_order": (ctypes.c_int, [sz, c_u32p, c_u32p, c_u32p, sz, c_u32p]),
        "fl_ranks_from_msf": (ctypes.c_int, [sz, c_u32p, sz, c_u32p, c_u32p, sz, c_u32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """Return the loaded native library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m fastlem_amd.build` "
                "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def check(code: int) -> None:
    if code != 0:
        msg = lib().fl_last_error()
        raise NativeError(code, msg.decode() if msg else "")


def ptr(a, ctype):
    """ctypes pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(ctype))
    return a.ctypes.data_as(ctypes.POINTER(ctype))
