"""Device-resident terrains: upload once, solve many times (C ABI
``fl_terrain_*``), plus the teacher-forced stage entry points used to check
each part of the iteration against the reference algorithm."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import TerrainDesc, check, ptr
from .generator import UNBOUNDED, ParameterArrays, context, make_desc


class DeviceTerrain:
    def __init__(self, row_ptr, col, length, area, erodibility, uplift, init_elevation, outlets,
                 max_slope=None, max_iteration: int = UNBOUNDED):
        class _M:
            pass

        m = _M()
        m.graph = lambda: (np.ascontiguousarray(row_ptr, np.uint32), np.ascontiguousarray(col, np.uint32),
                           np.ascontiguousarray(length, np.float64))
        g = m.graph()
        m.num = lambda: len(g[0]) - 1
        a = np.ascontiguousarray(area, np.float64)
        m.areas = lambda: a
        self._g = g
        n = m.num()
        self.n = n
        params = ParameterArrays(np.zeros(n), erodibility, uplift, np.zeros(n, bool), max_slope)
        self.init_elevation = np.ascontiguousarray(init_elevation, np.float64)
        self.outlets = np.ascontiguousarray(outlets, np.uint32)
        self._keep: list = [g, a]
        desc = make_desc(m, params, self.init_elevation, self.outlets, max_iteration, self._keep)
        # make_desc reads graph() again: keep the same arrays alive
        self._desc = desc
        self._ctx = context()
        h = ctypes.c_void_p()
        check(_native.lib().fl_terrain_create(self._ctx, ctypes.byref(desc), ctypes.byref(h)))
        self._h = h
        self.max_iteration = max_iteration

    @staticmethod
    def from_workload(w, init_elevation=None, max_iteration: int = UNBOUNDED) -> "DeviceTerrain":
        from .generator import jitter_elevations

        rp, col, ln = w.model.graph()
        init = jitter_elevations(w.base, 0) if init_elevation is None else init_elevation
        return DeviceTerrain(rp, col, ln, w.model.areas(), w.erodibility, w.uplift, init, w.outlets(),
                             w.max_slope, max_iteration)

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().fl_terrain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, max_iteration: int = None) -> int:
        it = ctypes.c_uint32()
        mi = self.max_iteration if max_iteration is None else max_iteration
        check(_native.lib().fl_terrain_solve(self._ctx, self._h, mi, ctypes.byref(it)))
        return it.value

    def elevations(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        check(_native.lib().fl_terrain_download(self._ctx, self._h, ptr(out, ctypes.c_double)))
        return out

    def last_stats(self) -> dict:
        st = _native.SolveStats()
        check(_native.lib().fl_terrain_last_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def device_bytes(self) -> int:
        b = ctypes.c_uint64()
        check(_native.lib().fl_terrain_device_bytes(self._h, ctypes.byref(b)))
        return b.value

    # ---- teacher-forced stages ----
    def stage_receivers(self, elev) -> np.ndarray:
        elev = np.ascontiguousarray(elev, np.float64)
        nxt = np.empty(self.n, np.uint32)
        check(_native.lib().fl_stage_receivers(self._ctx, self._h, ptr(elev, ctypes.c_double),
                                               ptr(nxt, ctypes.c_uint32)))
        return nxt

    def stage_stream_tree(self, elev):
        elev = np.ascontiguousarray(elev, np.float64)
        nxt = np.empty(self.n, np.uint32)
        sub = np.empty(self.n, np.uint32)
        flag = ctypes.c_int32()
        check(_native.lib().fl_stage_stream_tree(self._ctx, self._h, ptr(elev, ctypes.c_double),
                                                 ptr(nxt, ctypes.c_uint32), ptr(sub, ctypes.c_uint32),
                                                 ctypes.byref(flag)))
        return nxt, sub, bool(flag.value)

    def stage_basin_order(self, elev, outlet: int) -> np.ndarray:
        elev = np.ascontiguousarray(elev, np.float64)
        out = np.empty(self.n, np.uint32)
        ln = ctypes.c_uint32()
        check(_native.lib().fl_stage_basin_order(self._ctx, self._h, ptr(elev, ctypes.c_double), outlet,
                                                 ptr(out, ctypes.c_uint32), ctypes.byref(ln)))
        return out[:ln.value]

    def stage_iteration(self, elev):
        h = np.ascontiguousarray(elev, np.float64).copy()
        nxt = np.empty(self.n, np.uint32)
        da = np.empty(self.n)
        rt = np.empty(self.n)
        flag = ctypes.c_int32()
        check(_native.lib().fl_stage_iteration(self._ctx, self._h, ptr(h, ctypes.c_double),
                                               ptr(nxt, ctypes.c_uint32), ptr(da, ctypes.c_double),
                                               ptr(rt, ctypes.c_double), ctypes.byref(flag)))
        return h, nxt, da, rt, bool(flag.value)


def batch_descs(terrains, keep: list):
    """``fl_terrain_desc[]`` for a list of terrain dicts (keys row_ptr, col,
    length, area, erodibility, uplift, init_elevation, outlets and optionally
    max_slope, max_iteration); ``keep`` holds the arrays alive."""
    descs = (TerrainDesc * len(terrains))()
    for i, tr in enumerate(terrains):
        class _M:
            pass

        m = _M()
        g = (np.ascontiguousarray(tr["row_ptr"], np.uint32), np.ascontiguousarray(tr["col"], np.uint32),
             np.ascontiguousarray(tr["length"], np.float64))
        a = np.ascontiguousarray(tr["area"], np.float64)
        m.graph = (lambda g=g: g)
        m.num = (lambda g=g: len(g[0]) - 1)
        m.areas = (lambda a=a: a)
        n = m.num()
        params = ParameterArrays(np.zeros(n), tr["erodibility"], tr["uplift"], np.zeros(n, bool),
                                 tr.get("max_slope"))
        init = np.ascontiguousarray(tr["init_elevation"], np.float64)
        outlets = np.ascontiguousarray(tr["outlets"], np.uint32)
        keep.extend([g, a, init, outlets, params])
        descs[i] = make_desc(m, params, init, outlets, tr.get("max_iteration", UNBOUNDED), keep)
    return descs


class DeviceBatch:
    """Several terrains iterated together (C ABI ``fl_batch_*``): one launch
    sequence per iteration for all of them; each terrain stops at its own
    convergence or ``max_iteration`` and keeps the result of its solo run."""

    def __init__(self, terrains):
        """terrains: list of dicts with keys row_ptr, col, length, area,
        erodibility, uplift, init_elevation, outlets and optionally max_slope,
        max_iteration."""
        self._keep: list = []
        descs = batch_descs(terrains, self._keep)
        self.sizes = [int(d.n) for d in descs]
        self._descs = descs
        self._ctx = context()
        h = ctypes.c_void_p()
        check(_native.lib().fl_batch_create(self._ctx, descs, len(terrains), ctypes.byref(h)))
        self._h = h

    @staticmethod
    def terrain_of(w, init_elevation=None, max_iteration: int = UNBOUNDED) -> dict:
        from .generator import jitter_elevations

        rp, col, ln = w.model.graph()
        return dict(row_ptr=rp, col=col, length=ln, area=w.model.areas(), erodibility=w.erodibility,
                    uplift=w.uplift, outlets=w.outlets(), max_slope=w.max_slope, max_iteration=max_iteration,
                    init_elevation=jitter_elevations(w.base, 0) if init_elevation is None else init_elevation)

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().fl_terrain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, compact_min_idle: int = None) -> np.ndarray:
        """Iteration counts per terrain.  compact_min_idle: idle sites before
        the running terrains are gathered into a smaller union (the library's
        default when None)."""
        its = np.zeros(len(self.sizes), np.uint32)
        L = _native.lib()
        if compact_min_idle is not None:
            check(L.fl_ctx_set_batch_compaction(self._ctx, int(compact_min_idle)))
        try:
            check(L.fl_batch_solve(self._ctx, self._h, ptr(its, ctypes.c_uint32)))
        finally:
            if compact_min_idle is not None:
                check(L.fl_ctx_set_batch_compaction(self._ctx, 1 << 16))
        return its

    def elevations(self, index: int) -> np.ndarray:
        out = np.empty(self.sizes[index], np.float64)
        check(_native.lib().fl_batch_download(self._ctx, self._h, index, ptr(out, ctypes.c_double)))
        return out

    def last_stats(self) -> dict:
        st = _native.SolveStats()
        check(_native.lib().fl_terrain_last_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}
