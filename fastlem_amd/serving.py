"""Serving: generate() requests through the library's continuous-batching
request loop (C ABI ``fl_server_*``).

``Server(slots, hbm_budget_bytes)`` starts native workers on up to four
streams: preparation workers (validation, upload, visit order) and lanes,
each iterating a union of terrains.  ``submit`` enqueues a request and
returns a ticket at once; a request is admitted as soon as its HBM estimate
fits the budget, and once prepared it joins a lane at the lane's next
iteration boundary, sharing the launches of the terrains already running.
``wait`` returns (elevations, iterations, queue_ms, latency_ms).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import check, ptr
from .generator import ParameterArrays, make_desc


def request(terrain: dict):
    """The generate() call of a terrain dict (as for DeviceBatch): its
    ``fl_terrain_desc``, the arrays it points to (kept alive by the caller)
    and the output array."""

    class _M:
        pass

    m = _M()
    g = (np.ascontiguousarray(terrain["row_ptr"], np.uint32), np.ascontiguousarray(terrain["col"], np.uint32),
         np.ascontiguousarray(terrain["length"], np.float64))
    a = np.ascontiguousarray(terrain["area"], np.float64)
    m.graph = lambda: g
    m.num = lambda: len(g[0]) - 1
    m.areas = lambda: a
    n = m.num()
    params = ParameterArrays(np.zeros(n), terrain["erodibility"], terrain["uplift"], np.zeros(n, bool),
                             terrain.get("max_slope"))
    init = np.ascontiguousarray(terrain["init_elevation"], np.float64)
    outlets = np.ascontiguousarray(terrain["outlets"], np.uint32)
    keep = [g, a, params, init, outlets]
    desc = make_desc(m, params, init, outlets, terrain.get("max_iteration", 0xFFFFFFFF), keep)
    return desc, keep, np.empty(n, np.float64)


def request_bytes(terrain: dict) -> int:
    """HBM the generate() call of a terrain dict holds at its peak (the
    server's admission estimate, C ABI ``fl_request_device_bytes``)."""
    desc, keep, _ = request(terrain)
    b = ctypes.c_uint64()
    check(_native.lib().fl_request_device_bytes(ctypes.byref(desc), ctypes.byref(b)))
    return b.value


class Server:
    def __init__(self, slots: int = 8, hbm_budget_bytes: int = 64 << 30, device: int = 0):
        h = ctypes.c_void_p()
        check(_native.lib().fl_server_create(device, slots, hbm_budget_bytes, ctypes.byref(h)))
        self._h = h
        self._pending: dict = {}

    def submit(self, terrain: dict) -> int:
        """terrain: dict as for fastlem_amd.device.DeviceBatch."""
        desc, keep, out = request(terrain)
        ticket = ctypes.c_uint64()
        check(_native.lib().fl_server_submit(self._h, ctypes.byref(desc), ptr(out, ctypes.c_double),
                                             ctypes.byref(ticket)))
        self._pending[ticket.value] = (desc, keep, out)
        return ticket.value

    def wait(self, ticket: int):
        its = ctypes.c_uint32()
        q = ctypes.c_double()
        lat = ctypes.c_double()
        rc = _native.lib().fl_server_wait(self._h, ticket, ctypes.byref(its), ctypes.byref(q), ctypes.byref(lat))
        desc, keep, out = self._pending.pop(ticket)
        check(rc)
        return out, its.value, q.value, lat.value

    def stats(self):
        b = ctypes.c_uint64()
        r = ctypes.c_uint32()
        check(_native.lib().fl_server_stats(self._h, ctypes.byref(b), ctypes.byref(r)))
        c = _native.ServerCounts()
        check(_native.lib().fl_server_counters(self._h, ctypes.byref(c)))
        out = {"peak_bytes": b.value, "peak_running": r.value, "joins_idle_lane": c.joins_idle,
               "joins_busy_lane": c.joins_busy, "max_lane_terrains": c.max_lane_terrains}
        out.update({f: getattr(c, f) for f in ("steps", "lanes", "prep_workers", "prep_ms", "step_ms", "join_ms",
                                                "idle_ms")})
        return out

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().fl_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
