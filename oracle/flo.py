"""ctypes binding of oracle/fl_oracle.c (test infrastructure only).

The shared object is built by ``oracle/Makefile`` into ``oracle/build``;
``load()`` builds it with gcc when it is missing.
"""
from __future__ import annotations

import ctypes as c
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "build", "libfl_oracle.so")
NONE = 0xFFFFFFFF

_u32p = c.POINTER(c.c_uint32)
_f64p = c.POINTER(c.c_double)
_u8p = c.POINTER(c.c_uint8)


class Stats(c.Structure):
    _fields_ = [("iterations", c.c_uint32), ("lake_iterations", c.c_uint32), ("seconds_tree", c.c_double)]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return SO


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = c.CDLL(SO)
        sz = c.c_size_t
        L.flo_generate.restype = c.c_int
        L.flo_generate.argtypes = [sz, _u32p, _u32p, _f64p, _f64p, _f64p, _f64p, _u8p, _f64p, _f64p, _u32p, sz,
                                   c.c_uint32, _f64p, c.POINTER(Stats)]
        L.flo_receivers.restype = c.c_int
        L.flo_receivers.argtypes = [sz, _u32p, _u32p, _f64p, _f64p, _u32p, sz, _u32p]
        L.flo_find_roots.restype = c.c_int
        L.flo_find_roots.argtypes = [sz, _u32p, _u32p, sz, _u32p]
        L.flo_stream_tree.restype = c.c_int
        L.flo_stream_tree.argtypes = [sz, _u32p, _u32p, _f64p, _f64p, _u32p, sz, _u32p, _u32p]
        L.flo_remove_lakes.restype = c.c_int
        L.flo_remove_lakes.argtypes = [sz, _u32p, _u32p, _f64p, _u32p, sz, _u32p, _u32p]
        L.flo_visit_order.restype = c.c_long
        L.flo_visit_order.argtypes = [sz, _u32p, _u32p, _f64p, _u32p, sz, _u32p]
        L.flo_basin_order.restype = c.c_long
        L.flo_basin_order.argtypes = [sz, _u32p, _u32p, _f64p, _u32p, c.c_uint32, _u32p]
        L.flo_iteration.restype = c.c_int
        L.flo_iteration.argtypes = [sz, _u32p, _u32p, _f64p, _f64p, _f64p, _f64p, _u8p, _f64p, _u32p, sz, _f64p,
                                    _u32p, _u32p, _f64p, _f64p, c.POINTER(c.c_int)]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(c.POINTER(t))


class Graph:
    """CSR graph (rows = neighbors_of(i) in order)."""

    def __init__(self, row_ptr, col, length):
        self.row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        self.col = np.ascontiguousarray(col, np.uint32)
        self.length = np.ascontiguousarray(length, np.float64)

    @property
    def n(self):
        return len(self.row_ptr) - 1

    def args(self):
        return (c.c_size_t(self.n), _p(self.row_ptr, c.c_uint32), _p(self.col, c.c_uint32),
                _p(self.length, c.c_double))


def _ms(max_slope):
    if max_slope is None:
        return None, None
    ms = np.ascontiguousarray(max_slope, np.float64)
    has = np.ascontiguousarray(~np.isnan(ms), np.uint8)
    return has, np.ascontiguousarray(np.where(np.isnan(ms), 0.0, ms))


def generate(g: Graph, area, erod, uplift, init_elev, outlets, max_iteration=NONE, max_slope=None):
    """Returns (elevations, iterations, lake_iterations)."""
    L = load()
    area, erod, uplift, init_elev = (np.ascontiguousarray(a, np.float64) for a in (area, erod, uplift, init_elev))
    outlets = np.ascontiguousarray(outlets, np.uint32)
    has, ms = _ms(max_slope)
    out = np.empty(g.n, np.float64)
    st = Stats()
    rc = L.flo_generate(*g.args(), _p(area, c.c_double), _p(erod, c.c_double), _p(uplift, c.c_double),
                        _p(has, c.c_uint8), _p(ms, c.c_double), _p(init_elev, c.c_double),
                        _p(outlets, c.c_uint32), c.c_size_t(len(outlets)), c.c_uint32(max_iteration),
                        _p(out, c.c_double), c.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle generate failed: {rc}")
    return out, st.iterations, st.lake_iterations


def receivers(g: Graph, elev, outlets):
    L = load()
    elev = np.ascontiguousarray(elev, np.float64)
    outlets = np.ascontiguousarray(outlets, np.uint32)
    nxt = np.empty(g.n, np.uint32)
    rc = L.flo_receivers(*g.args(), _p(elev, c.c_double), _p(outlets, c.c_uint32), len(outlets),
                         _p(nxt, c.c_uint32))
    assert rc == 0
    return nxt


def find_roots(nxt, outlets):
    L = load()
    nxt = np.ascontiguousarray(nxt, np.uint32)
    outlets = np.ascontiguousarray(outlets, np.uint32)
    sub = np.empty(len(nxt), np.uint32)
    hl = L.flo_find_roots(len(nxt), _p(nxt, c.c_uint32), _p(outlets, c.c_uint32), len(outlets),
                          _p(sub, c.c_uint32))
    assert hl >= 0
    return sub, bool(hl)


def stream_tree(g: Graph, elev, outlets):
    """Returns (next, pre-join subroot, has_lake)."""
    L = load()
    elev = np.ascontiguousarray(elev, np.float64)
    outlets = np.ascontiguousarray(outlets, np.uint32)
    nxt = np.empty(g.n, np.uint32)
    sub = np.empty(g.n, np.uint32)
    hl = L.flo_stream_tree(*g.args(), _p(elev, c.c_double), _p(outlets, c.c_uint32), len(outlets),
                           _p(nxt, c.c_uint32), _p(sub, c.c_uint32))
    assert hl >= 0, hl
    return nxt, sub, bool(hl)


def remove_lakes(g: Graph, outlets, subroot, nxt):
    L = load()
    outlets = np.ascontiguousarray(outlets, np.uint32)
    subroot = np.ascontiguousarray(subroot, np.uint32)
    out = np.ascontiguousarray(nxt, np.uint32).copy()
    rc = L.flo_remove_lakes(*g.args(), _p(outlets, c.c_uint32), len(outlets), _p(subroot, c.c_uint32),
                            _p(out, c.c_uint32))
    assert rc == 0, rc
    return out


def visit_order(g: Graph, outlets):
    L = load()
    outlets = np.ascontiguousarray(outlets, np.uint32)
    order = np.empty(g.n, np.uint32)
    cnt = L.flo_visit_order(*g.args(), _p(outlets, c.c_uint32), len(outlets), _p(order, c.c_uint32))
    assert cnt >= 0, cnt
    return order[:cnt]


def basin_order(g: Graph, nxt, outlet):
    L = load()
    nxt = np.ascontiguousarray(nxt, np.uint32)
    out = np.empty(g.n, np.uint32)
    cnt = L.flo_basin_order(*g.args(), _p(nxt, c.c_uint32), c.c_uint32(outlet), _p(out, c.c_uint32))
    assert cnt >= 0
    return out[:cnt]


def iteration(g: Graph, area, erod, uplift, outlets, elev, max_slope=None):
    """One outer iteration.  Returns (new_elev, next, subroot, da, rt, changed, had_lake)."""
    L = load()
    area, erod, uplift = (np.ascontiguousarray(a, np.float64) for a in (area, erod, uplift))
    outlets = np.ascontiguousarray(outlets, np.uint32)
    h = np.ascontiguousarray(elev, np.float64).copy()
    has, ms = _ms(max_slope)
    nxt = np.empty(g.n, np.uint32)
    sub = np.empty(g.n, np.uint32)
    da = np.empty(g.n)
    rt = np.empty(g.n)
    hl = c.c_int()
    ch = L.flo_iteration(*g.args(), _p(area, c.c_double), _p(erod, c.c_double), _p(uplift, c.c_double),
                         _p(has, c.c_uint8), _p(ms, c.c_double), _p(outlets, c.c_uint32), len(outlets),
                         _p(h, c.c_double), _p(nxt, c.c_uint32), _p(sub, c.c_uint32), _p(da, c.c_double),
                         _p(rt, c.c_double), c.byref(hl))
    assert ch >= 0, ch
    return h, nxt, sub, da, rt, bool(ch), bool(hl.value)
