"""Terrains and checks of the allocation-free entry point
(fl_generate_f64_noalloc), shared by the CPU-emulation and GPU tests."""
import ctypes

import numpy as np

FL_ERR_TOO_LARGE = -4


def noalloc_workloads(scale: float = 1.0):
    """A fBm terrain of 60,000 sites, then a 3,000-site terrain with a 0.6
    rad max slope on half its sites (both scaled by `scale`; at full scale
    the oracle converges in 30 iterations: a max slope of 0.3 on half the
    sites can make the reference alternate forever)."""
    from fastlem_amd.workloads import Workload, config_fbm, config_plain

    nb, ns = int(60000 * scale), int(3000 * scale)
    big = config_fbm(nb, float(np.sqrt(nb)), seed=13)
    small = config_plain(ns, float(np.sqrt(ns)), 1.0)
    ms = np.full(small.n, np.nan)
    ms[: small.n // 2] = 0.6
    small = Workload("plain-3000-ms", small.model, small.erodibility, small.uplift, small.is_outlet, small.base, ms)
    return [big, small]


def call(L, fn, ctx, w):
    from fastlem_amd import _native
    from fastlem_amd.generator import ParameterArrays, jitter_elevations, make_desc

    keep: list = []
    params = ParameterArrays(w.base, w.erodibility, w.uplift, w.is_outlet, w.max_slope)
    # a cap both entry points honour alike: the calls end even on a terrain
    # that never converges
    d = make_desc(w.model, params, jitter_elevations(w.base, 0), np.ascontiguousarray(w.outlets(), np.uint32),
                  3000, keep)
    out = np.empty(w.n, np.float64)
    it = ctypes.c_uint32()
    rc = getattr(L, fn)(ctx, ctypes.byref(d), _native.ptr(out, ctypes.c_double), ctypes.byref(it))
    assert rc == 0, L.fl_last_error()
    return out, it.value


def too_large_desc():
    from fastlem_amd._native import TerrainDesc

    d = TerrainDesc()
    d.n = 2_000_000
    d.nnz = 12_000_000
    return d


def check_noalloc(L, ctx, scale: float = 1.0):
    for w in noalloc_workloads(scale):
        ref, rit = call(L, "fl_generate_f64", ctx, w)
        got, git = call(L, "fl_generate_f64_noalloc", ctx, w)
        assert git == rit, (w.name, git, rit)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), w.name
    out = np.empty(1)
    it = ctypes.c_uint32()
    d = too_large_desc()
    rc = L.fl_generate_f64_noalloc(ctx, ctypes.byref(d), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   ctypes.byref(it))
    assert rc == FL_ERR_TOO_LARGE
    assert b"static scratch" in L.fl_last_error()

