"""ctypes access to the CPU emulation build (test infrastructure only)."""
from __future__ import annotations

import ctypes

import numpy as np

from fastlem_amd import _native

from . import build as _build

_emu = {}


def lib(asan: bool = False):
    if asan not in _emu:
        _emu[asan] = _native._declare(ctypes.CDLL(_build.build(asan)))
    return _emu[asan]


def check(L, rc: int) -> None:
    if rc != 0:
        raise _native.NativeError(rc, L.fl_last_error().decode())


def generate(L, w, init, max_slope=None, max_iteration=0xFFFFFFFF):
    """fl_generate_f64 through the emulated library on workload ``w``."""
    from fastlem_amd.generator import ParameterArrays, make_desc

    params = ParameterArrays(w.base, w.erodibility, w.uplift, w.is_outlet, max_slope)
    keep = []
    d = make_desc(w.model, params, np.ascontiguousarray(init, np.float64),
                  np.ascontiguousarray(w.outlets(), np.uint32), max_iteration, keep)
    ctx = ctypes.c_void_p()
    check(L, L.fl_ctx_create(0, ctypes.byref(ctx)))
    out = np.empty(w.n, np.float64)
    it = ctypes.c_uint32()
    try:
        check(L, L.fl_generate_f64(ctx, ctypes.byref(d), _native.ptr(out, ctypes.c_double), ctypes.byref(it)))
    finally:
        L.fl_ctx_destroy(ctx)
    return out, it.value
