"""``TerrainGenerator`` -- mirror of src/lem/generator.rs:37-213.

``generate()`` validates exactly as the reference (generator.rs:91-116),
resolves outlets (:120-132), draws the elevation jitter with the restated
StdRng (:134-138) and hands the loop (:140-210) to the HIP device path
through the C ABI.  There is no CPU fallback: without the native library or a
GPU, ``generate()`` raises.
"""
from __future__ import annotations

import atexit
import ctypes
import threading
from typing import Optional, Sequence

import numpy as np

from . import _native
from ._native import TerrainDesc, check, ptr
from .parameters import TopographicalParameters

UNBOUNDED = 0xFFFFFFFF
F64_EPSILON = np.finfo(np.float64).eps


class GenerationError(Exception):
    """Mirror of ``GenerationError`` (generator.rs:19-26)."""


class InvalidNumberOfParameters(GenerationError):
    def __init__(self):
        super().__init__("The number of topographical parameters must be equal to the number of sites")


class ParametersNotSet(GenerationError):
    def __init__(self):
        super().__init__("You must set topographical parameters before generating terrain")


class ModelNotSet(GenerationError):
    def __init__(self):
        super().__init__("You must set `TerrainModel` before generating terrain")


def jitter_elevations(base: np.ndarray, seed: int = 0) -> np.ndarray:
    """base + StdRng::seed_from_u64(seed).gen::<f64>() * f64::EPSILON."""
    base = np.ascontiguousarray(base, np.float64)
    out = np.empty_like(base)
    _native.lib().fl_jitter_elevations(seed, len(base), ptr(base, ctypes.c_double), ptr(out, ctypes.c_double))
    return out


class _Context:
    """The calling thread's device context (one HIP stream on device 0).

    Calls on one context must not overlap (its stream and host slots are
    shared by everything the call enqueues), and ctypes releases the GIL
    during foreign calls, so every thread gets a context of its own; their
    work overlaps on the device.  Contexts live until the interpreter exits.
    """

    _local = threading.local()
    _lock = threading.Lock()
    _all: list = []

    @classmethod
    def get(cls):
        h = getattr(cls._local, "ctx", None)
        if h is None:
            h = ctypes.c_void_p()
            check(_native.lib().fl_ctx_create(0, ctypes.byref(h)))
            with cls._lock:
                cls._all.append(h)
            cls._local.ctx = h
        return h

    @classmethod
    def _destroy_all(cls):
        with cls._lock:
            for h in cls._all:
                _native.lib().fl_ctx_destroy(h)
            cls._all.clear()


atexit.register(_Context._destroy_all)


def context():
    return _Context.get()


class ParameterArrays:
    """Struct-of-arrays form of a ``Vec<TopographicalParameters>``."""

    def __init__(self, base, erodibility, uplift, is_outlet, max_slope=None):
        n = len(base)
        self.base = np.ascontiguousarray(base, np.float64)
        self.erodibility = np.ascontiguousarray(erodibility, np.float64)
        self.uplift = np.ascontiguousarray(uplift, np.float64)
        self.is_outlet = np.ascontiguousarray(is_outlet, bool)
        if max_slope is None:
            self.has_max_slope = None
            self.max_slope = None
        else:
            ms = np.ascontiguousarray(max_slope, np.float64)
            self.has_max_slope = np.ascontiguousarray(~np.isnan(ms), np.uint8)
            self.max_slope = np.where(np.isnan(ms), 0.0, ms)
        assert all(len(a) == n for a in (self.erodibility, self.uplift, self.is_outlet))

    def __len__(self):
        return len(self.base)

    @staticmethod
    def from_list(params: Sequence[TopographicalParameters]) -> "ParameterArrays":
        ms = [np.nan if p.max_slope is None else p.max_slope for p in params]
        has = any(p.max_slope is not None for p in params)
        return ParameterArrays([p.base_elevation for p in params], [p.erodibility for p in params],
                               [p.uplift_rate for p in params], [p.is_outlet for p in params],
                               ms if has else None)


def make_desc(model, params: ParameterArrays, init_elevation: np.ndarray, outlets: np.ndarray,
              max_iteration: int, keep: list) -> TerrainDesc:
    """Fill an ``fl_terrain_desc``; ``keep`` holds the arrays alive."""
    row_ptr, col, length = model.graph()
    arrays = [row_ptr, col, length, model.areas(), params.erodibility, params.uplift, init_elevation, outlets,
              params.has_max_slope, params.max_slope]
    keep.extend(arrays)
    d = TerrainDesc()
    d.n = model.num()
    d.nnz = len(col)
    d.row_ptr = ptr(row_ptr, ctypes.c_uint32)
    d.col = ptr(col, ctypes.c_uint32)
    d.length = ptr(length, ctypes.c_double)
    d.area = ptr(model.areas(), ctypes.c_double)
    d.erodibility = ptr(params.erodibility, ctypes.c_double)
    d.uplift_rate = ptr(params.uplift, ctypes.c_double)
    d.has_max_slope = ptr(params.has_max_slope, ctypes.c_uint8)
    d.max_slope = ptr(params.max_slope, ctypes.c_double)
    d.init_elevation = ptr(init_elevation, ctypes.c_double)
    d.outlets = ptr(outlets, ctypes.c_uint32)
    d.n_outlets = len(outlets)
    d.max_iteration = max_iteration
    return d


class TerrainGenerator:
    """Mirror of ``TerrainGenerator`` (generator.rs:37-213)."""

    def __init__(self):
        self.model = None
        self.parameters = None
        self.max_iteration: Optional[int] = None
        self.last_iterations: Optional[int] = None

    def set_model(self, model) -> "TerrainGenerator":
        self.model = model
        return self

    def set_parameters(self, parameters) -> "TerrainGenerator":
        if not isinstance(parameters, ParameterArrays):
            parameters = ParameterArrays.from_list(list(parameters))
        self.parameters = parameters
        return self

    def set_max_iteration(self, max_iteration: int) -> "TerrainGenerator":
        self.max_iteration = int(max_iteration)
        return self

    def resolve_outlets(self) -> np.ndarray:
        """generator.rs:120-132: flagged sites ascending, else the default outlets."""
        o = np.nonzero(self.parameters.is_outlet)[0].astype(np.uint32)
        if len(o) == 0:
            o = np.ascontiguousarray(self.model.default_outlets(), np.uint32)
        return o

    def generate(self):
        if self.model is None:
            raise ModelNotSet()
        num = self.model.num()
        if self.parameters is None:
            raise ParametersNotSet()
        if len(self.parameters) != num:
            raise InvalidNumberOfParameters()
        outlets = self.resolve_outlets()
        init = jitter_elevations(self.parameters.base, 0)
        maxit = UNBOUNDED if self.max_iteration is None else self.max_iteration
        keep: list = []
        desc = make_desc(self.model, self.parameters, init, outlets, maxit, keep)
        out = np.empty(num, np.float64)
        iters = ctypes.c_uint32()
        check(_native.lib().fl_generate_f64(context(), ctypes.byref(desc), ptr(out, ctypes.c_double),
                                            ctypes.byref(iters)))
        self.last_iterations = iters.value
        return self.model.create_terrain_from_result(out)
