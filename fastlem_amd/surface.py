"""2-D surface model: the host-side, one-time setup that feeds the solve.

Mirrors ``fastlem::models::surface`` (src/models/surface/{builder,model,
sites,terrain}.rs): ``Site2D``, ``TerrainModel2DBulider`` (the reference's
spelling), ``TerrainModel2D`` and ``Terrain2D``.  The graph is held as CSR
(``row_ptr``, ``col``, ``length``) whose rows list ``neighbors_of(i)`` in
insertion order, the form the device path consumes.

Delaunay triangulation comes from Qhull (scipy.spatial); cell areas, Lloyd
centroids and the adjacency rule are native host helpers.  The reference
obtains triangles from delaunator, so triangle order -- and with it row order,
the hull edges the ``a<b`` rule drops and the default-outlet order -- differs
from a delaunator-built model; the solve itself takes the CSR as given.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _native
from ._native import ptr


class ModelBuilderError(Exception):
    """Mirror of ``ModelBuilderError`` (builder.rs:13-21)."""


class SitesNotSet(ModelBuilderError):
    pass


class BoundsNotSet(ModelBuilderError):
    pass


class VoronoiError(ModelBuilderError):
    pass


@dataclass(frozen=True)
class Site2D:
    """A point of the plane (sites.rs:7-10)."""

    x: float
    y: float

    def distance(self, other: "Site2D") -> float:
        # sites.rs:27-29: ((dx)^2 + (dy)^2).sqrt()
        dx = self.x - other.x
        dy = self.y - other.y
        return float(np.sqrt(dx * dx + dy * dy))


def _as_xy(sites) -> tuple[np.ndarray, np.ndarray]:
    if isinstance(sites, tuple) and len(sites) == 2 and isinstance(sites[0], np.ndarray):
        return (np.ascontiguousarray(sites[0], np.float64), np.ascontiguousarray(sites[1], np.float64))
    xs = np.array([s.x for s in sites], np.float64)
    ys = np.array([s.y for s in sites], np.float64)
    return xs, ys


def delaunay_triangles(xs: np.ndarray, ys: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Counter-clockwise Delaunay triangles and the convex hull (CCW order)."""
    from scipy.spatial import Delaunay

    pts = np.column_stack([xs, ys])
    try:
        tri = Delaunay(pts)
    except Exception as exc:  # qhull failure (degenerate input)
        raise VoronoiError(str(exc)) from exc
    t = np.ascontiguousarray(tri.simplices.astype(np.int64))
    a, b, c = t[:, 0], t[:, 1], t[:, 2]
    cross = (xs[b] - xs[a]) * (ys[c] - ys[a]) - (ys[b] - ys[a]) * (xs[c] - xs[a])
    cw = cross < 0
    t[cw, 1], t[cw, 2] = c[cw], b[cw].copy()
    hull = _hull_ccw(tri)
    return np.ascontiguousarray(t), hull


def _hull_ccw(tri) -> np.ndarray:
    from scipy.spatial import ConvexHull

    return ConvexHull(tri.points).vertices.astype(np.int64)  # CCW in 2-D


def voronoi_cells(xs, ys, tris, bmin: Site2D, bmax: Site2D):
    """Clipped Voronoi cell areas and centroids (native host helper)."""
    n = len(xs)
    area = np.empty(n, np.float64)
    cx = np.empty(n, np.float64)
    cy = np.empty(n, np.float64)
    L = _native.lib()
    rc = L.fl_voronoi_cells(n, ptr(xs, ctypes.c_double), ptr(ys, ctypes.c_double), len(tris),
                            ptr(tris, ctypes.c_int64), bmin.x, bmin.y, bmax.x, bmax.y,
                            ptr(area, ctypes.c_double), ptr(cx, ctypes.c_double), ptr(cy, ctypes.c_double))
    if rc != 0:
        raise VoronoiError(f"voronoi cells failed ({rc})")
    return area, cx, cy


def build_adjacency(xs, ys, tris) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """CSR adjacency with the reference's ``a<b`` edge rule (builder.rs:249-268)."""
    n = len(xs)
    L = _native.lib()
    row_ptr = np.zeros(n + 1, np.uint32)
    rc = L.fl_build_adjacency(n, ptr(xs, ctypes.c_double), ptr(ys, ctypes.c_double), len(tris),
                              ptr(tris, ctypes.c_int64), ptr(row_ptr, ctypes.c_uint32), None, None)
    if rc != 0:
        raise VoronoiError(f"adjacency sizing failed ({rc})")
    nnz = int(row_ptr[n])
    col = np.empty(nnz, np.uint32)
    length = np.empty(nnz, np.float64)
    rc = L.fl_build_adjacency(n, ptr(xs, ctypes.c_double), ptr(ys, ctypes.c_double), len(tris),
                              ptr(tris, ctypes.c_int64), ptr(row_ptr, ctypes.c_uint32),
                              ptr(col, ctypes.c_uint32), ptr(length, ctypes.c_double))
    if rc != 0:
        raise VoronoiError(f"adjacency build failed ({rc})")
    return row_ptr, col, length


class TerrainModel2D:
    """Mirror of ``TerrainModel2D`` (model.rs:18-69) with a CSR graph.
    ``triangles`` (optional, [t, 3]): the sites' Delaunay triangles, kept
    from the build for the natural-neighbour interpolator of its terrains."""

    def __init__(self, xs, ys, areas, row_ptr, col, length, default_outlets, triangles=None):
        self.triangles = None if triangles is None else np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
        self.xs = np.ascontiguousarray(xs, np.float64)
        self.ys = np.ascontiguousarray(ys, np.float64)
        self._areas = np.ascontiguousarray(areas, np.float64)
        self.row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        self.col = np.ascontiguousarray(col, np.uint32)
        self.length = np.ascontiguousarray(length, np.float64)
        self._default_outlets = np.ascontiguousarray(default_outlets, np.uint32)

    # Model trait (traits.rs:13-20)
    def num(self) -> int:
        return len(self.row_ptr) - 1

    def sites(self) -> list[Site2D]:
        return [Site2D(float(x), float(y)) for x, y in zip(self.xs, self.ys)]

    def areas(self) -> np.ndarray:
        return self._areas

    def default_outlets(self) -> np.ndarray:
        return self._default_outlets

    def graph(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        return self.row_ptr, self.col, self.length

    def neighbors_of(self, i: int) -> list[tuple[int, float]]:
        b, e = int(self.row_ptr[i]), int(self.row_ptr[i + 1])
        return [(int(self.col[k]), float(self.length[k])) for k in range(b, e)]

    def create_terrain_from_result(self, elevations) -> "Terrain2D":
        """model.rs:62-68.  The reference builds the interpolator here, inside
        generate(); this one builds it on the first query (the same
        triangulation of the same sites, when the build kept it)."""
        return Terrain2D(self.xs, self.ys, np.asarray(elevations, np.float64), self.triangles)


class TerrainInterpolator2D:
    """Mirror of ``TerrainInterpolator2D`` (interpolator.rs:6-28):
    natural-neighbour (Sibson) interpolation over the sites, on the GPU (C ABI
    ``fl_interp_*``).  Its triangulation is the given one or, as the
    reference's ``Interpolator::new(sites)`` does, a Delaunay triangulation of
    the sites."""

    def __init__(self, xs, ys, triangles=None):
        from .generator import context

        self.xs = np.ascontiguousarray(xs, np.float64)
        self.ys = np.ascontiguousarray(ys, np.float64)
        if triangles is None:
            triangles, _ = delaunay_triangles(self.xs, self.ys)
        tri = np.ascontiguousarray(triangles, np.uint32).reshape(-1)
        h = ctypes.c_void_p()
        _native.check(_native.lib().fl_interp_create(context(), len(self.xs), ptr(self.xs, ctypes.c_double),
                                                      ptr(self.ys, ctypes.c_double), len(tri) // 3,
                                                      ptr(tri, ctypes.c_uint32), ctypes.byref(h)))
        self._h = h

    def interpolate_many(self, elevations, qx, qy) -> tuple[np.ndarray, np.ndarray]:
        """Values at the points (qx[k], qy[k]) and whether each has one
        (False outside the sites' convex hull, where the value is NaN)."""
        from .generator import context

        z = np.ascontiguousarray(elevations, np.float64)
        qx = np.ascontiguousarray(qx, np.float64).reshape(-1)
        qy = np.ascontiguousarray(qy, np.float64).reshape(-1)
        if len(qx) != len(qy) or len(z) != len(self.xs):
            raise ValueError("query coordinates or elevations of the wrong length")
        out = np.empty(len(qx), np.float64)
        found = np.empty(len(qx), np.uint8)
        _native.check(_native.lib().fl_interp_query(context(), self._h, ptr(z, ctypes.c_double), len(qx),
                                                     ptr(qx, ctypes.c_double), ptr(qy, ctypes.c_double),
                                                     ptr(out, ctypes.c_double), ptr(found, ctypes.c_uint8)))
        return out, found.astype(bool)

    def interpolate(self, elevations, site: Site2D) -> Optional[float]:
        """interpolator.rs:17-27: None outside the convex hull."""
        v, ok = self.interpolate_many(elevations, [site.x], [site.y])
        return float(v[0]) if ok[0] else None

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().fl_interp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Terrain2D:
    """Mirror of ``Terrain2D`` (terrain.rs:8-39): sites, elevations and the
    interpolated elevation anywhere in the sites' convex hull."""

    def __init__(self, xs, ys, elevations, triangles=None):
        self.xs = xs
        self.ys = ys
        self._elevations = elevations
        self._triangles = triangles
        self._interp: Optional[TerrainInterpolator2D] = None

    def sites(self) -> list[Site2D]:
        return [Site2D(float(x), float(y)) for x, y in zip(self.xs, self.ys)]

    def elevations(self) -> np.ndarray:
        return self._elevations

    def _interpolator(self) -> TerrainInterpolator2D:
        if self._interp is None:
            self._interp = TerrainInterpolator2D(self.xs, self.ys, self._triangles)
        return self._interp

    def get_elevation(self, site: Site2D) -> Optional[float]:
        """terrain.rs:36-38: the natural-neighbour interpolation of the
        elevations at `site`, None outside the sites' convex hull."""
        return self._interpolator().interpolate(self._elevations, site)

    def get_elevations(self, xs, ys) -> np.ndarray:
        """get_elevation for many points at once (one GPU launch): NaN where
        get_elevation returns None."""
        v, _ = self._interpolator().interpolate_many(self._elevations, xs, ys)
        return v


class TerrainModel2DBulider:
    """Mirror of ``TerrainModel2DBulider`` (builder.rs:31-322)."""

    def __init__(self):
        self._xs: Optional[np.ndarray] = None
        self._ys: Optional[np.ndarray] = None
        self._bmin: Optional[Site2D] = None
        self._bmax: Optional[Site2D] = None

    @staticmethod
    def from_random_sites(num: int, bound_min: Site2D, bound_max: Site2D) -> "TerrainModel2DBulider":
        """builder.rs:38-52 (StdRng::from_seed([0;32]), gen_range x then y)."""
        b = TerrainModel2DBulider()
        xs = np.empty(num, np.float64)
        ys = np.empty(num, np.float64)
        _native.lib().fl_random_sites(num, bound_min.x, bound_min.y, bound_max.x, bound_max.y,
                                      ptr(xs, ctypes.c_double), ptr(ys, ctypes.c_double))
        b._xs, b._ys = xs, ys
        b._bmin, b._bmax = bound_min, bound_max
        return b

    def set_sites(self, sites) -> "TerrainModel2DBulider":
        self._xs, self._ys = _as_xy(sites)
        return self

    def set_bounding_box(self, bound_min: Optional[Site2D], bound_max: Optional[Site2D]):
        self._bmin, self._bmax = bound_min, bound_max
        return self

    def add_edge_sites(self, edge_num_x: Optional[int] = None, edge_num_y: Optional[int] = None):
        """builder.rs:54-131: evenly spaced sites along the four box edges."""
        if self._xs is None:
            raise SitesNotSet()
        if self._bmin is None or self._bmax is None:
            raise BoundsNotSet()
        bmin, bmax = self._bmin, self._bmax
        n = len(self._xs)
        corners = [(bmin.x, bmin.y), (bmin.x, bmax.y), (bmax.x, bmax.y), (bmax.x, bmin.y)]
        ex, ey = [], []
        for i, (cx, cy) in enumerate(corners):
            nx, ny = corners[(i + 1) % 4]
            if i % 2 == 1:
                cnt = edge_num_x if edge_num_x is not None else int(
                    np.sqrt(n / (bmax.y - bmin.y) * (bmax.x - bmin.x)))
            else:
                cnt = edge_num_y if edge_num_y is not None else int(
                    np.sqrt(n / (bmax.x - bmin.x) * (bmax.y - bmin.y)))
            for j in range(cnt):
                t = j / cnt
                ex.append(cx * (1.0 - t) + nx * t)
                ey.append(cy * (1.0 - t) + ny * t)
        self._xs = np.concatenate([self._xs, np.array(ex, np.float64)])
        self._ys = np.concatenate([self._ys, np.array(ey, np.float64)])
        return self

    def _bounds(self) -> tuple[Site2D, Site2D]:
        if self._xs is None:
            raise SitesNotSet()
        bmin = self._bmin if self._bmin is not None else Site2D(float(self._xs.min()), float(self._ys.min()))
        bmax = self._bmax if self._bmax is not None else Site2D(float(self._xs.max()), float(self._ys.max()))
        return bmin, bmax

    def relaxate_sites(self, times: int) -> "TerrainModel2DBulider":
        """builder.rs:155-200: ``times`` Lloyd steps over box-clipped cells."""
        if times == 0:
            return self
        bmin, bmax = self._bounds()
        xs, ys = self._xs, self._ys
        for _ in range(times):
            tris, _ = delaunay_triangles(xs, ys)
            _, cx, cy = voronoi_cells(xs, ys, tris, bmin, bmax)
            xs, ys = cx, cy
        self._xs, self._ys = xs, ys
        return self

    def build(self) -> TerrainModel2D:
        """builder.rs:202-281: areas, a<b adjacency, hull as default outlets."""
        bmin, bmax = self._bounds()
        xs, ys = self._xs, self._ys
        tris, hull = delaunay_triangles(xs, ys)
        areas, _, _ = voronoi_cells(xs, ys, tris, bmin, bmax)
        row_ptr, col, length = build_adjacency(xs, ys, tris)
        return TerrainModel2D(xs, ys, areas, row_ptr, col, length, hull.astype(np.uint32), tris)
