"""Brute-force natural-neighbour (Sibson) interpolation -- test infrastructure.

The reference's Terrain2D::get_elevation (src/models/surface/terrain.rs:36-38)
calls TerrainInterpolator2D::interpolate (src/models/surface/interpolator.rs:
17-27), which hands the query to the third-party crate naturalneighbor 1.2.2
(Cargo.toml; not vendored under the reference, so it cannot be read or run
here).  Its published method is Sibson's natural-neighbour interpolation
over the sites: the value at q is sum_i w_i z_i, where w_i is the area that
q's Voronoi cell, were q inserted, takes from site i's cell, divided by the
area of q's cell; a query outside the sites' convex hull has no value
(`None`, interpolator.rs:26 maps the crate's errors to `None` as well).

This module restates that definition directly, without a triangulation: q's
cell is a large box clipped by the bisector half-planes of q against every
site, and the part taken from site i is q's cell clipped further by i's
bisectors against every other site.  O(n^2) half-plane clips per query, for
small cases only.  Only tests import it.
"""
from __future__ import annotations

import numpy as np


def _clip(poly, a, b, c):
    """Part of polygon `poly` (list of (x, y)) with a*x + b*y <= c."""
    out = []
    m = len(poly)
    for k in range(m):
        p, r = poly[k], poly[(k + 1) % m]
        fp = a * p[0] + b * p[1] - c
        fr = a * r[0] + b * r[1] - c
        if fp <= 0:
            out.append(p)
        if (fp < 0 < fr) or (fr < 0 < fp):
            t = fp / (fp - fr)
            out.append((p[0] + t * (r[0] - p[0]), p[1] + t * (r[1] - p[1])))
    return out


def _area(poly):
    if len(poly) < 3:
        return 0.0
    s = 0.0
    for k in range(len(poly)):
        x0, y0 = poly[k]
        x1, y1 = poly[(k + 1) % len(poly)]
        s += x0 * y1 - x1 * y0
    return 0.5 * abs(s)


def _closer_than(poly, p, s):
    """Clip to the half-plane of points at least as close to p as to s."""
    # |x - p|^2 <= |x - s|^2  <=>  2 (s - p) . x <= |s|^2 - |p|^2
    return _clip(poly, 2.0 * (s[0] - p[0]), 2.0 * (s[1] - p[1]), s[0] ** 2 + s[1] ** 2 - p[0] ** 2 - p[1] ** 2)


def in_hull(xs, ys, q) -> bool:
    from scipy.spatial import ConvexHull

    pts = np.column_stack([xs, ys])
    hull = ConvexHull(pts)
    # q inside (or on) every hull facet
    return bool(np.all(hull.equations[:, :2] @ np.asarray(q, float) + hull.equations[:, 2] <= 1e-12))


def sibson_weights(xs, ys, q):
    """{site: weight} of query q (inside the sites' convex hull)."""
    pts = list(zip(map(float, xs), map(float, ys)))
    q = (float(q[0]), float(q[1]))
    span = max(np.ptp(xs), np.ptp(ys)) + 1.0
    cx, cy = float(np.mean(xs)), float(np.mean(ys))
    big = 1e3 * span
    cell = [(cx - big, cy - big), (cx + big, cy - big), (cx + big, cy + big), (cx - big, cy + big)]
    for s in pts:
        cell = _closer_than(cell, q, s)
    total = _area(cell)
    w = {}
    for i, p in enumerate(pts):
        part = cell
        for j, s in enumerate(pts):
            if j != i:
                part = _closer_than(part, p, s)
                if len(part) < 3:
                    break
        a = _area(part)
        if a > 0.0:
            w[i] = a / total
    return w


def interpolate(xs, ys, values, q):
    """Sibson's natural-neighbour value at q, or None outside the hull."""
    for i in range(len(xs)):
        if xs[i] == q[0] and ys[i] == q[1]:
            return float(values[i])
    if not in_hull(xs, ys, q):
        return None
    return float(sum(wi * values[i] for i, wi in sibson_weights(xs, ys, q).items()))
