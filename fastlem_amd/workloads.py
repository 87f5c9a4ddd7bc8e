"""Synthetic workloads for tests and the benchmark.

Each workload is a model (sites, cell areas, CSR graph) plus per-site
parameters, built with this package's host setup:

* ``config_plain``: the reference's own CPU-runnable case -- 10,000 sites of
  ``from_random_sites`` in [0,100]^2, one Lloyd step, sites within 2.0 of the
  boundary are outlets, erodibility 1.0, default uplift / m / iterations
  (the setup of tests/landscape_evolution.rs and examples/landscape_evolution.rs
  with the outlet band of the benchmark definition).
* ``config_fbm``: N seeded uniform sites in [0,W]^2, one Lloyd step,
  erodibility = clip(10^fBm1, 0.1, 10), outlets where a second fBm field is
  below -0.2 ("ocean").  At N = 1,000,000 and W = 1000 this is the
  single-terrain benchmark workload.

fBm here is a sum of octaves of 2-D gradient noise on the integer lattice with
a splitmix64-derived gradient hash; it is this package's own definition (no
reference counterpart) and is fully determined by its seed.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from .surface import Site2D, TerrainModel2D, TerrainModel2DBulider


def _splitmix64(x: np.ndarray) -> np.ndarray:
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def gradient_noise(x: np.ndarray, y: np.ndarray, seed: int) -> np.ndarray:
    """2-D gradient noise in about [-1, 1] with quintic fade."""
    x0 = np.floor(x)
    y0 = np.floor(y)
    fx = x - x0
    fy = y - y0
    ix = x0.astype(np.int64)
    iy = y0.astype(np.int64)

    def grad(cx, cy, dx, dy):
        with np.errstate(over="ignore"):
            key = (cx.astype(np.uint64) * np.uint64(0x8DA6B343)) ^ (cy.astype(np.uint64) * np.uint64(0xD8163841)) \
                ^ np.uint64(seed * 0x2545F4914F6CDD1D & 0xFFFFFFFFFFFFFFFF)
            h = _splitmix64(key)
        ang = (h >> np.uint64(11)).astype(np.float64) * (2.0 * np.pi / 9007199254740992.0)
        return np.cos(ang) * dx + np.sin(ang) * dy

    def fade(t):
        return t * t * t * (t * (t * 6.0 - 15.0) + 10.0)

    n00 = grad(ix, iy, fx, fy)
    n10 = grad(ix + 1, iy, fx - 1.0, fy)
    n01 = grad(ix, iy + 1, fx, fy - 1.0)
    n11 = grad(ix + 1, iy + 1, fx - 1.0, fy - 1.0)
    u = fade(fx)
    v = fade(fy)
    nx0 = n00 + u * (n10 - n00)
    nx1 = n01 + u * (n11 - n01)
    return (nx0 + v * (nx1 - nx0)) * 1.41421356


FBM_GAIN = 1.36  # output scale: standard deviation about 0.24, so P(fbm < -0.2) is about 0.2


def fbm(x: np.ndarray, y: np.ndarray, seed: int, octaves: int = 6, persistence: float = 0.5,
        lacunarity: float = 2.0) -> np.ndarray:
    total = np.zeros_like(x)
    amp = 1.0
    freq = 1.0
    norm = 0.0
    for o in range(octaves):
        total += amp * gradient_noise(x * freq, y * freq, seed * 131 + o)
        norm += amp
        amp *= persistence
        freq *= lacunarity
    return total * (FBM_GAIN / norm)


@dataclass
class Workload:
    name: str
    model: TerrainModel2D
    erodibility: np.ndarray
    uplift: np.ndarray
    is_outlet: np.ndarray
    base: np.ndarray
    max_slope: Optional[np.ndarray] = None  # radians, NaN = None

    @property
    def n(self) -> int:
        return self.model.num()

    def outlets(self) -> np.ndarray:
        """Outlet resolution of generator.rs:120-132."""
        o = np.nonzero(self.is_outlet)[0].astype(np.uint32)
        return o if len(o) else self.model.default_outlets().astype(np.uint32)

    def parameters(self):
        from .parameters import TopographicalParameters

        ps = []
        for i in range(self.n):
            p = TopographicalParameters().set_erodibility(float(self.erodibility[i])) \
                .set_uplift_rate(float(self.uplift[i])).set_is_outlet(bool(self.is_outlet[i])) \
                .set_base_elevation(float(self.base[i]))
            if self.max_slope is not None and not np.isnan(self.max_slope[i]):
                p = p.set_max_slope(float(self.max_slope[i]))
            ps.append(p)
        return ps


def build_uniform_model(n: int, width: float, height: float, seed: int, lloyd: int = 1) -> TerrainModel2D:
    """n seeded uniform sites (numpy PCG64) in [0,width]x[0,height]."""
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0.0, width, n)
    ys = rng.uniform(0.0, height, n)
    b = TerrainModel2DBulider().set_sites((xs, ys)).set_bounding_box(Site2D(0.0, 0.0), Site2D(width, height))
    return b.relaxate_sites(lloyd).build()


def config_plain(n: int = 10000, width: float = 100.0, band: float = 2.0, max_slope: Optional[float] = None
                 ) -> Workload:
    m = TerrainModel2DBulider.from_random_sites(n, Site2D(0.0, 0.0), Site2D(width, width)) \
        .relaxate_sites(1).build()
    xs, ys = m.xs, m.ys
    is_out = (xs < band) | (ys < band) | (xs > width - band) | (ys > width - band)
    k = m.num()
    ms = None if max_slope is None else np.full(k, max_slope)
    return Workload("plain-%d" % n, m, np.ones(k), np.ones(k), is_out, np.zeros(k), ms)


def config_fbm(n: int = 1_000_000, width: float = 1000.0, seed: int = 7, ocean: float = -0.2,
               wavelength: Optional[float] = None) -> Workload:
    m = build_uniform_model(n, width, width, seed)
    lam = wavelength if wavelength is not None else width / 8.0
    x = m.xs / lam
    y = m.ys / lam
    k = np.clip(10.0 ** fbm(x, y, seed * 2 + 1), 0.1, 10.0)
    sea = fbm(x + 17.31, y + 5.77, seed * 2 + 2)
    is_out = sea < ocean
    cnt = m.num()
    return Workload("fbm-%d" % n, m, k, np.ones(cnt), is_out, np.zeros(cnt))


def grid_triangles(nx: int, ny: int, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
    """Counter-clockwise triangles of a jittered nx-by-ny grid (site id =
    iy * nx + ix): each cell quad is split along the diagonal that satisfies
    the empty-circumcircle test, i.e. the Delaunay choice for the quad."""
    ix, iy = np.meshgrid(np.arange(nx - 1, dtype=np.int64), np.arange(ny - 1, dtype=np.int64))
    a = (iy * nx + ix).ravel()      # (x, y)
    b = a + 1                        # (x+1, y)
    c = a + nx + 1                   # (x+1, y+1)
    d = a + nx                       # (x, y+1)
    # is d inside the circumcircle of (a, b, c)?  (a, b, c counter-clockwise)
    ax, ay = xs[a] - xs[d], ys[a] - ys[d]
    bx, by = xs[b] - xs[d], ys[b] - ys[d]
    cx, cy = xs[c] - xs[d], ys[c] - ys[d]
    det = (ax * ax + ay * ay) * (bx * cy - cx * by) - (bx * bx + by * by) * (ax * cy - cx * ay) \
        + (cx * cx + cy * cy) * (ax * by - bx * ay)
    flip = det > 0  # d inside: use diagonal b-d instead of a-c
    t1 = np.where(flip[:, None], np.stack([a, b, d], 1), np.stack([a, b, c], 1))
    t2 = np.where(flip[:, None], np.stack([b, c, d], 1), np.stack([a, c, d], 1))
    return np.ascontiguousarray(np.concatenate([t1, t2]))


def config_strip(nx: int = 16384, ny: int = 1024, jitter: float = 0.4, seed: int = 11) -> Workload:
    """Deep-tree stress: one site per unit cell of an nx-by-ny strip at
    (cx + 0.5 + U(-j, j), cy + 0.5 + U(-j, j)), outlets = the x = 0 column,
    erodibility 1.0 * (1 + U(-0.05, 0.05)), no Lloyd step.  Rivers run the
    whole strip (depth about nx)."""
    from .surface import TerrainModel2D, build_adjacency, voronoi_cells

    rng = np.random.default_rng(seed)
    n = nx * ny
    gx, gy = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64))
    xs = (gx.ravel() + 0.5 + rng.uniform(-jitter, jitter, n))
    ys = (gy.ravel() + 0.5 + rng.uniform(-jitter, jitter, n))
    tris = grid_triangles(nx, ny, xs, ys)
    area, _, _ = voronoi_cells(xs, ys, tris, Site2D(0.0, 0.0), Site2D(float(nx), float(ny)))
    rp, col, ln = build_adjacency(xs, ys, tris)
    del tris
    hull = np.concatenate([np.arange(nx), np.arange(1, ny) * nx + nx - 1, (ny - 1) * nx + np.arange(nx - 2, -1, -1),
                           np.arange(ny - 2, 0, -1) * nx]).astype(np.uint32)
    m = TerrainModel2D(xs, ys, area, rp, col, ln, hull)
    is_out = np.zeros(n, bool)
    is_out[np.arange(ny) * nx] = True
    k = 1.0 * (1.0 + rng.uniform(-0.05, 0.05, n))
    return Workload("strip-%dx%d" % (nx, ny), m, k, np.ones(n), is_out, np.zeros(n))


def _terrain_dict(args):
    """Worker: one terrain of the batched / served streams (no GPU use)."""
    kind, n, seed = args
    import numpy as _np

    from .generator import jitter_elevations
    from .surface import Site2D, TerrainModel2DBulider

    rng = _np.random.default_rng(seed)
    width = float(_np.sqrt(n))
    xs = rng.uniform(0.0, width, n)
    ys = rng.uniform(0.0, width, n)
    m = TerrainModel2DBulider().set_sites((xs, ys)).set_bounding_box(Site2D(0.0, 0.0), Site2D(width, width)).build()
    lam = max(width / 4.0, 1.0)
    k = _np.clip(10.0 ** fbm(m.xs / lam, m.ys / lam, seed * 2 + 1), 0.1, 10.0)
    if kind == "boundary":
        band = 2.0
        out = (m.xs < band) | (m.ys < band) | (m.xs > width - band) | (m.ys > width - band)
    else:  # ocean mask
        out = fbm(m.xs / lam + 17.31, m.ys / lam + 5.77, seed * 2 + 2) < -0.2
        if not out.any():
            out[_np.argmin(m.xs)] = True
    outlets = _np.nonzero(out)[0].astype(_np.uint32)
    rp, col, ln = m.graph()
    cnt = m.num()
    return dict(row_ptr=rp, col=col, length=ln, area=m.areas(), erodibility=k, uplift=_np.ones(cnt),
                outlets=outlets, init_elevation=jitter_elevations(_np.zeros(cnt), 0))


def terrain_stream(count: int, lo: int, hi: int, seed: int, kind: str, workers: int = 8, progress=None) -> list:
    """`count` independent terrains with site counts log-uniform in [lo, hi]
    (seeded): uniform-random sites in [0, sqrt(N)]^2, Delaunay graph,
    erodibility = clip(10^fBm, 0.1, 10), outlets on the boundary band
    (kind="boundary", the batched workload) or where a second fBm < -0.2
    (kind="ocean", the served workload).  Built in a spawn process pool; call
    before the GPU is initialised.  progress(done, count) is called as they
    arrive."""
    import concurrent.futures as cf
    import multiprocessing as mp

    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    jobs = [(kind, int(n), seed * 100003 + i) for i, n in enumerate(sizes)]
    out = []
    if workers <= 1:
        it = map(_terrain_dict, jobs)
        for t in it:
            out.append(t)
            if progress:
                progress(len(out), count)
        return out
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        for t in ex.map(_terrain_dict, jobs, chunksize=1):
            out.append(t)
            if progress:
                progress(len(out), count)
    return out
