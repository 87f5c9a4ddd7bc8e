"""Natural-neighbour interpolation (Terrain2D::get_elevation, terrain.rs:36-38;
TerrainInterpolator2D, interpolator.rs:11-27; the crate naturalneighbor 1.2.2
is not vendored, so its bits are parity unpinned).

* The brute-force oracle (oracle/sibson.py: Voronoi cells clipped from
  half-planes, no triangulation) is pinned by what Sibson's interpolant must
  satisfy: a site's own value at the site, weights that sum to one (a
  constant field is reproduced), linear precision (a + b x + c y is
  reproduced exactly), non-negative weights, and no value outside the
  sites' convex hull.
* The device kernels (grid-seeded walk, Bowyer-Watson cavity, Watson's
  decomposition) are compared with the oracle through the CPU emulation and
  on the GPU, and checked for the same properties at larger sizes.
"""
import ctypes

import numpy as np
import pytest

from oracle import sibson


def _points(n, seed, box=10.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, box, n), rng.uniform(0.0, box, n)


def _queries(xs, ys, seed, m=40):
    rng = np.random.default_rng(seed)
    qx = list(rng.uniform(xs.min() - 1.0, xs.max() + 1.0, m))
    qy = list(rng.uniform(ys.min() - 1.0, ys.max() + 1.0, m))
    # sites themselves, an edge midpoint and the hull corners' neighbourhood
    for k in (0, 3, 7):
        qx.append(float(xs[k]))
        qy.append(float(ys[k]))
    return np.array(qx), np.array(qy)


# ------------------------------------------------------------- the oracle ---

def test_oracle_sibson_properties():
    xs, ys = _points(40, 1)
    rng = np.random.default_rng(2)
    a, b, c = 0.7, -1.3, 2.1
    lin = a + b * xs + c * ys
    inside = 0
    for _ in range(25):
        q = rng.uniform(1.5, 8.5, 2)
        if not sibson.in_hull(xs, ys, q):
            continue
        inside += 1
        w = sibson.sibson_weights(xs, ys, q)
        assert all(v >= 0.0 for v in w.values())
        assert abs(sum(w.values()) - 1.0) < 1e-12
        pos = sum(v * np.array([xs[i], ys[i]]) for i, v in w.items())
        assert np.allclose(pos, q, atol=1e-10)
        assert abs(sibson.interpolate(xs, ys, lin, q) - (a + b * q[0] + c * q[1])) < 1e-9
    assert inside >= 15
    # at a site: its own value; outside: None
    z = rng.uniform(0, 1, len(xs))
    assert sibson.interpolate(xs, ys, z, (xs[5], ys[5])) == z[5]
    assert sibson.interpolate(xs, ys, z, (xs.max() + 1.0, ys.mean())) is None


def test_oracle_sibson_two_site_symmetry():
    """Four sites on a square: the centre takes a quarter from each."""
    xs = np.array([0.0, 1.0, 1.0, 0.0])
    ys = np.array([0.0, 0.0, 1.0, 1.0])
    w = sibson.sibson_weights(xs, ys, (0.5, 0.5))
    assert sorted(w) == [0, 1, 2, 3]
    assert all(abs(v - 0.25) < 1e-12 for v in w.values())


# ------------------------------------------------------------ the kernels ---

def _interp(L, ctx, xs, ys, tris, z, qx, qy):
    from fastlem_amd import _native

    tri = np.ascontiguousarray(tris, np.uint32).reshape(-1)
    h = ctypes.c_void_p()
    rc = L.fl_interp_create(ctx, len(xs), _native.ptr(xs, ctypes.c_double), _native.ptr(ys, ctypes.c_double),
                            len(tri) // 3, _native.ptr(tri, ctypes.c_uint32), ctypes.byref(h))
    assert rc == 0, L.fl_last_error()
    try:
        out = np.empty(len(qx))
        found = np.empty(len(qx), np.uint8)
        rc = L.fl_interp_query(ctx, h, _native.ptr(np.ascontiguousarray(z, np.float64), ctypes.c_double), len(qx),
                               _native.ptr(np.ascontiguousarray(qx, np.float64), ctypes.c_double),
                               _native.ptr(np.ascontiguousarray(qy, np.float64), ctypes.c_double),
                               _native.ptr(out, ctypes.c_double), _native.ptr(found, ctypes.c_uint8))
        assert rc == 0, L.fl_last_error()
    finally:
        L.fl_interp_destroy(h)
    return out, found.astype(bool)


def check_against_oracle(L, ctx, n=60, seeds=(3, 4)):
    from fastlem_amd.surface import delaunay_triangles

    for seed in seeds:
        xs, ys = _points(n, seed)
        tris, _ = delaunay_triangles(xs, ys)
        z = np.random.default_rng(seed + 10).uniform(-1.0, 3.0, n)
        qx, qy = _queries(xs, ys, seed + 20)
        got, ok = _interp(L, ctx, xs, ys, tris, z, qx, qy)
        for k in range(len(qx)):
            ref = sibson.interpolate(xs, ys, z, (qx[k], qy[k]))
            if ref is None:
                assert not ok[k], (seed, k)
            else:
                assert ok[k], (seed, k)
                assert abs(got[k] - ref) <= 1e-10 * (1.0 + abs(ref)), (seed, k, got[k], ref)


def check_properties(L, ctx, n, seed, m):
    """Linear precision and partition of unity at size: every query inside
    the hull reproduces a + b x + c y and 1; sites give their own value."""
    from fastlem_amd.surface import delaunay_triangles

    xs, ys = _points(n, seed, box=float(np.sqrt(n)))
    tris, hull = delaunay_triangles(xs, ys)
    rng = np.random.default_rng(seed + 1)
    qx = rng.uniform(xs.min(), xs.max(), m)
    qy = rng.uniform(ys.min(), ys.max(), m)
    a, b, c = 2.5, 0.75, -1.25
    lin = a + b * xs + c * ys
    got, ok = _interp(L, ctx, xs, ys, tris, lin, qx, qy)
    want = a + b * qx + c * qy
    assert ok.mean() > 0.9
    assert np.all(np.abs(got[ok] - want[ok]) <= 1e-9 * (1.0 + np.abs(want[ok])))
    one, ok1 = _interp(L, ctx, xs, ys, tris, np.ones(n), qx, qy)
    assert np.array_equal(ok, ok1)
    assert np.all(np.abs(one[ok] - 1.0) <= 1e-12)
    z = rng.uniform(0, 1, n)
    at, okk = _interp(L, ctx, xs, ys, tris, z, xs, ys)
    assert okk.all() and np.array_equal(at, z)
    # outside the hull: no value
    out, oko = _interp(L, ctx, xs, ys, tris, z, np.array([xs.min() - 1.0, xs.max() + 1.0]),
                       np.array([ys.mean(), ys.mean()]))
    assert not oko.any() and np.isnan(out).all()


@pytest.fixture(scope="module")
def emu():
    from cpu_emu import emu as emu_mod

    L = emu_mod.lib()
    ctx = ctypes.c_void_p()
    assert L.fl_ctx_create(0, ctypes.byref(ctx)) == 0
    yield L, ctx
    L.fl_ctx_destroy(ctx)


def test_emulated_interpolation_matches_oracle(emu):
    check_against_oracle(*emu)


def test_emulated_interpolation_properties(emu):
    check_properties(*emu, n=2000, seed=5, m=3000)


@pytest.mark.gpu
def test_interpolation_matches_oracle():
    from fastlem_amd import _native
    from fastlem_amd.generator import context

    check_against_oracle(_native.lib(), context(), n=80, seeds=(3, 4, 5))


@pytest.mark.gpu
def test_interpolation_properties_at_size():
    from fastlem_amd import _native
    from fastlem_amd.generator import context

    check_properties(_native.lib(), context(), n=200_000, seed=6, m=250_000)


@pytest.mark.gpu
def test_terrain_get_elevation(plain10k):
    """The drop-in: generate(), then get_elevation at sites (their own
    elevation), inside the hull (between its natural neighbours' values) and
    outside (None); get_elevations agrees with get_elevation."""
    from fastlem_amd.generator import TerrainGenerator
    from fastlem_amd.surface import Site2D

    w = plain10k
    t = TerrainGenerator().set_model(w.model).set_parameters(w.parameters()).generate()
    h = t.elevations()
    for k in (0, 17, 999):
        assert t.get_elevation(Site2D(float(w.model.xs[k]), float(w.model.ys[k]))) == h[k]
    assert t.get_elevation(Site2D(-5.0, 50.0)) is None
    rng = np.random.default_rng(9)
    qx, qy = rng.uniform(5, 95, 500), rng.uniform(5, 95, 500)
    many = t.get_elevations(qx, qy)
    assert not np.isnan(many).any()
    assert many.min() >= h.min() - 1e-12 and many.max() <= h.max() + 1e-12
    one = np.array([t.get_elevation(Site2D(float(x), float(y))) for x, y in zip(qx[:20], qy[:20])])
    assert np.array_equal(one, many[:20])
