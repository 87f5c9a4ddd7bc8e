"""Host-side product checks that need no GPU: the C ABI library loads and
exports every declared symbol, the one-time setup helpers (StdRng restatement,
Voronoi cells, adjacency rule, lake visit order) behave as specified, and the
Python mirror of TerrainGenerator raises the reference's errors."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fastlem_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(fl_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from fastlem_amd import _native

    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    from fastlem_amd import _native

    L = _native.lib()
    for n in declared_functions():
        assert getattr(L, n).argtypes is not None or n in ("fl_version",), n


def test_context_without_gpu_is_an_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from fastlem_amd import _native

    h = ctypes.c_void_p()
    rc = _native.lib().fl_ctx_create(0, ctypes.byref(h))
    assert rc == -6  # FL_ERR_NO_DEVICE
    assert b"no HIP device" in _native.lib().fl_last_error()


# ------------------------------------------------------------ StdRng ------

def test_chacha_block_rfc8439():
    """RFC 8439 section 2.3.2 block-function test vector (20 rounds)."""
    from fastlem_amd import _native

    st = np.zeros(16, np.uint32)
    st[:4] = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    st[4:12] = np.frombuffer(bytes(range(32)), "<u4")
    st[12] = 1
    st[13:16] = np.frombuffer(bytes.fromhex("000000090000004a00000000"), "<u4")
    out = np.zeros(16, np.uint32)
    _native.lib().fl_chacha_block(20, _native.ptr(st, ctypes.c_uint32), _native.ptr(out, ctypes.c_uint32))
    assert out.astype("<u4").tobytes().hex() == (
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")


def test_chacha20_zero_key_keystream():
    """Well-known first keystream block of ChaCha20 with zero key and nonce."""
    from fastlem_amd import _native

    st = np.zeros(16, np.uint32)
    st[:4] = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    out = np.zeros(16, np.uint32)
    _native.lib().fl_chacha_block(20, _native.ptr(st, ctypes.c_uint32), _native.ptr(out, ctypes.c_uint32))
    assert out.astype("<u4").tobytes().hex().startswith("76b8e0ada0f13d90405d6ae55386bd28")


def test_jitter_range_and_determinism():
    from fastlem_amd.generator import jitter_elevations

    base = np.linspace(0, 5, 1000)
    a = jitter_elevations(base, 0)
    b = jitter_elevations(base, 0)
    assert np.array_equal(a, b)
    d = a - base
    eps = np.finfo(np.float64).eps
    assert np.all(d >= 0) and np.all(d < eps * 1.0000001)
    # with base 0 the value is an exact multiple of 2^-53 * eps
    z = jitter_elevations(np.zeros(64), 0)
    k = z / (eps * 2.0 ** -53)
    assert np.array_equal(k, np.round(k))
    assert not np.array_equal(jitter_elevations(np.zeros(64), 1), z)


def test_random_sites_in_bounds():
    from fastlem_amd.surface import Site2D, TerrainModel2DBulider

    b = TerrainModel2DBulider.from_random_sites(500, Site2D(10.0, -5.0), Site2D(20.0, 5.0))
    assert np.all((b._xs >= 10) & (b._xs < 20) & (b._ys >= -5) & (b._ys < 5))
    b2 = TerrainModel2DBulider.from_random_sites(500, Site2D(10.0, -5.0), Site2D(20.0, 5.0))
    assert np.array_equal(b._xs, b2._xs)


# -------------------------------------------------------- model building ---

@pytest.fixture(scope="module")
def small_model():
    from fastlem_amd.surface import Site2D, TerrainModel2DBulider

    return TerrainModel2DBulider.from_random_sites(2000, Site2D(0, 0), Site2D(40, 20)).relaxate_sites(1).build()


def test_voronoi_areas_tile_the_box(small_model):
    assert np.all(small_model.areas() > 0)
    assert small_model.areas().sum() == pytest.approx(40 * 20, rel=1e-9)


def test_adjacency_rule(small_model):
    """Each interior Delaunay edge appears in two triangles with opposite
    orientation, so the a<b rule (builder.rs:257-265) adds it exactly once;
    rows are symmetric, duplicate-free and carry the Euclidean length."""
    m = small_model
    rp, col, ln = m.graph()
    n = m.num()
    pairs = set()
    for i in range(n):
        row = col[rp[i]:rp[i + 1]].tolist()
        assert len(row) == len(set(row)) and i not in row
        for k in range(rp[i], rp[i + 1]):
            j = int(col[k])
            pairs.add((i, j))
            assert ln[k] == np.sqrt((m.xs[i] - m.xs[j]) ** 2 + (m.ys[i] - m.ys[j]) ** 2)
    assert all((j, i) in pairs for i, j in pairs)
    from fastlem_amd.surface import delaunay_triangles

    tris, hull = delaunay_triangles(m.xs, m.ys)
    all_edges = set()
    for a, b, c in tris.tolist():
        for u, v in ((a, b), (b, c), (c, a)):
            all_edges.add((min(u, v), max(u, v)))
    # every kept edge is a Delaunay edge; only hull edges may be missing
    assert {(min(i, j), max(i, j)) for i, j in pairs} <= all_edges
    assert len(all_edges) - len(pairs) // 2 <= len(hull)


def test_product_visit_order_matches_oracle(small_model):
    from oracle import flo
    from fastlem_amd import _native

    m = small_model
    rp, col, ln = m.graph()
    rng = np.random.default_rng(5)
    for k in (1, 7, 200):
        outlets = rng.choice(m.num(), k, replace=False).astype(np.uint32)
        ref = flo.visit_order(flo.Graph(rp, col, ln), outlets)
        rank = np.empty(m.num(), np.uint32)
        cnt = ctypes.c_uint64()
        rc = _native.lib().fl_lake_visit_order(m.num(), _native.ptr(rp, ctypes.c_uint32),
                                               _native.ptr(col, ctypes.c_uint32), _native.ptr(ln, ctypes.c_double),
                                               _native.ptr(outlets, ctypes.c_uint32), len(outlets),
                                               _native.ptr(rank, ctypes.c_uint32), ctypes.byref(cnt))
        assert rc == 0 and cnt.value == len(ref)
        assert np.array_equal(rank[ref], np.arange(len(ref)))


def test_product_visit_order_with_ties_matches_oracle():
    """A regular grid: many equal edge lengths, so the heap's tie order matters."""
    from oracle import flo
    from fastlem_amd import _native
    from test_oracle_pinned import random_delaunay_graph

    g, _ = random_delaunay_graph(400, 1, grid=True)
    outlets = np.array([0, 37, 399, 123], np.uint32)
    ref = flo.visit_order(g, outlets)
    rank = np.empty(g.n, np.uint32)
    cnt = ctypes.c_uint64()
    rc = _native.lib().fl_lake_visit_order(g.n, _native.ptr(g.row_ptr, ctypes.c_uint32),
                                           _native.ptr(g.col, ctypes.c_uint32),
                                           _native.ptr(g.length, ctypes.c_double),
                                           _native.ptr(outlets, ctypes.c_uint32), len(outlets),
                                           _native.ptr(rank, ctypes.c_uint32), ctypes.byref(cnt))
    assert rc == 0
    assert np.array_equal(rank[ref], np.arange(len(ref)))


# ------------------------------------------------------- Python mirror ----

def test_generator_errors_match_reference():
    from fastlem_amd import (InvalidNumberOfParameters, ModelNotSet, ParametersNotSet, TerrainGenerator,
                             TopographicalParameters)
    from fastlem_amd.workloads import config_plain

    with pytest.raises(ModelNotSet):
        TerrainGenerator().generate()
    w = config_plain(300, 20.0, 1.0)
    with pytest.raises(ParametersNotSet):
        TerrainGenerator().set_model(w.model).generate()
    with pytest.raises(InvalidNumberOfParameters):
        TerrainGenerator().set_model(w.model).set_parameters([TopographicalParameters()] * 3).generate()


def test_parameter_defaults():
    from fastlem_amd import TopographicalParameters

    p = TopographicalParameters()
    assert (p.base_elevation, p.erodibility, p.uplift_rate, p.is_outlet, p.max_slope) == (0.0, 1.0, 1.0, False, None)
    q = p.set_erodibility(2.0).set_is_outlet(True).set_max_slope(0.3)
    assert (q.erodibility, q.is_outlet, q.max_slope) == (2.0, True, 0.3) and p.erodibility == 1.0


def test_outlet_resolution_falls_back_to_hull():
    from fastlem_amd import ParameterArrays, TerrainGenerator
    from fastlem_amd.workloads import config_plain

    w = config_plain(300, 20.0, 1.0)
    n = w.n
    g = TerrainGenerator().set_model(w.model).set_parameters(
        ParameterArrays(np.zeros(n), np.ones(n), np.ones(n), np.zeros(n, bool)))
    assert np.array_equal(g.resolve_outlets(), w.model.default_outlets())
    flags = np.zeros(n, bool)
    flags[[5, 2, 9]] = True
    g.set_parameters(ParameterArrays(np.zeros(n), np.ones(n), np.ones(n), flags))
    assert g.resolve_outlets().tolist() == [2, 5, 9]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


# ------------------------------------------- visit order, fast form ------

def _outlet_order(g, outlets):
    from fastlem_amd import _native

    out = np.empty(len(outlets), np.uint32)
    rc = _native.lib().fl_outlet_pop_order(g.n, _native.ptr(g.row_ptr, ctypes.c_uint32),
                                           _native.ptr(g.col, ctypes.c_uint32),
                                           _native.ptr(np.ascontiguousarray(outlets, np.uint32), ctypes.c_uint32),
                                           len(outlets), _native.ptr(out, ctypes.c_uint32))
    assert rc == 0
    return out


@pytest.mark.parametrize("seed,k,grid", [(0, 1, False), (1, 5, False), (2, 40, False), (3, 150, False),
                                         (4, 60, True), (5, 399, False), (6, 400, False), (7, 380, True),
                                         (8, 250, False), (9, 320, True)])
def test_outlet_phase_matches_heap_simulation(seed, k, grid):
    """The first k pops of the full Rust-heap simulation are the outlets; the
    shape-only simulation must give the same order (outlets adjacent to each
    other, many pushes, grids with equal lengths included)."""
    from oracle import flo
    from test_oracle_pinned import random_delaunay_graph

    g, _ = random_delaunay_graph(400, 300 + seed, grid=grid)
    rng = np.random.default_rng(seed)
    outlets = rng.choice(g.n, k, replace=False).astype(np.uint32)
    ref = flo.visit_order(g, outlets)[:k]
    assert np.array_equal(_outlet_order(g, outlets), ref)


@pytest.mark.parametrize("seed", range(6))
def test_outlet_phase_with_isolated_outlets(seed):
    """Outlets that push nothing (isolated sites) shrink the heap until its
    end holds outlets again: the order must follow the heap simulation
    through every switch between the walked and the simulated pops."""
    from oracle import flo
    from test_oracle_pinned import random_delaunay_graph

    g0, _ = random_delaunay_graph(120, 900 + seed)
    iso = 60 + 40 * seed
    n = g0.n + iso
    rp = np.concatenate([g0.row_ptr.astype(np.int64), np.full(iso, g0.row_ptr[-1], np.int64)])
    g = flo.Graph(rp, g0.col, g0.length)
    rng = np.random.default_rng(seed)
    outlets = np.concatenate([np.arange(g0.n, n), rng.choice(g0.n, 3 + 5 * seed, replace=False)])
    outlets = rng.permutation(outlets).astype(np.uint32)
    k = len(outlets)
    ref = flo.visit_order(g, outlets)[:k]
    assert np.array_equal(_outlet_order(g, outlets), ref)


def _contracted_msf(g, outlets):
    """Minimum spanning forest of the graph with the outlets merged (scipy)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree

    r0 = int(outlets[0])
    is_out = np.zeros(g.n, bool)
    is_out[outlets] = True
    owner = np.repeat(np.arange(g.n), np.diff(g.row_ptr.astype(np.int64)))
    a = np.where(is_out[owner], r0, owner)
    b = np.where(is_out[g.col], r0, g.col.astype(np.int64))
    keep = a < b
    best = {}
    for x, y, w in zip(a[keep], b[keep], g.length[keep]):
        if (x, y) not in best or w < best[(x, y)]:
            best[(x, y)] = w
    xs = np.array([k[0] for k in best]), np.array([k[1] for k in best])
    mst = minimum_spanning_tree(coo_matrix((list(best.values()), xs), shape=(g.n, g.n))).tocoo()
    order = np.argsort(mst.data, kind="stable")
    return mst.row[order].astype(np.uint32), mst.col[order].astype(np.uint32)


@pytest.mark.parametrize("seed,k", [(0, 1), (1, 4), (2, 30), (3, 120)])
def test_ranks_from_msf_match_heap_simulation(seed, k):
    from oracle import flo
    from fastlem_amd import _native
    from test_oracle_pinned import random_delaunay_graph

    g, _ = random_delaunay_graph(500, 400 + seed)
    rng = np.random.default_rng(seed)
    outlets = rng.choice(g.n, k, replace=False).astype(np.uint32)
    ref = flo.visit_order(g, outlets)
    ma, mb = _contracted_msf(g, outlets)
    first = _outlet_order(g, outlets)
    rank = np.empty(g.n, np.uint32)
    rc = _native.lib().fl_ranks_from_msf(g.n, _native.ptr(first, ctypes.c_uint32), k,
                                         _native.ptr(ma, ctypes.c_uint32), _native.ptr(mb, ctypes.c_uint32),
                                         len(ma), _native.ptr(rank, ctypes.c_uint32))
    assert rc == 0
    assert np.array_equal(rank[ref], np.arange(len(ref)))
    assert len(ref) == g.n


def test_server_budget_and_errors_without_gpu():
    """The request loop's admission control needs no GPU: a request larger
    than the budget is refused at submission; without a device the admitted
    request completes with FL_ERR_NO_DEVICE instead of hanging."""
    import torch

    from fastlem_amd import _native
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.serving import Server, request_bytes
    from fastlem_amd.workloads import config_plain

    w = config_plain(500, 22.0, 1.0)
    need = request_bytes(DeviceBatch.terrain_of(w))
    assert need > 500 * 300  # device arrays (~300 B per site) and the visit order's scratch
    small = Server(slots=2, hbm_budget_bytes=need - 1)  # just too small for this terrain
    with pytest.raises(_native.NativeError) as e:
        small.submit(DeviceBatch.terrain_of(w))
    assert e.value.code == -7  # FL_ERR_BUDGET
    small.close()
    if torch.cuda.is_available():
        return
    srv = Server(slots=2, hbm_budget_bytes=1 << 30)
    t = srv.submit(DeviceBatch.terrain_of(w))
    with pytest.raises(_native.NativeError) as e:
        srv.wait(t)
    assert e.value.code == -6  # FL_ERR_NO_DEVICE
    srv.close()
