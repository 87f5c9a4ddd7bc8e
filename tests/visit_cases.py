"""Graphs for the visit-order checks (lake traversal, stream_tree.rs:184-240),
shared by the CPU-emulation and GPU tests: the library's ranks
(fl_stage_visit_order) against the oracle's heap traversal on

* random Delaunay graphs with distinct lengths and 1..60 outlets (the
  spanning-forest form);
* an exact grid (equal lengths: the heap simulation);
* a graph whose two entries of one edge carry different lengths (legal C-ABI
  input; the reference keys each push by the popped site's own entry);
* two disjoint graphs with outlets in one only (the other is never reached);
* all of them in one batch, each ranked as if alone.
"""
from __future__ import annotations

import ctypes

import numpy as np

from oracle import flo
from test_oracle_pinned import random_delaunay_graph

NONE = 0xFFFFFFFF


def _terrain(g, outlets):
    n = g.n
    return dict(row_ptr=g.row_ptr, col=g.col, length=g.length, area=np.ones(n), erodibility=np.ones(n),
                uplift=np.ones(n), init_elevation=np.zeros(n), outlets=np.ascontiguousarray(outlets, np.uint32))


def _disjoint(g1, g2):
    n1 = g1.n
    rp = np.concatenate([g1.row_ptr[:-1], g2.row_ptr + g1.row_ptr[-1]]).astype(np.uint32)
    col = np.concatenate([g1.col, g2.col + n1]).astype(np.uint32)
    return flo.Graph(rp, col, np.concatenate([g1.length, g2.length]))


def cases():
    out = []
    for seed, n, k in [(0, 300, 1), (1, 500, 5), (2, 700, 40), (3, 900, 60)]:
        g, _ = random_delaunay_graph(n, 900 + seed)
        rng = np.random.default_rng(seed)
        out.append(("delaunay-%d-%d" % (n, k), _terrain(g, rng.choice(g.n, k, replace=False))))
    g, _ = random_delaunay_graph(400, 7, grid=True)
    out.append(("grid", _terrain(g, np.array([0, 19, 380], np.uint32))))
    g, _ = random_delaunay_graph(400, 8)
    ln = g.length.copy()
    # one edge's second entry differs in its last bit
    i = 5
    j = int(g.col[g.row_ptr[i]])
    for q in range(g.row_ptr[j], g.row_ptr[j + 1]):
        if g.col[q] == i:
            ln[q] = np.nextafter(ln[q], np.inf)
    out.append(("asymmetric", _terrain(flo.Graph(g.row_ptr, g.col, ln), np.array([3, 150], np.uint32))))
    g1, _ = random_delaunay_graph(300, 9)
    g2, _ = random_delaunay_graph(200, 10)
    out.append(("unreached-part", _terrain(_disjoint(g1, g2), np.array([4, 77], np.uint32))))
    return out


def expected_ranks(tr):
    g = flo.Graph(tr["row_ptr"], tr["col"], tr["length"])
    order = flo.visit_order(g, tr["outlets"])
    rank = np.full(g.n, NONE, np.uint32)
    rank[order] = np.arange(len(order), dtype=np.uint32)
    return rank


def stage_ranks(L, ctx, terrains):
    """fl_stage_visit_order of the terrains as one batch (one terrain: a
    terrain object) through library L; the ranks per terrain and the
    solve statistics' order_from_forest flag."""
    from fastlem_amd import _native
    from fastlem_amd.device import batch_descs

    keep: list = []
    descs = batch_descs(terrains, keep)
    h = ctypes.c_void_p()
    if len(terrains) == 1:
        rc = L.fl_terrain_create(ctx, ctypes.byref(descs[0]), ctypes.byref(h))
    else:
        rc = L.fl_batch_create(ctx, descs, len(terrains), ctypes.byref(h))
    assert rc == 0, L.fl_last_error()
    try:
        n = sum(len(t["row_ptr"]) - 1 for t in terrains)
        out = np.empty(n, np.uint32)
        rc = L.fl_stage_visit_order(ctx, h, _native.ptr(out, ctypes.c_uint32))
        assert rc == 0, L.fl_last_error()
        st = _native.SolveStats()
        assert L.fl_terrain_last_stats(h, ctypes.byref(st)) == 0
    finally:
        L.fl_terrain_destroy(h)
    res, s0 = [], 0
    for t in terrains:
        m = len(t["row_ptr"]) - 1
        res.append(out[s0:s0 + m])
        s0 += m
    return res, st.order_from_forest


def check_all(L, ctx):
    cs = cases()
    for name, tr in cs:
        (got,), fast = stage_ranks(L, ctx, [tr])
        assert np.array_equal(got, expected_ranks(tr)), name
        assert fast == (0 if name in ("grid", "asymmetric") else 1), (name, fast)
    # one batch: every terrain ranked as alone; a terrain with ties or
    # asymmetric lengths does not stop the others from taking the forest
    terrains = [tr for _, tr in cs]
    got, _ = stage_ranks(L, ctx, terrains)
    for (name, tr), r in zip(cs, got):
        assert np.array_equal(r, expected_ranks(tr)), "batch: " + name
    # the same mesh twice in a batch (equal lengths across terrains only)
    tr = cs[1][1]
    got, fast = stage_ranks(L, ctx, [tr, dict(tr)])
    assert fast == 1
    for r in got:
        assert np.array_equal(r, expected_ranks(tr))
