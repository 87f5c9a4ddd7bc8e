"""Device path vs the CPU oracle: teacher-forced stages and end-to-end runs.

Integer outputs (receivers, chain ends) must match exactly; floating-point
outputs (drainage area, response time, elevations) must match bit for bit,
because the device evaluates every expression in the reference's order.
"""
import numpy as np
import pytest

from conftest import graph_of, init_of

pytestmark = pytest.mark.gpu


def _dev(w, **kw):
    from fastlem_amd.device import DeviceTerrain

    return DeviceTerrain.from_workload(w, **kw)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def test_receivers_teacher_forced(plain10k):
    from oracle import flo

    w = plain10k
    g = graph_of(w)
    t = _dev(w)
    rng = np.random.default_rng(1)
    for elev in (init_of(w), rng.uniform(0, 1, w.n), np.round(rng.uniform(0, 4, w.n))):
        ref = flo.receivers(g, elev, w.outlets())
        got = t.stage_receivers(elev)
        assert np.array_equal(ref, got)


def test_stream_tree_lakes_teacher_forced(plain10k):
    from oracle import flo

    w = plain10k
    g = graph_of(w)
    t = _dev(w)
    elev = init_of(w)
    nref, sref, lake = flo.stream_tree(g, elev, w.outlets())
    assert lake
    ngot, sgot, lgot = t.stage_stream_tree(elev)
    assert lgot == lake
    assert np.array_equal(sref, sgot), "pre-join chain ends differ"
    assert np.array_equal(nref, ngot), f"{int((nref != ngot).sum())} receivers differ after lake joining"


def test_iterations_teacher_forced(plain10k):
    from oracle import flo

    w = plain10k
    g = graph_of(w)
    t = _dev(w)
    elev = init_of(w)
    for it in range(6):
        h_ref, n_ref, _, da_ref, rt_ref, ch_ref, _ = flo.iteration(g, w.model.areas(), w.erodibility, w.uplift,
                                                                   w.outlets(), elev)
        h_got, n_got, da_got, rt_got, ch_got = t.stage_iteration(elev)
        assert np.array_equal(n_ref, n_got), f"iteration {it}: receivers differ"
        assert _bits_equal(da_ref, da_got), f"iteration {it}: drainage area differs"
        assert _bits_equal(rt_ref, rt_got), f"iteration {it}: response time differs"
        assert _bits_equal(h_ref, h_got), f"iteration {it}: elevations differ"
        assert ch_ref == ch_got
        elev = h_ref


def test_generate_plain_end_to_end(plain10k):
    from oracle import flo

    w = plain10k
    g = graph_of(w)
    init = init_of(w)
    ref, iters, _ = flo.generate(g, w.model.areas(), w.erodibility, w.uplift, init, w.outlets())
    t = _dev(w)
    got_iters = t.solve()
    got = t.elevations()
    assert got_iters == iters
    assert _bits_equal(ref, got)


def test_generate_fbm_end_to_end(fbm20k):
    from oracle import flo

    w = fbm20k
    g = graph_of(w)
    init = init_of(w)
    ref, iters, _ = flo.generate(g, w.model.areas(), w.erodibility, w.uplift, init, w.outlets())
    t = _dev(w)
    assert t.solve() == iters
    assert _bits_equal(ref, t.elevations())


def test_generate_max_slope(plain10k):
    """tests/landscape_evolution.rs:29 sets max_slope = 3.14 * 0.1."""
    from oracle import flo
    from fastlem_amd.workloads import Workload

    w0 = plain10k
    ms = np.full(w0.n, 3.14 * 0.1)
    w = Workload(w0.name, w0.model, w0.erodibility, w0.uplift, w0.is_outlet, w0.base, ms)
    g = graph_of(w)
    init = init_of(w)
    ref, iters, _ = flo.generate(g, w.model.areas(), w.erodibility, w.uplift, init, w.outlets(), max_slope=ms)
    t = _dev(w)
    assert t.solve() == iters
    assert _bits_equal(ref, t.elevations())


def test_solve_is_deterministic(fbm20k):
    t = _dev(fbm20k)
    t.solve()
    a = t.elevations()
    t.solve()
    assert _bits_equal(a, t.elevations())


def test_visit_order_from_forest_is_used_and_exact(fbm20k):
    """Random geometry: lengths distinct, the spanning-forest form applies."""
    t = _dev(fbm20k)
    t.solve()
    st = t.last_stats()
    assert st["order_from_forest"] == 1 and st["lake_iterations"] >= 1


def _grid_terrain(side, outlet_band, jitter=0.0, seed=0):
    from test_oracle_pinned import random_delaunay_graph

    g, pts = random_delaunay_graph(side * side, seed, grid=True)
    n = g.n
    xs, ys = pts[:, 0], pts[:, 1]
    out = np.nonzero((xs < outlet_band) | (ys < outlet_band))[0].astype(np.uint32)
    return g, out


def test_generate_grid_with_equal_lengths():
    """A regular grid: many equal edge lengths, so the exact heap simulation
    decides the lake joins (ties beyond the outlet phase)."""
    from oracle import flo
    from fastlem_amd.device import DeviceTerrain
    from fastlem_amd.generator import jitter_elevations

    g, out = _grid_terrain(40, 1.0)
    n = g.n
    init = jitter_elevations(np.zeros(n), 0)
    area = np.ones(n)
    K = np.ones(n)
    ref, iters, _ = flo.generate(g, area, K, K, init, out)
    t = DeviceTerrain(g.row_ptr, g.col, g.length, area, K, K, init, out)
    assert t.solve() == iters
    assert t.last_stats()["order_from_forest"] == 0
    assert _bits_equal(ref, t.elevations())


def test_generate_drop_in_default_outlets():
    """TerrainGenerator.generate() (fl_generate_f64) with no flagged outlet:
    the hull becomes the outlet set (generator.rs:127-129)."""
    from oracle import flo
    from fastlem_amd import ParameterArrays, TerrainGenerator
    from fastlem_amd.generator import jitter_elevations
    from fastlem_amd.workloads import config_plain

    w = config_plain(3000, 55.0, 1.0)
    n = w.n
    gen = TerrainGenerator().set_model(w.model).set_parameters(
        ParameterArrays(np.zeros(n), np.full(n, 2.0), np.full(n, 0.5), np.zeros(n, bool)))
    terrain = gen.generate()
    hull = w.model.default_outlets()
    ref, iters, _ = flo.generate(graph_of(w), w.model.areas(), np.full(n, 2.0), np.full(n, 0.5),
                                 jitter_elevations(np.zeros(n), 0), hull)
    assert gen.last_iterations == iters
    assert _bits_equal(ref, terrain.elevations())


def test_unreachable_component_keeps_elevation():
    from oracle import flo
    from fastlem_amd.device import DeviceTerrain
    from test_oracle_pinned import csr_from_edges, random_delaunay_graph

    g1, _ = random_delaunay_graph(300, 21)
    g2, _ = random_delaunay_graph(200, 22)
    # disjoint union: the second component has no outlet
    edges = []
    for g, off in ((g1, 0), (g2, g1.n)):
        for i in range(g.n):
            for k in range(g.row_ptr[i], g.row_ptr[i + 1]):
                j = int(g.col[k])
                if i < j:
                    edges.append((i + off, j + off, float(g.length[k])))
    g = csr_from_edges(g1.n + g2.n, edges)
    n = g.n
    rng = np.random.default_rng(4)
    init = rng.uniform(0, 1e-9, n)
    out = np.array([0, 7, 19], np.uint32)
    area, K = np.ones(n), np.ones(n)
    ref, iters, _ = flo.generate(g, area, K, K, init, out)
    t = DeviceTerrain(g.row_ptr, g.col, g.length, area, K, K, init, out)
    assert t.solve() == iters
    got = t.elevations()
    assert _bits_equal(ref, got)
    assert np.array_equal(got[g1.n:], init[g1.n:])


def test_max_iteration_is_honoured(plain10k):
    from oracle import flo

    w = plain10k
    init = init_of(w)
    t = _dev(w)
    for cap in (1, 3, 0):
        ref, iters, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init, w.outlets(),
                                     max_iteration=cap)
        assert t.solve(max_iteration=cap) == iters == cap  # re-solving the same terrain
        assert _bits_equal(ref, t.elevations())
    assert _bits_equal(t.elevations(), init)  # cap 0: nothing ran


@pytest.fixture(scope="module")
def strip():
    from fastlem_amd.workloads import config_strip

    return config_strip(512, 64)


def test_deep_strip_end_to_end(strip):
    """Outlets on one short side only: rivers ~900 sites deep, chains that
    span many 64-site blocks of the wavefront kernels."""
    from oracle import flo

    w = strip
    ref, iters, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets())
    t = _dev(w)
    assert t.solve() == iters
    assert _bits_equal(ref, t.elevations())


def test_deep_strip_teacher_forced_with_max_slope(strip):
    from oracle import flo
    from fastlem_amd.workloads import Workload

    w0 = strip
    ms = np.where(np.arange(w0.n) % 3 == 0, 0.05, np.nan)  # a third of the sites clamp
    w = Workload(w0.name, w0.model, w0.erodibility, w0.uplift, w0.is_outlet, w0.base, ms)
    g = graph_of(w)
    t = _dev(w)
    elev = init_of(w)
    for it in range(8):
        h_ref, n_ref, _, da_ref, rt_ref, ch_ref, _ = flo.iteration(g, w.model.areas(), w.erodibility, w.uplift,
                                                                   w.outlets(), elev, ms)
        h_got, n_got, da_got, rt_got, ch_got = t.stage_iteration(elev)
        assert np.array_equal(n_ref, n_got), it
        assert _bits_equal(da_ref, da_got), it
        assert _bits_equal(rt_ref, rt_got), it
        assert _bits_equal(h_ref, h_got), it
        elev = h_ref


def test_batch_matches_solo_runs(plain10k, fbm20k, strip):
    """A mixed batch: every terrain equals its own oracle run bit for bit,
    with its own iteration count (caps and convergence per terrain)."""
    from oracle import flo
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import Workload, config_fbm, config_plain

    small = config_plain(800, 28.0, 1.0)
    ms = np.full(fbm20k.n, 0.2)
    fbm_ms = Workload("fbm-ms", fbm20k.model, fbm20k.erodibility, fbm20k.uplift, fbm20k.is_outlet, fbm20k.base, ms)
    g_grid, out_grid = _grid_terrain(30, 1.0)
    n_grid = g_grid.n
    grid = dict(row_ptr=g_grid.row_ptr, col=g_grid.col, length=g_grid.length, area=np.ones(n_grid),
                erodibility=np.ones(n_grid), uplift=np.ones(n_grid), outlets=out_grid,
                init_elevation=np.random.default_rng(3).uniform(0, 1e-12, n_grid))
    terrains = [DeviceBatch.terrain_of(plain10k), DeviceBatch.terrain_of(small, max_iteration=4),
                DeviceBatch.terrain_of(fbm_ms), grid, DeviceBatch.terrain_of(strip),
                DeviceBatch.terrain_of(config_fbm(5000, 70.0, seed=9)),
                dict(DeviceBatch.terrain_of(small), outlets=np.zeros(0, np.uint32))]
    b = DeviceBatch(terrains)
    its = b.solve()
    for i, tr in enumerate(terrains):
        g = flo.Graph(tr["row_ptr"], tr["col"], tr["length"])
        ref, ref_it, _ = flo.generate(g, tr["area"], tr["erodibility"], tr["uplift"], tr["init_elevation"],
                                      tr["outlets"], max_iteration=tr.get("max_iteration", 0xFFFFFFFF),
                                      max_slope=tr.get("max_slope"))
        assert its[i] == ref_it, (i, its[i], ref_it)
        assert _bits_equal(ref, b.elevations(i)), i


def test_served_results_equal_solo_results(plain10k, fbm20k, strip):
    """Continuous batching: requests of mixed sizes admitted as slots and HBM
    free up; every result is bit-identical to its solo oracle run."""
    from oracle import flo
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.serving import Server
    from fastlem_amd.workloads import config_fbm, config_plain

    terrains = [DeviceBatch.terrain_of(w) for w in
                (plain10k, fbm20k, strip, config_plain(800, 28.0, 1.0), config_fbm(5000, 70.0, seed=9))]
    terrains += [DeviceBatch.terrain_of(config_fbm(3000, 55.0, seed=s)) for s in (21, 22, 23)]
    budget = sum(len(t["row_ptr"]) * 400 + len(t["col"]) * 24 + (1 << 20) for t in terrains[:3])
    srv = Server(slots=4, hbm_budget_bytes=budget)  # not everything fits at once
    tickets = [srv.submit(t) for t in terrains]
    results = [srv.wait(k) for k in tickets]
    assert srv.stats()["peak_running"] >= 2
    srv.close()
    for tr, (h, its, q_ms, lat_ms) in zip(terrains, results):
        g = flo.Graph(tr["row_ptr"], tr["col"], tr["length"])
        ref, ref_it, _ = flo.generate(g, tr["area"], tr["erodibility"], tr["uplift"], tr["init_elevation"],
                                      tr["outlets"])
        assert its == ref_it
        assert _bits_equal(ref, h)
        assert lat_ms >= q_ms >= 0


def test_basin_traversal_order(fbm20k):
    """Tree order: the device's basin traversal equals the reference BFS
    (drainage_basin.rs:13-36) for the largest basins, before and after
    lake joining."""
    from oracle import flo

    w = fbm20k
    g = graph_of(w)
    t = _dev(w)
    elev = init_of(w)
    for it in range(3):
        nxt, _, _ = flo.stream_tree(g, elev, w.outlets())
        roots = np.array([_root(nxt, v) for v in range(w.n)])
        sizes = np.bincount(roots, minlength=w.n)
        for o in np.argsort(-sizes)[:4]:
            ref = flo.basin_order(g, nxt, int(o))
            got = t.stage_basin_order(elev, int(o))
            assert np.array_equal(ref, got), (it, o)
        elev = flo.iteration(g, w.model.areas(), w.erodibility, w.uplift, w.outlets(), elev)[0]


def _root(nxt, v):
    while nxt[v] != v:
        v = nxt[v]
    return v


def test_comb_many_light_children():
    """Every site of a 24-site heavy chain (the long-chain path) folds six
    light children after its heavy child, beyond the four kept per site."""
    from conftest import comb_terrain
    from fastlem_amd.device import DeviceTerrain
    from oracle import flo

    t = comb_terrain()
    d = DeviceTerrain(t["row_ptr"], t["col"], t["length"], t["area"], t["erodibility"], t["uplift"],
                      t["init_elevation"], t["outlets"])
    it = d.solve()
    g = flo.Graph(t["row_ptr"], t["col"], t["length"])
    ref, rit, _ = flo.generate(g, t["area"], t["erodibility"], t["uplift"], t["init_elevation"], t["outlets"])
    assert it == rit
    assert _bits_equal(d.elevations(), ref)


def test_batch_compaction_matches_solo_runs():
    """The union is gathered into smaller ones as terrains stop (forced at
    this size); results and counts equal the solo runs."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm, config_plain
    from oracle import flo

    ws = [config_plain(400, 20.0), config_fbm(2500, 50.0, seed=9), config_plain(900, 30.0),
          config_fbm(1200, 35.0, seed=4), config_plain(3000, 55.0)]
    caps = [0xFFFFFFFF, 0xFFFFFFFF, 5, 0xFFFFFFFF, 0xFFFFFFFF]
    b = DeviceBatch([DeviceBatch.terrain_of(w, max_iteration=c) for w, c in zip(ws, caps)])
    for rep in range(2):  # a second solve starts again from the full union
        its = b.solve(compact_min_idle=1)
        assert b.last_stats()["compactions"] >= 1
        for i, (w, cap) in enumerate(zip(ws, caps)):
            ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w),
                                       w.outlets(), max_iteration=cap)
            assert its[i] == rit, (rep, i)
            assert _bits_equal(b.elevations(i), ref), (rep, i)


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_pipelining_modes_agree(fbm20k, mode):
    """Synchronous iterations (mode 0), minimal speculation with most
    iterations redone (mode 2) and pipelined iterations replayed from
    captured graphs (mode 3) against the oracle; mode 1 (the default) is
    what every other test runs."""
    import ctypes

    from fastlem_amd import _native
    from fastlem_amd.generator import context
    from oracle import flo

    w = fbm20k
    L = _native.lib()
    _native.check(L.fl_ctx_set_pipelining(context(), mode))
    try:
        t = _dev(w)
        it = t.solve()
        st = t.last_stats()
    finally:
        _native.check(L.fl_ctx_set_pipelining(context(), 1))
    ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets())
    assert it == rit
    assert _bits_equal(t.elevations(), ref)
    if mode == 2:
        assert st["redone"] > 0


def test_profiled_solve_is_exact_and_timed(fbm20k):
    """A profiled solve (every launch carries start / end events stamped by
    its dispatch) gives the oracle's elevations, and the profile holds every
    per-iteration kernel with positive time and the launch counts the
    iteration count implies."""
    import ctypes

    from fastlem_amd import _native
    from fastlem_amd.generator import context
    from oracle import flo

    w = fbm20k
    L = _native.lib()
    ctx = context()
    _native.check(L.fl_ctx_set_profiling(ctx, 1))
    try:
        _native.check(L.fl_ctx_reset_profile(ctx))
        t = _dev(w)
        it = t.solve()
        redone = t.last_stats()["redone"]
        prof = {}
        for nm in ("k_receivers", "k_children", "k_chain_perm", "k_da_fold", "k_tau_fold<false>", "k_da_tail",
                   "k_tau_tail<false>", "k_tau_scatter"):
            cnt, ms, units = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64()
            _native.check(L.fl_ctx_kernel_time(ctx, nm.encode(), ctypes.byref(cnt), ctypes.byref(ms),
                                               ctypes.byref(units)))
            prof[nm] = (cnt.value, ms.value, units.value)
    finally:
        _native.check(L.fl_ctx_set_profiling(ctx, 0))
    ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets())
    assert it == rit
    assert _bits_equal(t.elevations(), ref)
    for nm, (cnt, ms, units) in prof.items():
        assert cnt > 0 and ms > 0.0 and units > 0, (nm, cnt, ms, units)
    # one receivers launch per iteration, per redone iteration and per
    # speculative iteration cancelled at convergence
    assert it <= prof["k_receivers"][0] <= it + redone + 3
    assert prof["k_receivers"][2] == prof["k_receivers"][0] * w.n
    assert prof["k_da_fold"][0] >= it
    # the deep rounds of an iteration share one launch per sweep
    assert prof["k_da_tail"][0] == prof["k_tau_tail<false>"][0] <= prof["k_receivers"][0]


def test_visit_order_matches_heap_traversal():
    """The device visit order equals the oracle's heap traversal on random
    Delaunay graphs, an equal-length grid, asymmetric lengths and an
    unreachable component, alone and in batches."""
    import visit_cases
    from fastlem_amd import _native
    from fastlem_amd.generator import context

    visit_cases.check_all(_native.lib(), context())


def test_generate_from_python_threads(plain10k, fbm20k):
    """The drop-in generate() called from several Python threads at once
    (ctypes releases the GIL): each thread runs on its own context and every
    result equals the oracle's bit for bit."""
    import threading

    from oracle import flo
    from fastlem_amd.generator import TerrainGenerator

    ws = [plain10k, fbm20k, plain10k, fbm20k]
    results = [None] * len(ws)

    def run(k):
        w = ws[k]
        g = TerrainGenerator().set_model(w.model).set_parameters(w.parameters())
        results[k] = (g.generate().elevations(), g.last_iterations)

    th = [threading.Thread(target=run, args=(k,)) for k in range(len(ws))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for w, res in zip(ws, results):
        assert res is not None
        h, its = res
        ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets())
        assert its == rit
        assert _bits_equal(ref, h)


def test_tunings_give_identical_results(fbm20k):
    """On the GPU: every launch tuning gives the same elevations bit for bit
    as the table's, and equals the oracle."""
    import ctypes

    from oracle import flo
    from fastlem_amd import _native
    from fastlem_amd.generator import context

    w = fbm20k
    t = _dev(w)
    L = _native.lib()
    ctx = context()
    ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets())
    try:
        for tun in [None, (64, 1, 0, 1, 65536), (1024, 100000, 64, 1, 65536), (128, 3, 8, 0, 65536),
                    (512, 12, 1000, 3, 65536)]:
            if tun is None:
                assert L.fl_ctx_set_tuning(ctx, None) == 0
            else:
                tt = _native.Tuning(*tun)
                assert L.fl_ctx_set_tuning(ctx, ctypes.byref(tt)) == 0
            assert t.solve() == rit, tun
            assert _bits_equal(ref, t.elevations()), tun
    finally:
        L.fl_ctx_set_tuning(ctx, None)
