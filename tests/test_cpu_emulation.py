"""The kernel and host sources compiled for the CPU (tests/cpu_emu: g++
against an emulation of the HIP subset they use) and run through the C ABI
against the oracle.  This checks the kernels' logic where no GPU is present;
it says nothing about speed, and races between concurrently running threads
are not exercised (workgroups run one after another).  The product never
loads this build."""
import ctypes

import numpy as np
import pytest

from conftest import graph_of, init_of


@pytest.fixture(scope="module")
def emu_lib():
    from cpu_emu import emu

    return emu.lib()


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _oracle(w, ms=None, cap=0xFFFFFFFF):
    from oracle import flo

    return flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init_of(w), w.outlets(),
                        max_slope=ms, max_iteration=cap)


@pytest.mark.parametrize("kind", ["plain", "plain-max-slope", "fbm", "strip-max-slope"])
def test_emulated_generate_matches_oracle(emu_lib, kind):
    from cpu_emu import emu
    from fastlem_amd.workloads import config_fbm, config_plain, config_strip

    w = {"plain": lambda: config_plain(2000), "plain-max-slope": lambda: config_plain(2000),
         "fbm": lambda: config_fbm(3000, 55.0, seed=5), "strip-max-slope": lambda: config_strip(48, 24)}[kind]()
    ms = None
    if kind.endswith("max-slope"):
        ms = np.full(w.n, np.nan)
        ms[: w.n // 2] = 0.3
    got, it = emu.generate(emu_lib, w, init_of(w), ms)
    ref, rit, _ = _oracle(w, ms)
    assert it == rit
    assert _bits_equal(got, ref)


def test_emulated_capped_oscillation(emu_lib):
    """A terrain on which the reference alternates between two states
    forever (from iteration 82): only max_iteration ends the loop."""
    from cpu_emu import emu
    from fastlem_amd.workloads import config_plain

    w = config_plain(10000)
    ms = np.full(w.n, np.nan)
    ms[: w.n // 2] = 0.3
    got, it = emu.generate(emu_lib, w, init_of(w), ms, max_iteration=90)
    ref, rit, _ = _oracle(w, ms, cap=90)
    assert it == rit == 90
    assert _bits_equal(got, ref)


def _emu_solve(L, terrs, compact=None, recompute_all=None, rehint=None, pipelining=None):
    """fl_batch_create/solve/download through the emulated library."""
    from cpu_emu import emu
    from fastlem_amd import _native
    from fastlem_amd.device import batch_descs

    keep = []
    descs = batch_descs(terrs, keep)
    ctx = ctypes.c_void_p()
    emu.check(L, L.fl_ctx_create(0, ctypes.byref(ctx)))
    b = ctypes.c_void_p()
    outs = []
    try:
        emu.check(L, L.fl_batch_create(ctx, descs, len(terrs), ctypes.byref(b)))
        if compact is not None:
            emu.check(L, L.fl_ctx_set_batch_compaction(ctx, compact))
        if recompute_all is not None:
            emu.check(L, L.fl_ctx_set_recompute_all(ctx, recompute_all))
        if rehint is not None:
            emu.check(L, L.fl_ctx_set_rehint_rounds(ctx, rehint))
        if pipelining is not None:
            emu.check(L, L.fl_ctx_set_pipelining(ctx, pipelining))
        its = np.zeros(len(terrs), np.uint32)
        emu.check(L, L.fl_batch_solve(ctx, b, _native.ptr(its, ctypes.c_uint32)))
        for i, t in enumerate(terrs):
            out = np.empty(len(t["row_ptr"]) - 1)
            emu.check(L, L.fl_batch_download(ctx, b, i, _native.ptr(out, ctypes.c_double)))
            outs.append(out)
        st = _native.SolveStats()
        emu.check(L, L.fl_terrain_last_stats(b, ctypes.byref(st)))
        L.fl_terrain_destroy(b)
        _emu_solve.last_compactions = st.compactions
        _emu_solve.last_stats = st
    finally:
        L.fl_ctx_destroy(ctx)
    return outs, its


@pytest.mark.parametrize("spine,leaves", [(24, 6), (300, 1), (200, 3)])
def test_emulated_comb_many_light_children(emu_lib, spine, leaves):
    """Every site of a heavy chain (the long-chain path) folds its light
    children after its heavy child: six, beyond the four kept per site, on a
    24-site chain; one or three on chains of several 64-position tiles (the
    tiles loaded ahead of the fold)."""
    from conftest import comb_terrain
    from oracle import flo

    t = comb_terrain(spine=spine, leaves=leaves)
    (got,), (it,) = _emu_solve(emu_lib, [t])
    g = flo.Graph(t["row_ptr"], t["col"], t["length"])
    ref, rit, _ = flo.generate(g, t["area"], t["erodibility"], t["uplift"], t["init_elevation"], t["outlets"])
    assert it == rit
    assert _bits_equal(got, ref)


def test_emulated_create_paths(emu_lib):
    """Terrains whose rows hold at most 64 entries are checked, put in the
    clustered locality order and relabelled on the device; a longer row (a
    comb spine site with 63 leaves and two spine neighbours) sends the upload
    to the host path, for a whole batch too.  Every path gives the oracle's
    bits."""
    from conftest import comb_terrain
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm
    from oracle import flo

    def oracle(t):
        g = flo.Graph(t["row_ptr"], t["col"], t["length"])
        ref, rit, _ = flo.generate(g, t["area"], t["erodibility"], t["uplift"], t["init_elevation"], t["outlets"])
        return ref, rit

    fbm = DeviceBatch.terrain_of(config_fbm(3000, 55.0, seed=4))
    at64, at65 = comb_terrain(spine=6, leaves=62, seed=1), comb_terrain(spine=6, leaves=63, seed=2)
    assert np.diff(at64["row_ptr"]).max() == 64 and np.diff(at65["row_ptr"]).max() == 65
    for terrs, on_device in (([fbm], 1), ([at64], 1), ([at65], 0), ([fbm, at64], 1), ([fbm, at65, at64], 0)):
        outs, its = _emu_solve(emu_lib, terrs)
        st = _emu_solve.last_stats
        assert st.create_on_device == on_device, [len(t["row_ptr"]) - 1 for t in terrs]
        if on_device:
            n = sum(len(t["row_ptr"]) - 1 for t in terrs)
            assert n / 20 <= st.create_clusters <= n / 2, st.create_clusters
        for t, got, it in zip(terrs, outs, its):
            ref, rit = oracle(t)
            assert it == rit
            assert _bits_equal(got, ref)


def test_emulated_dirty_basins_only(emu_lib):
    """By default an iteration recomputes only the basins whose drainage tree
    changed; recomputing every basin (the reference's work) gives the same
    bits and iteration counts, alone and in a batch with caps, and counts
    every reachable site in every iteration."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm, config_plain

    fbm = DeviceBatch.terrain_of(config_fbm(4000, 63.0, seed=3))
    plain = DeviceBatch.terrain_of(config_plain(1500), max_iteration=7)
    for terrs in ([fbm], [fbm, plain]):
        outs_all, its_all = _emu_solve(emu_lib, terrs, recompute_all=1)
        every = _emu_solve.last_stats.recomputed_sites
        outs, its = _emu_solve(emu_lib, terrs, recompute_all=0)
        some = _emu_solve.last_stats.recomputed_sites
        assert list(its) == list(its_all)
        for a, b in zip(outs, outs_all):
            assert _bits_equal(a, b)
        if len(terrs) == 1:
            n = len(fbm["row_ptr"]) - 1
            assert every <= n * its[0] and every > 0.9 * n * its[0], (every, n, its[0])
        assert some < 0.8 * every, (some, every)


def test_emulated_first_iteration_light_depth(emu_lib):
    """A solve's first iteration picks heavy children by subtree size (no
    drainage area exists yet), so every light edge at least halves the
    subtree: its sweep rounds (light depth + 1) are at most floor(log2 n) + 1,
    alone and in a batch with a frozen terrain's stale pointers."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm, config_plain

    fbm = DeviceBatch.terrain_of(config_fbm(4000, 63.0, seed=5))
    plain = DeviceBatch.terrain_of(config_plain(1500), max_iteration=2)
    for terrs in ([fbm], [plain, fbm]):
        n = sum(len(t["row_ptr"]) - 1 for t in terrs)
        _emu_solve(emu_lib, terrs, compact=0)
        st = _emu_solve.last_stats
        assert 2 <= st.first_sweep_rounds <= int(np.log2(n)) + 1, (st.first_sweep_rounds, n)
        assert st.max_sweep_rounds >= st.first_sweep_rounds


@pytest.mark.parametrize("pipelining", [1, 0])
def test_emulated_rehinted_decompositions(emu_lib, pipelining):
    """An iteration whose decomposition is deeper than the bound is decomposed
    again with subtree sizes: with a bound of one round every iteration is,
    in the synchronous and the pipelined form (where a confirmed deep flight
    stops the enqueueing until a synchronous iteration re-hints), and the bits
    and iteration counts stay the oracle's, alone and in a batch with caps."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm, config_plain
    from oracle import flo

    fbm = DeviceBatch.terrain_of(config_fbm(3000, 55.0, seed=8))
    plain = DeviceBatch.terrain_of(config_plain(1200), max_iteration=6)
    for terrs in ([fbm], [plain, fbm]):
        outs, its = _emu_solve(emu_lib, terrs, rehint=1, pipelining=pipelining)
        st = _emu_solve.last_stats
        assert st.rehinted > 0
        for t, got, it in zip(terrs, outs, its):
            g = flo.Graph(t["row_ptr"], t["col"], t["length"])
            ref, rit, _ = flo.generate(g, t["area"], t["erodibility"], t["uplift"], t["init_elevation"],
                                       t["outlets"], max_iteration=t.get("max_iteration", 0xFFFFFFFF))
            assert it == rit
            assert _bits_equal(got, ref)


def test_emulated_batch_matches_solo(emu_lib):
    """The batched union (per-terrain convergence, caps and freeze masks)
    against each terrain's oracle run."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm, config_plain

    ws = [config_plain(1500), config_fbm(2500, 50.0, seed=9), config_plain(800, 60.0)]
    caps = [0xFFFFFFFF, 0xFFFFFFFF, 3]
    terrs = [DeviceBatch.terrain_of(w, max_iteration=c) for w, c in zip(ws, caps)]
    outs, its = _emu_solve(emu_lib, terrs)
    for i, (w, cap) in enumerate(zip(ws, caps)):
        ref, rit, _ = _oracle(w, cap=cap)
        assert its[i] == rit
        assert _bits_equal(outs[i], ref)


def test_emulated_batch_compaction(emu_lib):
    """Terrains that converge at different iterations, with a cap: the union
    is compacted as they stop (forced at this size); every result and count
    still equals its solo run."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_fbm, config_plain

    ws = [config_plain(400, 20.0), config_fbm(2500, 50.0, seed=9), config_plain(900, 30.0),
          config_fbm(1200, 35.0, seed=4), config_plain(3000, 55.0)]
    caps = [0xFFFFFFFF, 0xFFFFFFFF, 5, 0xFFFFFFFF, 0xFFFFFFFF]
    terrs = [DeviceBatch.terrain_of(w, max_iteration=c) for w, c in zip(ws, caps)]
    outs, its = _emu_solve(emu_lib, terrs, compact=1)
    assert _emu_solve.last_compactions >= 1
    for i, (w, cap) in enumerate(zip(ws, caps)):
        ref, rit, _ = _oracle(w, cap=cap)
        assert its[i] == rit, i
        assert _bits_equal(outs[i], ref), i


@pytest.mark.parametrize("kind", ["fbm", "strip"])
def test_emulated_pipelining_modes_agree(emu_lib, kind):
    """The pipelined iteration (the next one enqueued while the current one
    runs, round counts taken from the last confirmed iteration; launched
    directly in mode 1, replayed from a captured graph in mode 3), its redo
    path (mode 2 speculates one round of each kind, so most iterations are
    redone) and the synchronous one (mode 0) give the same bits, counts and
    oracle results."""
    from cpu_emu import emu
    from fastlem_amd import _native
    from fastlem_amd.device import DeviceBatch, batch_descs
    from fastlem_amd.workloads import config_fbm, config_strip

    w = config_fbm(3000, 55.0, seed=5) if kind == "fbm" else config_strip(48, 24)
    L = emu_lib
    runs = {}
    for mode in (0, 1, 2, 3):
        keep = []
        d = batch_descs([DeviceBatch.terrain_of(w)], keep)
        ctx = ctypes.c_void_p()
        emu.check(L, L.fl_ctx_create(0, ctypes.byref(ctx)))
        try:
            emu.check(L, L.fl_ctx_set_pipelining(ctx, mode))
            t = ctypes.c_void_p()
            emu.check(L, L.fl_terrain_create(ctx, ctypes.byref(d[0]), ctypes.byref(t)))
            it = ctypes.c_uint32()
            emu.check(L, L.fl_terrain_solve(ctx, t, 0xFFFFFFFF, ctypes.byref(it)))
            out = np.empty(w.n)
            emu.check(L, L.fl_terrain_download(ctx, t, _native.ptr(out, ctypes.c_double)))
            st = _native.SolveStats()
            emu.check(L, L.fl_terrain_last_stats(t, ctypes.byref(st)))
            L.fl_terrain_destroy(t)
        finally:
            L.fl_ctx_destroy(ctx)
        runs[mode] = (it.value, out, st.host_syncs, st.redone)
    ref, rit, _ = _oracle(w)
    for mode in (0, 1, 2, 3):
        assert runs[mode][0] == rit
        assert _bits_equal(runs[mode][1], ref)
    assert runs[1][2] < runs[0][2]  # fewer host synchronisations
    assert runs[2][3] > 0  # the redo path ran


def _create_error(L, terrains):
    """(status, message) of fl_batch_create on the terrain dicts (emulated
    library: validation and relabelling are host code)."""
    from cpu_emu import emu
    from fastlem_amd.device import batch_descs

    keep = []
    d = batch_descs(terrains, keep)
    ctx = ctypes.c_void_p()
    emu.check(L, L.fl_ctx_create(0, ctypes.byref(ctx)))
    h = ctypes.c_void_p()
    rc = L.fl_batch_create(ctx, d, len(terrains), ctypes.byref(h))
    msg = L.fl_last_error().decode()
    if rc == 0:
        L.fl_terrain_destroy(h)
    L.fl_ctx_destroy(ctx)
    return rc, msg


def _symmetric_edit(t, i, j_old, j_new):
    """Terrain dict t with row i's entry j_old replaced by j_new (one-sided)."""
    t = dict(t)
    col = np.array(t["col"], np.uint32)
    rp = t["row_ptr"]
    k = rp[i] + int(np.nonzero(col[rp[i]:rp[i + 1]] == j_old)[0][0])
    col[k] = j_new
    t["col"] = col
    return t


def test_emulated_validation_reports_the_first_error(emu_lib):
    """Malformed graphs are rejected at upload with FL_ERR_INVALID (-1): the
    message names the first failing row, as a sequential scan would, although
    rows are checked on host threads (the 300k-site terrain takes the
    threaded path); in a batch, the first failing terrain is named."""
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.workloads import config_plain

    L = emu_lib
    big = DeviceBatch.terrain_of(config_plain(300_000, 548.0))
    small = DeviceBatch.terrain_of(config_plain(500, 22.0, 1.0))
    rp, col = big["row_ptr"], big["col"]
    assert _create_error(L, [big]) == (0, "")
    # a self-loop at row 250000 and a duplicate at row 120000: the duplicate's row comes first
    a = _symmetric_edit(big, 250_000, int(col[rp[250_000]]), 250_000)
    r1 = 120_000
    a = _symmetric_edit(a, r1, int(col[rp[r1] + 1]), int(col[rp[r1]]))
    rc, msg = _create_error(L, [a])
    assert rc == -1 and msg == f"duplicate edge {r1}-{int(col[rp[r1]])}", msg
    # an out-of-range index
    b = _symmetric_edit(big, 200_000, int(col[rp[200_000]]), 300_000)
    rc, msg = _create_error(L, [b])
    assert rc == -1 and msg == f"col[{rp[200_000]}]=300000 out of range", msg
    # one-sided edges (rows 180000 and 90000 lose a neighbour and gain a
    # distant site): the first entry in row order whose reverse is missing,
    # found by brute force, is the one reported
    far = 5
    c = _symmetric_edit(big, 180_000, int(col[rp[180_000]]), far)
    c = _symmetric_edit(c, 90_000, int(col[rp[90_000]]), far)
    n = len(rp) - 1
    cc = np.asarray(c["col"], np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    keys = np.sort(rows * n + cc)
    rev = cc * n + rows
    pos = np.minimum(np.searchsorted(keys, rev), len(keys) - 1)
    bad_k = int(np.nonzero(keys[pos] != rev)[0][0])
    rc, msg = _create_error(L, [c])
    assert rc == -1 and msg == f"edge {int(rows[bad_k])}-{int(cc[bad_k])} is not symmetric", msg
    # a batch: the first failing terrain is named
    bad = dict(small)
    bad["outlets"] = np.concatenate([small["outlets"], small["outlets"][:1]]).astype(np.uint32)
    rc, msg = _create_error(L, [small, small, bad, a])
    assert rc == -1 and msg == f"terrain 2: duplicate outlet {int(small['outlets'][0])}", msg


def test_emulated_visit_order_matches_heap_traversal(emu_lib):
    """The device visit order (spanning forest, Euler tours, heavier-ancestor
    jumping, preorder ranking; exact simulation for ties and asymmetric
    lengths) equals the oracle's heap traversal, alone and in batches."""
    import visit_cases

    ctx = ctypes.c_void_p()
    assert emu_lib.fl_ctx_create(0, ctypes.byref(ctx)) == 0
    try:
        visit_cases.check_all(emu_lib, ctx)
    finally:
        emu_lib.fl_ctx_destroy(ctx)


def test_emulated_noalloc_matches_generate(emu_lib):
    """fl_generate_f64_noalloc (static scratch) equals fl_generate_f64, two
    calls in a row with a smaller second terrain; too large is refused."""
    import noalloc_cases

    ctx = ctypes.c_void_p()
    assert emu_lib.fl_ctx_create(0, ctypes.byref(ctx)) == 0
    try:
        noalloc_cases.check_noalloc(emu_lib, ctx, scale=0.05)
    finally:
        emu_lib.fl_ctx_destroy(ctx)


def test_emulated_server_continuous_batching(emu_lib):
    """The serving loop through the CPU emulation: requests of mixed sizes,
    caps (0 and small) and max slopes, submitted with gaps and under a budget
    that holds only part of them, join lanes at iteration boundaries; every
    result and iteration count equals the oracle's solo run."""
    import time

    from oracle import flo
    from fastlem_amd.device import DeviceBatch
    from fastlem_amd.serving import request
    from fastlem_amd.workloads import config_fbm, config_plain, terrain_stream

    from cpu_emu import emu

    terrains = terrain_stream(8, 200, 1500, 77, "ocean", workers=1)
    small = config_plain(300, 17.0, 1.0)
    ms = np.full(small.n, np.nan)
    ms[: small.n // 2] = 0.6
    terrains.append(dict(DeviceBatch.terrain_of(small), max_slope=ms))
    terrains.append(DeviceBatch.terrain_of(config_fbm(900, 30.0, seed=4), max_iteration=3))
    terrains.append(DeviceBatch.terrain_of(config_plain(250, 16.0, 1.0), max_iteration=0))
    L = emu_lib
    need = []
    reqs = []
    for t in terrains:
        desc, keep, out = request(t)
        b = ctypes.c_uint64()
        assert L.fl_request_device_bytes(ctypes.byref(desc), ctypes.byref(b)) == 0
        need.append(b.value)
        reqs.append((desc, keep, out))
    budget = 3 * max(need) + 2 * sorted(need)[len(need) // 2]
    srv = ctypes.c_void_p()
    assert L.fl_server_create(0, 4, budget, ctypes.byref(srv)) == 0
    tickets = []
    try:
        for k, (desc, keep, out) in enumerate(reqs):
            if k % 3 == 2:
                time.sleep(0.05)
            tk = ctypes.c_uint64()
            emu.check(L, L.fl_server_submit(srv, ctypes.byref(desc), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            ctypes.byref(tk)))
            tickets.append(tk.value)
        its = []
        for tk in tickets:
            it = ctypes.c_uint32()
            qms = ctypes.c_double()
            lat = ctypes.c_double()
            emu.check(L, L.fl_server_wait(srv, tk, ctypes.byref(it), ctypes.byref(qms), ctypes.byref(lat)))
            its.append(it.value)
        pb = ctypes.c_uint64()
        pr = ctypes.c_uint32()
        assert L.fl_server_stats(srv, ctypes.byref(pb), ctypes.byref(pr)) == 0
        assert pb.value <= budget and pr.value >= 2
    finally:
        L.fl_server_destroy(srv)
    for t, (desc, keep, out), it in zip(terrains, reqs, its):
        g = flo.Graph(t["row_ptr"], t["col"], t["length"])
        ref, rit, _ = flo.generate(g, t["area"], t["erodibility"], t["uplift"], t["init_elevation"], t["outlets"],
                                   max_iteration=t.get("max_iteration", 0xFFFFFFFF), max_slope=t.get("max_slope"))
        assert it == rit
        assert _bits_equal(out, ref)


def _solve_with_tuning(L, ctx, h, tuning, n):
    from fastlem_amd import _native

    if tuning is None:
        assert L.fl_ctx_set_tuning(ctx, None) == 0
    else:
        t = _native.Tuning(*tuning)
        assert L.fl_ctx_set_tuning(ctx, ctypes.byref(t)) == 0, L.fl_last_error()
    it = ctypes.c_uint32()
    assert L.fl_terrain_solve(ctx, h, 0xFFFFFFFF, ctypes.byref(it)) == 0, L.fl_last_error()
    out = np.empty(n)
    assert L.fl_terrain_download(ctx, h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return out, it.value


def test_emulated_site_limit(emu_lib):
    """Terrains beyond 2^31 - 1 sites (chain positions and light depths of
    the pointer jumping stay below 2^31) are refused with FL_ERR_TOO_LARGE
    before any array is read or any memory taken, by every entry point that
    takes a description; batches and served unions sum their sites first."""
    from fastlem_amd import _native

    L = emu_lib
    d = _native.TerrainDesc()
    d.n = 1 << 31  # the arrays stay null: the size is refused first
    b = ctypes.c_uint64()
    assert L.fl_request_device_bytes(ctypes.byref(d), ctypes.byref(b)) == -4
    assert b"too large" in L.fl_last_error()
    ctx = ctypes.c_void_p()
    assert L.fl_ctx_create(0, ctypes.byref(ctx)) == 0
    try:
        h = ctypes.c_void_p()
        assert L.fl_terrain_create(ctx, ctypes.byref(d), ctypes.byref(h)) == -4
        assert b"2^31-1" in L.fl_last_error()
        out = np.empty(1)
        it = ctypes.c_uint32()
        assert L.fl_generate_f64_noalloc(ctx, ctypes.byref(d), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         ctypes.byref(it)) == -4
    finally:
        L.fl_ctx_destroy(ctx)


def test_emulated_tunings_give_identical_results(emu_lib):
    """Every launch tuning -- workgroup size, the short/long chain switch
    (all chains short, all long), tail launches off and wide, pipelining
    modes -- gives the same elevations bit for bit; the autotuner, which
    checks exactly that for each candidate, completes."""
    from fastlem_amd import _native
    from fastlem_amd.device import batch_descs, DeviceBatch
    from fastlem_amd.workloads import config_fbm

    L = emu_lib
    w = config_fbm(2500, 50.0, seed=8)
    keep = []
    d = batch_descs([DeviceBatch.terrain_of(w)], keep)
    ctx = ctypes.c_void_p()
    assert L.fl_ctx_create(0, ctypes.byref(ctx)) == 0
    h = ctypes.c_void_p()
    assert L.fl_terrain_create(ctx, ctypes.byref(d[0]), ctypes.byref(h)) == 0
    try:
        ref, rit = _solve_with_tuning(L, ctx, h, None, w.n)
        for t in [(64, 1, 0, 1, 65536), (1024, 100000, 64, 1, 65536), (128, 3, 8, 0, 65536),
                  (512, 12, 1000, 3, 65536), (256, 6, 16, 1, 65536)]:
            got, it = _solve_with_tuning(L, ctx, h, t, w.n)
            assert it == rit and _bits_equal(got, ref), t
        assert L.fl_ctx_set_tuning(ctx, None) == 0
        bad = _native.Tuning(100, 6, 16, 1, 65536)
        assert L.fl_ctx_set_tuning(ctx, ctypes.byref(bad)) == -1  # block not a multiple of 64
        best = _native.Tuning()
        bms, tms = ctypes.c_double(), ctypes.c_double()
        tried = ctypes.c_uint32()
        assert L.fl_autotune(ctx, h, 1, ctypes.byref(best), ctypes.byref(bms), ctypes.byref(tms),
                             ctypes.byref(tried)) == 0, L.fl_last_error()
        assert tried.value >= 15 and bms.value <= tms.value
        table = _native.Tuning()
        assert L.fl_ctx_tuning_for(ctx, w.n, 1, ctypes.byref(table)) == 0 and table.site_block >= 64
    finally:
        L.fl_terrain_destroy(h)
        L.fl_ctx_destroy(ctx)
