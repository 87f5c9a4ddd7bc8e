"""Pinning the CPU oracle (oracle/fl_oracle.c) before trusting it.

The reference has no golden vectors and no assertions (its tests render
PNGs), and it cannot be built here (Rust), so each oracle function is pinned
by checks that a plausible mistake would fail:

* receivers        brute-force definition (lexicographic max of (slope, -slot))
                   on hand graphs and random graphs, plus invariants;
* chain ends       brute-force walk; has_lake = some non-outlet pit;
* lake removal     an independent Prim-order formulation on graphs with
                   distinct edge lengths (brute-force Prim) and a hand-derived
                   Rust BinaryHeap trace for the equal-key (outlet) phase;
* basin BFS order  definition (level-major, parent position, slot);
* drainage area    closed forms (outlet total = basin area sum) and an
                   association-sensitive fixture (reverse row order);
* response time /  the closed-form steady state on a uniform chain, outlet
  elevation        cancellation, the max-slope clamp identity;
* the loop         fixed point: one more iteration changes no bit.
"""
import math

import numpy as np
import pytest

from oracle import flo


def csr_from_edges(n, edges):
    """Rows in insertion order, as terrain-graph's add_edge builds them."""
    rows = [[] for _ in range(n)]
    for a, b, w in edges:
        rows[a].append((b, w))
        rows[b].append((a, w))
    rp = np.zeros(n + 1, np.uint32)
    col, ln = [], []
    for i, r in enumerate(rows):
        for j, w in r:
            col.append(j)
            ln.append(w)
        rp[i + 1] = len(col)
    return flo.Graph(rp, np.array(col, np.uint32), np.array(ln, np.float64))


def rows_of(g):
    return [[(int(g.col[k]), float(g.length[k])) for k in range(g.row_ptr[i], g.row_ptr[i + 1])] for i in range(g.n)]


def random_delaunay_graph(n, seed, grid=False):
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    if grid:
        side = int(math.ceil(math.sqrt(n)))
        xs, ys = np.meshgrid(np.arange(side, dtype=float), np.arange(side, dtype=float))
        pts = np.column_stack([xs.ravel(), ys.ravel()])[:n]
        pts = pts + rng.uniform(-1e-3, 1e-3, pts.shape) * 0  # exact grid: equal lengths
    else:
        pts = rng.uniform(0, 10, (n, 2))
    tri = Delaunay(pts)
    edges = []
    seen = set()
    for t in tri.simplices:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            key = (min(a, b), max(a, b))
            if key in seen:
                continue
            seen.add(key)
            w = float(np.hypot(*(pts[a] - pts[b])))
            edges.append((int(a), int(b), w))
    return csr_from_edges(len(pts), edges), pts


# ------------------------------------------------------------- receivers ---

def brute_receivers(g, h, outlets):
    out = set(int(o) for o in outlets)
    nxt = np.arange(g.n, dtype=np.uint32)
    for i, row in enumerate(rows_of(g)):
        if i in out:
            continue
        cands = []
        for slot, (j, w) in enumerate(row):
            if h[i] > h[j]:
                s = (h[i] - h[j]) / w
                if s > 0.0 and not math.isnan(s):
                    cands.append((s, -slot, j))
        if cands:
            nxt[i] = max(cands)[2]
    return nxt


def test_receivers_hand_graphs():
    # chain 0-1-2-3, outlet 0
    g = csr_from_edges(4, [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.0)])
    assert list(flo.receivers(g, [0.0, 1.0, 2.0, 3.0], [0])) == [0, 0, 1, 2]
    # star: centre 0 higher than three neighbours at equal slope -> first slot
    g = csr_from_edges(4, [(0, 3, 2.0), (0, 1, 2.0), (0, 2, 2.0)])
    assert flo.receivers(g, [4.0, 2.0, 2.0, 2.0], [1])[0] == 3
    # plateau: equal heights -> no receiver (pit)
    g = csr_from_edges(3, [(0, 1, 1.0), (1, 2, 1.0)])
    assert list(flo.receivers(g, [1.0, 1.0, 1.0], [2])) == [0, 1, 2]
    # steeper but farther: slope, not drop, decides
    g = csr_from_edges(3, [(0, 1, 1.0), (0, 2, 10.0)])
    assert flo.receivers(g, [5.0, 4.0, 0.0], [1])[0] == 1
    # outlets never get a receiver even when higher
    g = csr_from_edges(2, [(0, 1, 1.0)])
    assert list(flo.receivers(g, [9.0, 0.0], [0])) == [0, 1]


@pytest.mark.parametrize("seed", range(6))
def test_receivers_random_match_brute_force(seed):
    g, _ = random_delaunay_graph(120, seed)
    rng = np.random.default_rng(seed)
    h = rng.uniform(0, 1, g.n)
    if seed % 2:
        h = np.round(h * 4) / 4  # many equal heights and equal slopes
    outlets = np.sort(rng.choice(g.n, 10, replace=False)).astype(np.uint32)
    got = flo.receivers(g, h, outlets)
    assert np.array_equal(got, brute_receivers(g, h, outlets))
    # invariants: receiver is a neighbour (or self), strictly lower
    rows = rows_of(g)
    for i in range(g.n):
        j = int(got[i])
        if j != i:
            assert j in [a for a, _ in rows[i]]
            assert h[j] < h[i]


# ------------------------------------------------------------ chain ends ---

def test_find_roots_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(5):
        g, _ = random_delaunay_graph(80, int(rng.integers(1000)))
        h = rng.uniform(0, 1, g.n)
        outlets = np.sort(rng.choice(g.n, 5, replace=False)).astype(np.uint32)
        nxt = flo.receivers(g, h, outlets)
        sub, lake = flo.find_roots(nxt, outlets)
        ends = []
        for i in range(g.n):
            j = i
            while nxt[j] != j:
                j = nxt[j]
            ends.append(j)
        assert list(sub) == ends
        out = set(outlets.tolist())
        assert lake == any(int(nxt[i]) == i and i not in out for i in range(g.n))


# ---------------------------------------------------------- lake removal ---

def brute_prim_order(g, outlets):
    """Prim from the outlet set with distinct edge lengths: the outlets, then
    repeatedly the unvisited site behind the shortest edge out of the
    visited set.  Returns the non-outlet part of the order."""
    rows = rows_of(g)
    visited = set(int(o) for o in outlets)
    order = []
    while True:
        best = None
        for i in visited:
            for j, w in rows[i]:
                if j not in visited and (best is None or w < best[0]):
                    best = (w, j)
        if best is None:
            return order
        visited.add(best[1])
        order.append(best[1])


@pytest.mark.parametrize("seed", range(4))
def test_visit_order_is_prim_order_for_distinct_lengths(seed):
    g, _ = random_delaunay_graph(90, 100 + seed)
    rng = np.random.default_rng(seed)
    outlets = rng.choice(g.n, 6, replace=False).astype(np.uint32)
    order = flo.visit_order(g, outlets)
    assert sorted(order[:6].tolist()) == sorted(outlets.tolist())
    assert order[6:].tolist() == brute_prim_order(g, outlets)


def test_heap_tie_order_hand_derived():
    """Outlets [A,B,C,D] = sites [0,1,2,3] all at key 0; only A has an
    unvisited non-outlet neighbour (site 4), pushed when A pops.  Tracing std
    BinaryHeap (pop: last element to the root, sift_down_to_bottom taking the
    right child on equal keys, then sift_up) gives the pop order A, C, D, B:
      [A,B,C,D] -pop A-> [C,B,D] -push 4-> [C,B,D,4] -pop C-> [D,B,4]
      -pop D-> [B,4] -pop B-> [4] -pop 4."""
    g = csr_from_edges(5, [(0, 4, 1.0)])
    order = flo.visit_order(g, np.array([0, 1, 2, 3], np.uint32))
    assert order.tolist() == [0, 2, 3, 1, 4]


def rank_based_lake_join(g, outlets, nxt, sub):
    """The device algorithm, restated independently in Python: each lake
    (chain ending in a non-outlet pit) joins through the smallest
    (visit rank of i, slot) over edges i->j with j in the lake and i outside
    it; then the path j -> pit is reversed."""
    order = flo.visit_order(g, outlets)
    rank = {int(v): r for r, v in enumerate(order)}
    out = set(int(o) for o in outlets)
    rows = rows_of(g)
    best = {}
    for i in range(g.n):
        if i not in rank:
            continue
        for slot, (j, _) in enumerate(rows[i]):
            s = int(sub[j])
            if s != int(sub[i]) and s not in out:
                key = (rank[i], slot)
                if s not in best or key < best[s]:
                    best[s] = key
    nxt = nxt.copy()
    for s, (r, slot) in best.items():
        i = int(order[r])
        j = rows[i][slot][0]
        k, nk = j, i
        while nxt[k] != k:
            tmp = int(nxt[k])
            nxt[k] = nk
            nk, k = k, tmp
        nxt[k] = nk
    return nxt


@pytest.mark.parametrize("seed,grid", [(0, False), (1, False), (2, True), (3, True), (4, False)])
def test_lake_removal_matches_rank_formulation(seed, grid):
    g, _ = random_delaunay_graph(150, 200 + seed, grid=grid)
    rng = np.random.default_rng(seed)
    h = rng.uniform(0, 1, g.n)  # pits everywhere
    outlets = np.sort(rng.choice(g.n, 8, replace=False)).astype(np.uint32)
    nxt_ref, sub, lake = flo.stream_tree(g, h, outlets)
    assert lake
    nxt0 = flo.receivers(g, h, outlets)
    assert np.array_equal(rank_based_lake_join(g, outlets, nxt0, sub), nxt_ref)
    # the result is a forest rooted at outlets only, along graph edges
    rows = rows_of(g)
    out = set(outlets.tolist())
    for v in range(g.n):
        j, steps = v, 0
        while nxt_ref[j] != j:
            assert int(nxt_ref[j]) in [a for a, _ in rows[j]]
            j = int(nxt_ref[j])
            steps += 1
            assert steps <= g.n
        assert j in out


def test_unreachable_lake_stays_a_pit():
    # component {0,1,2} has the outlet; {3,4} has no outlet
    g = csr_from_edges(5, [(0, 1, 1.0), (1, 2, 1.0), (3, 4, 1.0)])
    h = np.array([0.0, 0.5, 0.2, 0.3, 0.6])
    nxt, _, lake = flo.stream_tree(g, h, [0])
    assert lake
    assert nxt[3] == 3 and nxt[4] == 3  # untouched
    assert nxt[2] == 1 and nxt[1] == 0  # joined through the outlet side


# ------------------------------------------------------------ basin order ---

def test_basin_order_definition():
    g, _ = random_delaunay_graph(100, 7)
    rng = np.random.default_rng(7)
    outlets = np.array([3], np.uint32)
    h = rng.uniform(0, 1, g.n)
    nxt, _, _ = flo.stream_tree(g, h, outlets)
    order = flo.basin_order(g, nxt, 3).tolist()
    rows = rows_of(g)
    expect = [3]
    q = 0
    while q < len(expect):
        u = expect[q]
        expect.extend(j for j, _ in rows[u] if nxt[j] == u and j != u)
        q += 1
    assert order == expect
    assert len(order) == g.n  # everything drains to the single outlet


# --------------------------------------------------------- drainage area ---

def run_iteration(g, h, outlets, area=None, K=None, U=None, ms=None):
    n = g.n
    area = np.ones(n) if area is None else np.asarray(area, float)
    K = np.ones(n) if K is None else np.asarray(K, float)
    U = np.ones(n) if U is None else np.asarray(U, float)
    return flo.iteration(g, area, K, U, outlets, h, ms)


def test_drainage_area_totals():
    g, _ = random_delaunay_graph(200, 11)
    rng = np.random.default_rng(11)
    outlets = np.sort(rng.choice(g.n, 7, replace=False)).astype(np.uint32)
    area = rng.uniform(0.5, 2.0, g.n)
    h, nxt, sub, da, rt, ch, lake = run_iteration(g, rng.uniform(0, 1, g.n), outlets, area=area)
    roots = []
    for v in range(g.n):
        j = v
        while nxt[j] != j:
            j = int(nxt[j])
        roots.append(j)
    roots = np.array(roots)
    for o in outlets:
        members = area[roots == o]
        assert da[o] == pytest.approx(math.fsum(members), rel=len(members) * 2.3e-16)
    assert math.fsum(da[outlets]) == pytest.approx(math.fsum(area), rel=1e-13)


def test_drainage_area_association_order():
    """Site 0 (outlet) has children 1 and 2 in row order [1, 2].  The
    reference adds them in reverse BFS order: A_0 = (a_0 + A_2) + A_1.  The
    areas are chosen so that the forward order gives different bits."""
    a0, a1, a2 = 1.0, 2.0 ** -53, 2.0 ** -52
    fwd = (a0 + a1) + a2
    rev = (a0 + a2) + a1
    assert fwd != rev
    g = csr_from_edges(3, [(0, 1, 1.0), (0, 2, 1.0)])
    h = np.array([0.0, 1.0, 1.0])
    _, nxt, _, da, _, _, _ = run_iteration(g, h, [0], area=[a0, a1, a2])
    assert list(nxt) == [0, 0, 0]
    assert da[0] == rev


# ---------------------------------------------- response time / elevation ---

def test_uniform_chain_closed_form():
    """Chain of n sites spaced d apart, outlet at 0, area a, K, U constant:
    A_r = (n - r) a, tau_r - tau_0 = sum_{q=1..r} d / (K sqrt(A_q)),
    h_r = U (tau_r - tau_0) (the steady state h = U/K * integral dx / A^m)."""
    n, d, a, K, U = 40, 0.7, 1.3, 0.8, 1.7
    g = csr_from_edges(n, [(i, i + 1, d) for i in range(n - 1)])
    h0 = np.arange(n, dtype=float)  # strictly increasing away from the outlet
    h, nxt, _, da, rt, _, _ = run_iteration(g, h0, [0], area=np.full(n, a), K=np.full(n, K),
                                            U=np.full(n, U))
    for r in range(n):
        assert da[r] == pytest.approx((n - r) * a, rel=1e-13)
        expect = U * math.fsum(d / (K * math.sqrt((n - q) * a)) for q in range(1, r + 1))
        assert h[r] == pytest.approx(expect, rel=1e-12, abs=1e-15)
    assert h[0] == 0.0  # the outlet keeps its elevation (tau_o - tau_o = 0)
    assert rt[0] == pytest.approx(1.0 / (K * math.sqrt(n * a)))  # outlet distance is 1.0


def test_outlet_keeps_elevation():
    g, _ = random_delaunay_graph(100, 13)
    rng = np.random.default_rng(13)
    outlets = np.sort(rng.choice(g.n, 5, replace=False)).astype(np.uint32)
    h0 = rng.uniform(0, 1, g.n)
    h, *_ = run_iteration(g, h0, outlets)
    assert np.array_equal(h[outlets], h0[outlets])


def test_max_slope_clamp_identity():
    n, d = 30, 1.0
    g = csr_from_edges(n, [(i, i + 1, d) for i in range(n - 1)])
    sigma = 0.05
    h, nxt, *_ = run_iteration(g, np.arange(n, dtype=float), [0], ms=np.full(n, sigma))
    t = math.tan(sigma)
    for r in range(1, n):
        slope = (h[r] - h[nxt[r]]) / d
        assert slope <= t * (1 + 1e-12)
    # near the outlet the drainage area is large and the slope small; upstream
    # the clamp binds: h_r - h_next = tan(sigma) * d exactly as written
    r = n - 1
    assert h[r] == h[r - 1] + t * d


def test_fixed_point_is_idempotent():
    g, _ = random_delaunay_graph(150, 17)
    rng = np.random.default_rng(17)
    outlets = np.sort(rng.choice(g.n, 6, replace=False)).astype(np.uint32)
    init = rng.uniform(0, 1e-12, g.n)
    final, iters, lakes = flo.generate(g, np.ones(g.n), np.ones(g.n), np.ones(g.n), init, outlets)
    assert lakes == 1 and iters > 1
    h, _, _, _, _, changed, lake = run_iteration(g, final, outlets)
    assert not changed and not lake
    assert np.array_equal(h.view(np.uint64), final.view(np.uint64))


def test_max_iteration_caps_the_loop():
    g, _ = random_delaunay_graph(120, 19)
    rng = np.random.default_rng(19)
    outlets = np.array([0, 5], np.uint32)
    init = rng.uniform(0, 1e-12, g.n)
    _, full, _ = flo.generate(g, np.ones(g.n), np.ones(g.n), np.ones(g.n), init, outlets)
    _, capped, _ = flo.generate(g, np.ones(g.n), np.ones(g.n), np.ones(g.n), init, outlets, max_iteration=2)
    assert full > 2 and capped == 2
    _, zero, _ = flo.generate(g, np.ones(g.n), np.ones(g.n), np.ones(g.n), init, outlets, max_iteration=0)
    assert zero == 0
