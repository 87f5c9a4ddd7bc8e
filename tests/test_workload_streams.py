"""The batched and served forms on their own workload generators, against the
oracle.

* Batched (512 log-uniform terrains of 1,000..262,144 sites in one union,
  boundary outlets, fBm erodibility): a reduced stream of the same generator,
  24 terrains of 1,000..20,000 sites, solved as one union; every terrain's
  elevations and iteration count must equal its solo oracle run bit for bit.
* Served (a Poisson stream of generate() requests, log-uniform sizes,
  ocean-mask outlets, fBm erodibility, HBM-budget admission): a reduced
  stream of 16 requests of 4,096..32,768 sites with exponential gaps between
  submissions and a budget that holds only part of the stream, so requests
  queue, are admitted as others finish and overlap on the device; every
  result must equal its solo oracle run bit for bit.

The terrains are built on the host before the first GPU call of each test
(no process pool: workers=1).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _oracle(tr):
    from oracle import flo

    g = flo.Graph(tr["row_ptr"], tr["col"], tr["length"])
    ref, ref_it, _ = flo.generate(g, tr["area"], tr["erodibility"], tr["uplift"], tr["init_elevation"],
                                  tr["outlets"], max_iteration=tr.get("max_iteration", 0xFFFFFFFF),
                                  max_slope=tr.get("max_slope"))
    return ref, ref_it


@pytest.fixture(scope="module")
def batched_stream():
    from fastlem_amd.workloads import terrain_stream

    return terrain_stream(24, 1000, 20000, 4242, "boundary", workers=1)


@pytest.fixture(scope="module")
def served_stream():
    from fastlem_amd.workloads import terrain_stream

    return terrain_stream(16, 4096, 32768, 5151, "ocean", workers=1)


def test_batched_stream_matches_solo_oracle_runs(batched_stream):
    from fastlem_amd.device import DeviceBatch

    terrains = batched_stream
    sizes = [len(t["row_ptr"]) - 1 for t in terrains]
    print(f"batched stream: {len(terrains)} terrains, {sum(sizes)} sites, {min(sizes)}..{max(sizes)}", flush=True)
    b = DeviceBatch(terrains)
    its = b.solve()
    stats = b.last_stats()
    for i, tr in enumerate(terrains):
        ref, ref_it = _oracle(tr)
        assert its[i] == ref_it, (i, sizes[i], its[i], ref_it)
        assert _bits_equal(ref, b.elevations(i)), (i, sizes[i])
    # a second solve of the same union (visit orders recomputed) is identical
    from fastlem_amd import _native

    _native.lib().fl_terrain_invalidate_cache(b._h)
    its2 = b.solve()
    assert list(its2) == list(its)
    assert stats["iterations"] == max(its)


def test_served_poisson_stream_matches_solo_oracle_runs(served_stream):
    from fastlem_amd.serving import Server, request_bytes

    terrains = served_stream
    sizes = [len(t["row_ptr"]) - 1 for t in terrains]
    print(f"served stream: {len(terrains)} requests, {sum(sizes)} sites, {min(sizes)}..{max(sizes)}", flush=True)
    need = sorted((request_bytes(t) for t in terrains), reverse=True)
    budget = need[0] + need[1] + need[2]  # a fraction of the stream: the rest queues
    assert budget < sum(need)
    srv = Server(slots=4, hbm_budget_bytes=budget)
    gaps = np.random.default_rng(17).exponential(0.01, len(terrains))  # Poisson arrivals, 100 per second
    tickets = []
    for tr, gap in zip(terrains, gaps):
        time.sleep(float(gap))
        tickets.append(srv.submit(tr))
    results = [srv.wait(k) for k in tickets]
    st = srv.stats()
    srv.close()
    assert st["peak_running"] >= 2, st
    assert st["peak_bytes"] <= budget, st
    for i, (tr, (h, its, q_ms, lat_ms)) in enumerate(zip(terrains, results)):
        ref, ref_it = _oracle(tr)
        assert its == ref_it, (i, sizes[i], its, ref_it)
        assert _bits_equal(ref, h), (i, sizes[i])
        assert lat_ms >= q_ms >= 0
