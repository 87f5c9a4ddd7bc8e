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
