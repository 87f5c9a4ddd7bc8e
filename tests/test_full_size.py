"""Parity at full size: the 1M-site fbm terrain and the 16.7M-site deep strip.

- 1M fbm terrain (sites uniform in [0,1000]^2, one Lloyd step, fBm
  erodibility, ocean outlets): the whole solve against the oracle, bit for bit.
- 16.7M strip (16384 x 1024 jittered grid, outlets on the x = 0 column, rivers
  about 16k sites deep): the first iteration (lake joining included) against
  the oracle's, bit for bit; the solve converges to a state that the oracle's
  iteration maps to itself (a fixed point of the reference's update), and a
  solve started there stops after one iteration with nothing changed.

Each case takes up to a few minutes (the oracle is single-threaded C); they
carry the `fullsize` marker so quick runs can leave them out.
"""
import sys

import numpy as np
import pytest

from conftest import graph_of, init_of

pytestmark = [pytest.mark.gpu, pytest.mark.fullsize]


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _say(msg):
    sys.stderr.write(f"[full-size] {msg}\n")
    sys.stderr.flush()


@pytest.fixture(scope="module")
def fbm1m():
    from fastlem_amd.workloads import config_fbm

    return config_fbm(1_000_000, 1000.0)


@pytest.fixture(scope="module")
def strip16m():
    from fastlem_amd.workloads import config_strip

    return config_strip(16384, 1024)


@pytest.mark.timeout(900)
def test_fbm_1m_end_to_end(fbm1m):
    from fastlem_amd.device import DeviceTerrain
    from oracle import flo

    w = fbm1m
    init = init_of(w)
    t = DeviceTerrain.from_workload(w, init_elevation=init)
    it = t.solve()
    _say(f"fbm 1M: device solve {it} iterations; oracle running")
    ref, rit, _ = flo.generate(graph_of(w), w.model.areas(), w.erodibility, w.uplift, init, w.outlets())
    assert it == rit
    assert _bits_equal(ref, t.elevations())
    t.close()


@pytest.mark.timeout(900)
def test_strip_16m_first_iteration_teacher_forced(strip16m):
    from fastlem_amd.device import DeviceTerrain
    from oracle import flo

    w = strip16m
    elev = init_of(w)
    t = DeviceTerrain.from_workload(w, init_elevation=elev)
    h_got, n_got, da_got, rt_got, ch_got = t.stage_iteration(elev)
    t.close()
    _say("strip 16.7M: device iteration 1 done; oracle running")
    h_ref, n_ref, _, da_ref, rt_ref, ch_ref, lake = flo.iteration(graph_of(w), w.model.areas(), w.erodibility,
                                                                  w.uplift, w.outlets(), elev)
    assert lake, "the jittered initial surface has lakes: the joining path runs"
    assert np.array_equal(n_ref, n_got), f"{int((n_ref != n_got).sum())} receivers differ"
    assert _bits_equal(da_ref, da_got)
    assert _bits_equal(rt_ref, rt_got)
    assert _bits_equal(h_ref, h_got)
    assert ch_ref == ch_got


@pytest.mark.timeout(900)
def test_strip_16m_converges_to_a_fixed_point(strip16m):
    from fastlem_amd.device import DeviceTerrain
    from oracle import flo

    w = strip16m
    init = init_of(w)
    t = DeviceTerrain.from_workload(w, init_elevation=init)
    it = t.solve(20000)
    assert 1 < it < 20000
    h = t.elevations()
    t.close()
    assert np.all(np.isfinite(h))
    assert _bits_equal(h[w.outlets()], init[w.outlets()])  # outlets keep their elevations
    _say(f"strip 16.7M: converged in {it} iterations; re-solving from the result")
    t2 = DeviceTerrain.from_workload(w, init_elevation=h)
    assert t2.solve() == 1
    assert _bits_equal(h, t2.elevations())
    t2.close()
    _say("strip 16.7M: oracle iteration at the converged state")
    h_ref, _, _, _, _, changed, _ = flo.iteration(graph_of(w), w.model.areas(), w.erodibility, w.uplift,
                                                  w.outlets(), h)
    assert not changed
    assert _bits_equal(h_ref, h)
