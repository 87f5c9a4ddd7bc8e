import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "fullsize: parity at the workloads' full sizes (minutes each)")


@pytest.fixture(scope="session")
def plain10k():
    from fastlem_amd.workloads import config_plain

    return config_plain(10000)


@pytest.fixture(scope="session")
def fbm20k():
    from fastlem_amd.workloads import config_fbm

    return config_fbm(20000, 141.0, seed=3)


def graph_of(w):
    from oracle.flo import Graph

    rp, col, ln = w.model.graph()
    return Graph(rp, col, ln)


def init_of(w):
    from fastlem_amd.generator import jitter_elevations

    return jitter_elevations(w.base, 0)


def comb_terrain(spine: int = 24, leaves: int = 6, seed: int = 0) -> dict:
    """A spine s_0..s_{spine-1} draining to the outlet s_0, each spine site
    with `leaves` leaf children listed in its row before its receiver and
    its upstream spine neighbour.  The spine is one heavy chain of `spine`
    sites and each of its sites folds `leaves` light children after the
    heavy one, the rare case of the drainage-area kernels."""
    rng = np.random.default_rng(seed)
    n = spine * (1 + leaves)
    rows = [[] for _ in range(n)]
    leaf = lambda i, j: spine + i * leaves + j
    for i in range(spine):
        for j in range(leaves):
            rows[i].append(leaf(i, j))
            rows[leaf(i, j)].append(i)
        if i > 0:
            rows[i].append(i - 1)
        if i + 1 < spine:
            rows[i].append(i + 1)
    row_ptr = np.zeros(n + 1, np.uint32)
    row_ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.uint32)
    length = rng.uniform(0.5, 1.5, len(col))
    # symmetric lengths
    lk = {}
    for u in range(n):
        for k in range(row_ptr[u], row_ptr[u + 1]):
            key = (min(u, int(col[k])), max(u, int(col[k])))
            lk.setdefault(key, length[k])
            length[k] = lk[key]
    area = np.ones(n)
    area[:spine] = 50.0
    elev = np.zeros(n)
    elev[:spine] = np.arange(spine) * 1.0
    for i in range(spine):
        for j in range(leaves):
            elev[leaf(i, j)] = i + 0.5 + 0.01 * j
    return dict(row_ptr=row_ptr, col=col, length=length, area=area, erodibility=np.ones(n), uplift=np.ones(n),
                init_elevation=elev, outlets=np.array([0], np.uint32))
