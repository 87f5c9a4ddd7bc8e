import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


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
