import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

GOLDEN = os.path.join(TESTS, "golden")
DATASETS = os.path.join(GOLDEN, "datasets")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 (MI355X) device")


@pytest.fixture(scope="session")
def ctx():
    import bc_amd
    return bc_amd.Context(0)


def probe_depth(g):
    """BFS depth from the highest-degree vertex (the lowest id among equals):
    the probe the engine runs when a graph is uploaded."""
    import numpy as np
    deg = np.diff(g.row_ptr)
    s = int(np.argmax(deg))
    dist = np.full(g.n, -1, np.int64)
    dist[s] = 0
    front = np.array([s])
    d = 0
    while len(front):
        cnt = deg[front]
        idx = np.repeat(g.row_ptr[front] - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(cnt.sum())
        nb = np.unique(g.col[idx])
        nb = nb[dist[nb] < 0]
        if len(nb) == 0:
            break
        d += 1
        dist[nb] = d
        front = nb
    return d
