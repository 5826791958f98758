import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def golden(name):
    return np.load(os.path.join(GOLDEN, "ref_%s.npz" % name))


@pytest.fixture(scope="session")
def ctx():
    import sequoia_amd as sq
    c = sq.Context(0)
    yield c
    c.close()


def tree_depth(nodes, root):
    """Levels from the root (1) to the deepest node, by breadth-first expansion
    of the interior nodes' child ranges."""
    frontier = np.array([root], np.int64)
    depth = 0
    while len(frontier):
        depth += 1
        inner = frontier[nodes["flags"][frontier] == 0]
        cnt = nodes["count"][inner].astype(np.int64)
        first = nodes["child"][inner].astype(np.int64)
        frontier = np.repeat(first, cnt) + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    return depth


def nodes_equal(a, b, fields=("min", "count", "flags", "max", "child")):
    """Indices where two Node1 arrays differ bitwise in the given fields."""
    bad = np.zeros(len(a), bool)
    for f in fields:
        x, y = a[f], b[f]
        if x.ndim == 2:
            bad |= (x.view(np.uint32) != y.view(np.uint32)).any(1)
        else:
            bad |= x != y
    return np.nonzero(bad)[0]
