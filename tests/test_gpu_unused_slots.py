"""Morton builds with refit take no zero fill of the node array: the emission
writes every slot, the unused ones (fan tails, the array's last slot) as
zeros. The test hook SEQUOIA_TEST_POISON_NODES (read at context creation)
fills those node arrays with all-ones words before each build, so a slot the
emission misses differs from the oracle's zero slot. Plain builds (both
layouts, top-phase thresholds 0 and 64, scenes of 1 to 1M triangles) and
batched builds are compared with the oracle node for node."""
import os

import numpy as np
import pytest

from conftest import nodes_equal
from oracle import binding as ob
from scenes import GOLDEN_SCENES

import sequoia_amd as sq

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def pctx():
    os.environ["SEQUOIA_TEST_POISON_NODES"] = "1"
    try:
        c = sq.Context(0)
    finally:
        del os.environ["SEQUOIA_TEST_POISON_NODES"]
    yield c
    c.close()


def check_plain(c, tris, kind, bits, layout=sq.LAYOUT_REFERENCE, threshold=0):
    s = c.upload(tris)
    b = c.build(s, kind, layout=layout, subtree_threshold=threshold)
    got = b.download()
    b.free()
    s.free()
    want, _, _, _ = ob.build_morton(tris, bits, layout=layout)
    assert len(got) == len(want)
    assert len(nodes_equal(got, want)) == 0, (len(tris), kind, layout, threshold)


def test_plain_builds_write_every_slot(pctx):
    rng = np.random.default_rng(29)
    for n in (1, 2, 3, 5, 7, 8, 9, 15, 64, 2047, 2048, 2049, 30000):
        tris = rng.random((n, 9), dtype=np.float32)
        say("poisoned build", n)
        for layout in (sq.LAYOUT_REFERENCE, sq.LAYOUT_FIXED):
            for threshold in (0, 64):
                check_plain(pctx, tris, "morton63", 21, layout, threshold)
    dup = np.tile(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), (5000, 1))
    check_plain(pctx, dup, "morton63", 21)
    for name in ("ico69632", "clus3000"):
        for kind, bits in (("morton", 10), ("morton2", 20), ("morton63", 21)):
            say("poisoned build", name, kind)
            for threshold in (0, 64):
                check_plain(pctx, GOLDEN_SCENES[name](), kind, bits, threshold=threshold)
    say("poisoned build 1M soup")
    check_plain(pctx, sq.gen_soup(1_000_000, 2, 0.01), "morton63", 21)


def test_batched_builds_write_every_slot(pctx):
    scenes = [sq.gen_soup(1, 3, 0.01), sq.gen_soup(5, 11, 0.01), sq.gen_soup(9, 12, 0.01),
              sq.gen_icosphere(1280, 1, 0.1), sq.gen_clusters(3000, 16, 5), sq.gen_soup(70000, 22, 0.01),
              np.repeat(sq.gen_soup(1, 4, 0.01), 700, axis=0)]
    b = sq.Batch(pctx, scenes)
    try:
        for kind, bits in (("morton63", 21), ("morton", 10)):
            say("poisoned batch", kind)
            b.build(kind)
            for i, t in enumerate(scenes):
                want, _, _, _ = ob.build_morton(t, bits)
                got = b.bvh(i).download()
                assert len(got) == len(want)
                assert len(nodes_equal(got, want)) == 0, (kind, i)
    finally:
        b.free()
