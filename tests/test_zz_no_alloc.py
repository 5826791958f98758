"""Allocation-free tracing (sq_trace_rays_noalloc) against sq_trace_rays with
device rays and outputs: hit ids, t and per-ray counts bit-identical on two
calls in a row, the second smaller, for both semantics; work totals equal;
a tree too deep for the fixed scratch is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

import sequoia_amd as sq
from sequoia_amd import _lib

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def clusters():
    tris = sq.gen_clusters(200_000, 16, seed=7)
    rays = sq.gen_secondary_rays(tris, 300_000, seed=8)
    return tris, rays


def _fill(ctx, dev, n, value):
    ctx.to_device(dev, np.full(n, value, np.uint32))


@pytest.mark.parametrize("semantics", [sq.TRACE_EXHAUSTIVE, sq.TRACE_CLOSEST])
def test_noalloc_equals_workspace_trace(clusters, semantics):
    tris, rays = clusters
    with sq.Context(0) as ctx:
        scene = ctx.upload(tris)
        bvh = ctx.build(scene, "morton63")
        say("noalloc semantics", semantics, "depth", bvh.info()[2])
        nmax = len(rays)
        d_rays = ctx.alloc(nmax * 32)
        ctx.to_device(d_rays, rays)
        bufs = {k: ctx.alloc(nmax * 4) for k in ("id0", "t0", "c0", "id1", "t1", "c1")}
        try:
            for n in (nmax, 70_001):
                say("noalloc n", n)
                for k in bufs:
                    _fill(ctx, bufs[k], nmax, 0xA5A5A5A5)
                ctx.trace_rays_device(bvh, d_rays, n, bufs["id0"], bufs["t0"], bufs["c0"], semantics=semantics,
                                      count_work=True)
                want_work = ctx.timings()
                ctx.trace_rays_device(bvh, d_rays, n, bufs["id1"], bufs["t1"], bufs["c1"], semantics=semantics,
                                      count_work=True, noalloc=True)
                got_work = ctx.timings()
                for a, b in (("id0", "id1"), ("t0", "t1"), ("c0", "c1")):
                    want = ctx.to_host(bufs[a], nmax * 4, np.uint32)
                    got = ctx.to_host(bufs[b], nmax * 4, np.uint32)
                    assert np.array_equal(got[:n], want[:n]), (a, n)
                    # nothing written past the n rays
                    assert (got[n:] == 0xA5A5A5A5).all(), (b, n)
                hits = ctx.to_host(bufs["id1"], n * 4, np.uint32)
                assert (hits != sq.MISS).any()
                assert got_work["node_tests"] == want_work["node_tests"] > 0
                assert got_work["tri_tests"] == want_work["tri_tests"]
        finally:
            for p in bufs.values():
                ctx.free(p)
            ctx.free(d_rays)
        bvh.free()
        scene.free()


def chain_tree(levels):
    """A tree whose interior node i lists (leaf i, interior node i+1): depth
    levels + 1, every triangle reachable once."""
    n_int = levels
    nodes = np.zeros(1 + 2 * n_int, sq.NODE_DTYPE)
    box_min, box_max = np.zeros(3, np.float32), np.ones(3, np.float32)
    # slot 0: root interior; its list at slots 1, 2; slot 2 interior -> list at 3, 4; ...
    for i in range(n_int):
        slot = 0 if i == 0 else 2 * i
        nodes[slot]["min"], nodes[slot]["max"] = box_min, box_max
        nodes[slot]["count"], nodes[slot]["flags"], nodes[slot]["child"] = 2, 0, 2 * i + 1
        leaf = 2 * i + 1
        nodes[leaf]["min"], nodes[leaf]["max"] = box_min, box_max
        nodes[leaf]["count"], nodes[leaf]["flags"], nodes[leaf]["child"] = 1, 1, i
    last = 2 * n_int
    nodes[last]["min"], nodes[last]["max"] = box_min, box_max
    nodes[last]["count"], nodes[last]["flags"], nodes[last]["child"] = 1, 1, n_int
    return nodes


def test_noalloc_rejects_a_tree_too_deep_for_its_scratch():
    levels = 3000
    tris = np.tile(np.array([[0, 0, 0.5, 1, 0, 0.5, 0, 1, 0.5]], np.float32), (levels + 1, 1))
    with sq.Context(0) as ctx:
        scene = ctx.upload(tris)
        bvh = ctx.bvh_from_host(scene, chain_tree(levels), 0)
        assert bvh.info()[2] == levels + 1
        d_rays = ctx.alloc(64 * 32)
        d_id = ctx.alloc(64 * 4)
        try:
            p = _lib.Context._params(sq.TRACE_EXHAUSTIVE, sq.HEATMAP_RAY, 0, 0, False)
            out = _lib.TraceOut(d_id, None, None, None)
            rc = sq.lib().sq_trace_rays_noalloc(ctx.h, bvh.h, d_rays, 64, C.byref(p), C.byref(out))
            assert rc == sq.ERR_BUDGET, rc
            assert b"scratch" in sq.lib().sq_last_error()
        finally:
            ctx.free(d_id)
            ctx.free(d_rays)
        bvh.free()
        scene.free()
