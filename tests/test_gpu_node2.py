"""Node2 trees on the GPU: the device conversion writes the same bytes as the
CPU restatement (nodes, root record, reordered triangles, original indices)
for Morton, SAH and grid trees, reference and fixed layouts; traversal counts
and hits equal the restatement's for ray lists and cameras, exhaustive and
closest; closest hits equal the Node1 tree's."""
import numpy as np
import pytest

from oracle import binding as ob
from scenes import GOLDEN_SCENES
from test_gpu_trace import check_hits

import sequoia_amd as sq

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


CASES = [
    ("ico69632", "recursive", sq.LAYOUT_REFERENCE),
    ("clus3000", "morton63", sq.LAYOUT_REFERENCE),
    ("clus3000", "morton63", sq.LAYOUT_FIXED),
    ("soup2000", "grid_sse", sq.LAYOUT_REFERENCE),
    ("ico1280", "morton", sq.LAYOUT_REFERENCE),
]


def convert_both(ctx, tris, kind, layout, max_leaf=0):
    s = ctx.upload(tris)
    b = ctx.build(s, kind, layout=layout)
    nodes = b.download()
    t2 = ctx.node2(b, max_leaf=max_leaf)
    got = t2.download()
    want = ob.node2_convert(nodes, b.root, tris, max_leaf=max_leaf)
    return s, b, t2, got, want


def same_root(g, w):
    return (list(g.min) == [float(x) for x in w["min"]] and list(g.max) == [float(x) for x in w["max"]]
            and (g.kind, g.child, g.count) == (int(w["kind"]), int(w["child"]), int(w["count"])))


@pytest.mark.parametrize("name,kind,layout", CASES)
def test_node2_conversion_bit_exact(ctx, name, kind, layout):
    say("node2 convert", name, kind, layout)
    tris = GOLDEN_SCENES[name]()
    s, b, t2, got, want = convert_both(ctx, tris, kind, layout)
    assert len(got[0]) == len(want[0])
    assert got[0].tobytes() == want[0].tobytes()
    assert same_root(got[1], want[1])
    assert np.array_equal(got[2].view(np.uint32), want[2].view(np.uint32))
    assert np.array_equal(got[3], want[3])
    t2.free(); b.free(); s.free()


@pytest.mark.parametrize("name,kind,layout", CASES)
def test_node2_trace_matches_oracle(ctx, name, kind, layout):
    say("node2 trace", name, kind, layout)
    tris = GOLDEN_SCENES[name]()
    s, b, t2, got, want = convert_both(ctx, tris, kind, layout)
    n2, r2, tris2, perm = want
    rays = sq.gen_secondary_rays(tris, 20000, seed=5, tmin=1e-4)
    g = ctx.trace_rays(t2, rays)
    o = ob.node2_trace(n2, r2, tris2, perm, rays)
    assert np.array_equal(g["counts"], o["counts"])
    check_hits(g["hit_id"], g["hit_t"], o["hit_id"], o["hit_t"])
    cl = ctx.trace_rays(t2, rays, semantics=sq.TRACE_CLOSEST)
    check_hits(cl["hit_id"], cl["hit_t"], o["hit_id"], o["hit_t"])
    # and the Node1 tree's closest hits
    n1 = ctx.trace_rays(b, rays, semantics=sq.TRACE_CLOSEST)
    diff = np.nonzero(n1["hit_id"] != cl["hit_id"])[0]
    assert len(diff) <= max(2, len(rays) // 2000), len(diff)
    # camera rays: per-pixel counts summed over samples, heatmap bytes
    cam = sq.pinhole_camera(tris, 96, 64, 60.0, 3.0, spp=2, jitter=1, seed=8)
    gc = ctx.trace_camera(t2, cam)
    oc = ob.node2_trace(n2, r2, tris2, perm, ob.camera_rays(cam))
    assert np.array_equal(gc["counts"], oc["counts"].reshape(-1, 2).sum(1))
    check_hits(gc["hit_id"], gc["hit_t"], oc["hit_id"], oc["hit_t"])
    t2.free(); b.free(); s.free()


@pytest.mark.parametrize("max_leaf", [1, 2, 3])
def test_node2_leaf_group_sizes(ctx, max_leaf):
    say("node2 max_leaf", max_leaf)
    tris = GOLDEN_SCENES["clus3000"]()
    s, b, t2, got, want = convert_both(ctx, tris, "morton63", sq.LAYOUT_REFERENCE, max_leaf)
    assert got[0].tobytes() == want[0].tobytes()
    assert np.array_equal(got[3], want[3])
    rays = sq.gen_secondary_rays(tris, 5000, seed=6, tmin=1e-4)
    g = ctx.trace_rays(t2, rays)
    o = ob.node2_trace(*want, rays)
    assert np.array_equal(g["counts"], o["counts"])
    check_hits(g["hit_id"], g["hit_t"], o["hit_id"], o["hit_t"])
    t2.free(); b.free(); s.free()


def test_node2_large_soup(ctx):
    """1M-triangle soup, 63-bit Morton tree: conversion bit-exact, closest
    hits of the headline camera equal the Node1 tree's on a 1/16 crop."""
    say("node2 1M")
    tris = sq.gen_soup(1_000_000, 2, 0.01)
    s, b, t2, got, want = convert_both(ctx, tris, "morton63", sq.LAYOUT_REFERENCE)
    assert got[0].tobytes() == want[0].tobytes()
    assert np.array_equal(got[3], want[3])
    cam = sq.pinhole_camera(tris, 480, 270, 60.0, 3.0, spp=1, jitter=1, seed=2)
    a = ctx.trace_camera(t2, cam, semantics=sq.TRACE_CLOSEST, want_counts=False, want_rgb=False)
    c = ctx.trace_camera(b, cam, semantics=sq.TRACE_CLOSEST, want_counts=False, want_rgb=False)
    assert (a["hit_id"] != c["hit_id"]).sum() <= 5
    t2.free(); b.free(); s.free()


def test_node2_rejects_unsupported(ctx):
    tris = GOLDEN_SCENES["ico1280"]()
    s = ctx.upload(tris)
    b = ctx.build(s, "recursive")
    t2 = ctx.node2(b)
    cam = sq.pinhole_camera(tris, 32, 32, 60.0, 3.0, spp=1)
    with pytest.raises(sq.SqError):
        ctx.trace_camera(t2, cam, heatmap=sq.HEATMAP_PACKET, width=8, want_hits=False)
    t2.free(); b.free(); s.free()
