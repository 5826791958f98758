"""Wide (Node<K>, K = 4 and 8) trees on the GPU: the device conversion writes
the same bytes as the CPU restatement (nodes, root record, levels) for
Morton, SAH and grid trees, collapsed or not; traversal counts and hits equal
the restatement's for ray lists and cameras, exhaustive and closest; without
collapsing, an SAH tree's per-ray counts are the reference's own (golden
SHA-256 of the compiled reference's traverse); closest hits equal the Node1
tree's, at full size too."""
import hashlib

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
    ("clus3000", "morton", sq.LAYOUT_REFERENCE),
    ("soup2000", "grid_sse", sq.LAYOUT_REFERENCE),
    ("clus3000", "grid", sq.LAYOUT_REFERENCE),
]


def convert_both(ctx, tris, kind, layout, k, collapse=True):
    s = ctx.upload(tris)
    b = ctx.build(s, kind, layout=layout)
    nodes = b.download()
    tk = ctx.nodek(b, k, collapse)
    got = tk.download()
    want = ob.nodek_convert(nodes, b.root, k, collapse)
    return s, b, tk, got, want


def same_root(g, w):
    return (list(g.min) == [float(x) for x in w["min"]] and list(g.max) == [float(x) for x in w["max"]]
            and (g.kind, g.child, g.count) == (int(w["kind"]), int(w["child"]), int(w["count"])))


@pytest.mark.parametrize("collapse", [True, False])
@pytest.mark.parametrize("k", [4, 8])
@pytest.mark.parametrize("name,kind,layout", CASES)
def test_nodek_conversion_bit_exact(ctx, name, kind, layout, k, collapse):
    say("nodek convert", name, kind, k, collapse)
    tris = GOLDEN_SCENES[name]()
    s, b, tk, got, want = convert_both(ctx, tris, kind, layout, k, collapse)
    assert len(got[0]) == len(want[0])
    assert got[0].tobytes() == want[0].tobytes()
    assert same_root(got[1], want[1])
    assert tk.info()["levels"] == want[2]
    tk.free(); b.free(); s.free()


@pytest.mark.parametrize("k", [4, 8])
@pytest.mark.parametrize("name,kind,layout", CASES)
def test_nodek_trace_matches_oracle(ctx, name, kind, layout, k):
    say("nodek trace", name, kind, k)
    tris = GOLDEN_SCENES[name]()
    s, b, tk, got, want = convert_both(ctx, tris, kind, layout, k)
    wide, r, _ = want
    rays = sq.gen_secondary_rays(tris, 20000, seed=5, tmin=1e-4)
    g = ctx.trace_rays(tk, rays)
    o = ob.nodek_trace(wide, r, tris, rays)
    assert np.array_equal(g["counts"], o["counts"])
    check_hits(g["hit_id"], g["hit_t"], o["hit_id"], o["hit_t"])
    cl = ctx.trace_rays(tk, rays, semantics=sq.TRACE_CLOSEST)
    check_hits(cl["hit_id"], cl["hit_t"], o["hit_id"], o["hit_t"])
    # and the Node1 tree's closest hits
    n1 = ctx.trace_rays(b, rays, semantics=sq.TRACE_CLOSEST)
    diff = np.nonzero(n1["hit_id"] != cl["hit_id"])[0]
    assert len(diff) <= max(2, len(rays) // 2000), len(diff)
    # camera rays: per-pixel counts summed over samples, heatmap bytes
    cam = sq.pinhole_camera(tris, 96, 64, 60.0, 3.0, spp=2, jitter=1, seed=8)
    gc = ctx.trace_camera(tk, cam)
    oc = ob.nodek_trace(wide, r, tris, ob.camera_rays(cam))
    assert np.array_equal(gc["counts"], oc["counts"].reshape(-1, 2).sum(1))
    check_hits(gc["hit_id"], gc["hit_t"], oc["hit_id"], oc["hit_t"])
    cc = ctx.trace_camera(tk, cam, semantics=sq.TRACE_CLOSEST, want_counts=False, want_rgb=False)
    check_hits(cc["hit_id"], cc["hit_t"], oc["hit_id"], oc["hit_t"])
    tk.free(); b.free(); s.free()


@pytest.mark.parametrize("name", ["ico1280", "soup2000", "clus3000"])
def test_uncollapsed_counts_equal_reference(ctx, name, manifest):
    """8-bin recursive SAH trees hold lists of up to 8 nodes: uncollapsed
    Node<8> trees test exactly the reference's boxes."""
    say("nodek reference counts", name)
    tris = GOLDEN_SCENES[name]()
    tr = manifest["traces"][name]
    s = ctx.upload(tris)
    b = ctx.build(s, "recursive")
    tk = ctx.nodek(b, 8, collapse=False)
    cam = sq.reference_camera(tris, tr["w"], tr["h"])
    g = ctx.trace_rays(tk, ob.camera_rays(cam))
    assert hashlib.sha256(g["counts"].tobytes()).hexdigest() == tr["counts_sha256"]
    gc = ctx.trace_camera(tk, cam)
    assert hashlib.sha256(gc["rgb"].tobytes()).hexdigest() == tr["basic_rgb_sha256"]
    tk.free(); b.free(); s.free()


@pytest.mark.parametrize("k", [4, 8])
def test_nodek_large_soup(ctx, k):
    """1M-triangle soup, 63-bit Morton tree: conversion bit-exact; closest
    hits of a jittered camera equal the Node1 tree's; a 20k-ray sample of
    secondary rays equals the oracle, counts included."""
    say("nodek 1M", k)
    tris = sq.gen_soup(1_000_000, 2, 0.01)
    s, b, tk, got, want = convert_both(ctx, tris, "morton63", sq.LAYOUT_REFERENCE, k)
    assert got[0].tobytes() == want[0].tobytes()
    cam = sq.pinhole_camera(tris, 960, 540, 60.0, 3.0, spp=2, jitter=1, seed=2)
    a = ctx.trace_camera(tk, cam, semantics=sq.TRACE_CLOSEST, want_counts=False, want_rgb=False)
    c = ctx.trace_camera(b, cam, semantics=sq.TRACE_CLOSEST, want_counts=False, want_rgb=False)
    assert (a["hit_id"] != c["hit_id"]).sum() <= 5
    rays = sq.gen_secondary_rays(tris, 20000, seed=3, tmin=1e-4)
    g = ctx.trace_rays(tk, rays)
    o = ob.nodek_trace(want[0], want[1], tris, rays)
    assert np.array_equal(g["counts"], o["counts"])
    check_hits(g["hit_id"], g["hit_t"], o["hit_id"], o["hit_t"])
    tk.free(); b.free(); s.free()


@pytest.mark.parametrize("kind", ["morton63", "recursive"])
def test_two_nodes_per_step(ctx, kind):
    """Node<4> steps over two nodes give the one-node results: exhaustive
    counts, hits and heatmaps, closest hits."""
    say("nodek two per step", kind)
    tris = GOLDEN_SCENES["ico69632"]()
    s = ctx.upload(tris)
    b = ctx.build(s, kind)
    tk = ctx.nodek(b, 4)
    cam = sq.pinhole_camera(tris, 200, 150, 60.0, 3.0, spp=4, jitter=1, seed=3)
    for sem in (sq.TRACE_EXHAUSTIVE, sq.TRACE_CLOSEST):
        one = ctx.trace_camera(tk, cam, semantics=sem, lists_per_step=1)
        two = ctx.trace_camera(tk, cam, semantics=sem, lists_per_step=2)
        assert np.array_equal(one["hit_id"], two["hit_id"])
        assert np.array_equal(one["hit_t"].view(np.uint32), two["hit_t"].view(np.uint32))
        if sem == sq.TRACE_EXHAUSTIVE:
            assert np.array_equal(one["counts"], two["counts"])
            assert np.array_equal(one["rgb"], two["rgb"])
    rays = sq.gen_secondary_rays(tris, 20000, seed=4, tmin=1e-4)
    one = ctx.trace_rays(tk, rays, semantics=sq.TRACE_CLOSEST, lists_per_step=1)
    two = ctx.trace_rays(tk, rays, semantics=sq.TRACE_CLOSEST, lists_per_step=2)
    assert np.array_equal(one["hit_id"], two["hit_id"])
    t8 = ctx.nodek(b, 8)
    with pytest.raises(sq.SqError):
        ctx.trace_rays(t8, rays, lists_per_step=2)
    with pytest.raises(sq.SqError):
        ctx.trace_rays(tk, rays, lists_per_step=3)
    t8.free(); tk.free(); b.free(); s.free()


def test_nodek_rejects_unsupported(ctx):
    tris = GOLDEN_SCENES["ico1280"]()
    s = ctx.upload(tris)
    b = ctx.build(s, "recursive")
    for k in (2, 3, 16):
        with pytest.raises(sq.SqError):
            ctx.nodek(b, k)
    tk = ctx.nodek(b, 4)
    cam = sq.pinhole_camera(tris, 32, 32, 60.0, 3.0, spp=1)
    with pytest.raises(sq.SqError):
        ctx.trace_camera(tk, cam, heatmap=sq.HEATMAP_PACKET, width=8, want_hits=False)
    with pytest.raises(sq.SqError):
        ctx.trace_camera(tk, cam, tri_test=1)
    tk.free(); b.free(); s.free()
