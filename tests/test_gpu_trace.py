"""GPU traversal against the oracle: per-ray node-test counts and heatmaps
bit-exact on identical trees; closest hits (id bit-exact, t bit-exact)."""
import hashlib

import numpy as np
import pytest

from oracle import binding as ob
from scenes import GOLDEN_SCENES

import sequoia_amd as sq

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_hits(got_id, got_t, want_id, want_t):
    same = got_id == want_id
    # ids may differ only where two hits lie within 1 ulp in t
    bad = ~same
    if bad.any():
        gt, wt = got_t[bad], want_t[bad]
        ulp = np.abs(gt.view(np.int32).astype(np.int64) - wt.view(np.int32).astype(np.int64))
        assert (ulp <= 1).all(), (np.nonzero(bad)[0][:5], got_id[bad][:5], want_id[bad][:5])
    hit = same & (want_id != sq.MISS)
    assert np.array_equal(got_t[hit].view(np.uint32), want_t[hit].view(np.uint32))


@pytest.mark.parametrize("kernel", [sq.KERNEL_RAYS, sq.KERNEL_PACKETS])
@pytest.mark.parametrize("name", ["ico1280", "soup2000", "clus3000", "ico69632"])
def test_reference_camera_on_reference_tree(ctx, name, manifest, kernel):
    say("reference camera", name, kernel)
    tris = GOLDEN_SCENES[name]()
    tr = manifest["traces"][name]
    nodes, root = ob.build_sah(tris, 8)       # equals the reference's recursive tree (pinned)
    scene = ctx.upload(tris)
    bvh = ctx.bvh_from_host(scene, nodes, root)
    cam = sq.reference_camera(tris, tr["w"], tr["h"])
    got = ctx.trace_camera(bvh, cam, kernel=kernel)
    assert sha(got["counts"]) == tr["counts_sha256"]
    assert sha(got["rgb"]) == tr["basic_rgb_sha256"]
    want = ob.trace_camera(nodes, root, tris, cam)
    check_hits(got["hit_id"], got["hit_t"], want["hit_id"], want["hit_t"])


@pytest.mark.parametrize("kernel", [sq.KERNEL_RAYS, sq.KERNEL_PACKETS])
@pytest.mark.parametrize("kind", ["morton63", "morton"])
@pytest.mark.parametrize("name", ["ico1280", "clus3000", "ico69632"])
def test_jittered_pinhole_on_gpu_tree(ctx, name, kind, kernel):
    say("jittered pinhole", name, kind, kernel)
    tris = GOLDEN_SCENES[name]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, kind)
    nodes = bvh.download()
    cam = sq.pinhole_camera(tris, 200, 120, 60.0, 3.0, spp=4, jitter=1, seed=9)
    got = ctx.trace_camera(bvh, cam, kernel=kernel)
    want = ob.trace_camera(nodes, bvh.root, tris, cam)
    assert np.array_equal(got["counts"], want["counts"])
    assert np.array_equal(got["rgb"], want["rgb"])
    check_hits(got["hit_id"], got["hit_t"], want["hit_id"], want["hit_t"])


@pytest.mark.parametrize("lists", [0, 4])
@pytest.mark.parametrize("spp", [1, 4, 5])
def test_ragged_image_tiles(ctx, spp, lists):
    """Image edges that cut camera tiles (203 x 117: 4x4 tiles for 4+
    samples, 8x8 below, odd and even tile-row counts for the rows dealt from
    the middle outwards) against the oracle, both step widths."""
    say("ragged image", spp, lists)
    tris = GOLDEN_SCENES["clus3000"]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    nodes = bvh.download()
    cam = sq.pinhole_camera(tris, 203, 117, 60.0, 3.0, spp=spp, jitter=1, seed=11)
    got = ctx.trace_camera(bvh, cam, lists_per_step=lists)
    want = ob.trace_camera(nodes, bvh.root, tris, cam)
    assert np.array_equal(got["counts"], want["counts"])
    assert np.array_equal(got["rgb"], want["rgb"])
    check_hits(got["hit_id"], got["hit_t"], want["hit_id"], want["hit_t"])
    cl = ctx.trace_camera(bvh, cam, semantics=sq.TRACE_CLOSEST)
    check_hits(cl["hit_id"], cl["hit_t"], want["hit_id"], want["hit_t"])


@pytest.mark.parametrize("kind", ["morton63", "recursive"])
def test_closest_steps_with_finite_tmax(ctx, kind):
    """Camera rays whose tmax cuts through the scene: the two-list closest
    step (Node1) and the two-Node2 closest step equal the exhaustive hits and
    the oracle's, with culling and tmax acting together."""
    say("closest tmax", kind)
    tris = GOLDEN_SCENES["clus3000"]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, kind)
    nodes = bvh.download()
    cam = sq.pinhole_camera(tris, 160, 96, 60.0, 3.0, spp=4, jitter=1, seed=21)
    cam.tmax = 3.0  # about the distance to the scene-box centre
    want = ob.trace_camera(nodes, bvh.root, tris, cam)
    ex = ctx.trace_camera(bvh, cam)
    assert np.array_equal(ex["counts"], want["counts"])
    check_hits(ex["hit_id"], ex["hit_t"], want["hit_id"], want["hit_t"])
    assert (want["hit_id"] != sq.MISS).any() and (want["hit_id"] == sq.MISS).any()
    cl = ctx.trace_camera(bvh, cam, semantics=sq.TRACE_CLOSEST)
    check_hits(cl["hit_id"], cl["hit_t"], want["hit_id"], want["hit_t"])
    t2 = ctx.node2(bvh)
    c2 = ctx.trace_camera(t2, cam, semantics=sq.TRACE_CLOSEST)
    check_hits(c2["hit_id"], c2["hit_t"], want["hit_id"], want["hit_t"])
    t2.free()


@pytest.mark.parametrize("name", ["ico69632", "clus3000"])
def test_closest_mode_matches_exhaustive_hits(ctx, name):
    say("closest", name)
    tris = GOLDEN_SCENES[name]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63", layout=sq.LAYOUT_FIXED)
    cam = sq.pinhole_camera(tris, 256, 160, 60.0, 3.0, spp=2, jitter=1, seed=3)
    ex = ctx.trace_camera(bvh, cam)
    cl = ctx.trace_camera(bvh, cam, semantics=sq.TRACE_CLOSEST)
    check_hits(cl["hit_id"], cl["hit_t"], ex["hit_id"], ex["hit_t"])
    assert cl["counts"].sum() <= ex["counts"].sum()


def test_ray_list_secondary_rays(ctx):
    say("secondary rays")
    tris = sq.gen_clusters(20000, 64, 3)
    rays = sq.gen_secondary_rays(tris, 20000, seed=3, tmin=1e-4)
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    nodes = bvh.download()
    want = ob.trace_rays(nodes, bvh.root, tris, rays)
    for lists in (0, 4):  # two child lists per step by default for ray lists; four on request
        got = ctx.trace_rays(bvh, rays, lists_per_step=lists)
        assert np.array_equal(got["counts"], want["counts"]), lists
        check_hits(got["hit_id"], got["hit_t"], want["hit_id"], want["hit_t"])
    cl = ctx.trace_rays(bvh, rays, semantics=sq.TRACE_CLOSEST)
    check_hits(cl["hit_id"], cl["hit_t"], want["hit_id"], want["hit_t"])


def test_ray_list_tile_sizes(ctx):
    """Ray lists of 1 to 1.1M rays: tiles of 64, 128, 192 and 256 rays (one
    tile per wave for short lists), ragged last tiles, against the oracle."""
    say("ray list tile sizes")
    tris = GOLDEN_SCENES["clus3000"]()
    rays = sq.gen_secondary_rays(tris, 1_100_000, seed=5, tmin=1e-4)
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    nodes = bvh.download()
    want = ob.trace_rays(nodes, bvh.root, tris, rays)
    for n in (1, 65, 300_000, 700_000, 1_100_000):
        got = ctx.trace_rays(bvh, rays[:n])
        assert np.array_equal(got["counts"], want["counts"][:n]), n
        check_hits(got["hit_id"], got["hit_t"], want["hit_id"][:n], want["hit_t"][:n])


def test_camera_rays_equal_uploaded_oracle_rays(ctx):
    """GPU-generated camera rays produce the same results as the oracle's ray bits."""
    tris = GOLDEN_SCENES["ico1280"]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    cam = sq.pinhole_camera(tris, 64, 48, 60.0, 3.0, spp=2, jitter=1, seed=1)
    a = ctx.trace_camera(bvh, cam, want_rgb=False)
    b = ctx.trace_rays(bvh, ob.camera_rays(cam))
    assert np.array_equal(a["hit_id"], b["hit_id"])
    assert np.array_equal(a["hit_t"].view(np.uint32), b["hit_t"].view(np.uint32))
    assert np.array_equal(a["counts"], b["counts"].reshape(-1, 2).sum(1))


def test_small_stack_spills_correctly(ctx):
    say("spill")
    tris = GOLDEN_SCENES["ico69632"]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    cam = sq.pinhole_camera(tris, 128, 96, 60.0, 3.0)
    b = ctx.trace_camera(bvh, cam, stack_slots=32)
    for lists in (2, 4):  # spills under both step widths
        a = ctx.trace_camera(bvh, cam, stack_slots=2, lists_per_step=lists)
        assert np.array_equal(a["counts"], b["counts"]) and np.array_equal(a["hit_id"], b["hit_id"]), lists
    # closest hits: camera rays (two lists per step, up to four pushes) and a
    # ray list (one list per step) through a two-entry LDS stack
    cb = ctx.trace_camera(bvh, cam, semantics=sq.TRACE_CLOSEST, stack_slots=32)
    ca = ctx.trace_camera(bvh, cam, semantics=sq.TRACE_CLOSEST, stack_slots=2)
    for k in ("counts", "hit_id", "hit_t"):
        assert np.array_equal(ca[k], cb[k]), k
    check_hits(ca["hit_id"], ca["hit_t"], b["hit_id"], b["hit_t"])
    rays = sq.gen_secondary_rays(tris, 20000, seed=7, tmin=1e-4)
    rb = ctx.trace_rays(bvh, rays, semantics=sq.TRACE_CLOSEST, stack_slots=32)
    ra = ctx.trace_rays(bvh, rays, semantics=sq.TRACE_CLOSEST, stack_slots=2)
    for k in ("counts", "hit_id", "hit_t"):
        assert np.array_equal(ra[k], rb[k]), k


def test_lists_per_step_rejects_other_values(ctx):
    tris = GOLDEN_SCENES["ico1280"]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    cam = sq.pinhole_camera(tris, 16, 16, 60.0, 3.0)
    with pytest.raises(sq.SqError):
        ctx.trace_camera(bvh, cam, lists_per_step=3)


@pytest.mark.parametrize("name", ["ico1280", "soup2000", "clus3000", "ico69632"])
def test_packet_and_bundle_heatmaps_match_reference(ctx, name, manifest):
    say("packet/bundle", name)
    tris = GOLDEN_SCENES[name]()
    tr = manifest["traces"][name]
    nodes, root = ob.build_sah(tris, 8)
    scene = ctx.upload(tris)
    bvh = ctx.bvh_from_host(scene, nodes, root)
    cam = sq.reference_camera(tris, tr["w"], tr["h"])
    for K in (1, 2, 4, 8):
        got = ctx.trace_camera(bvh, cam, heatmap=sq.HEATMAP_PACKET, width=K, want_hits=False)
        assert sha(got["rgb"]) == tr["packet%d_rgb_sha256" % K], K
    for K in (16, 64):
        got = ctx.trace_camera(bvh, cam, heatmap=sq.HEATMAP_BUNDLE, width=K, want_hits=False)
        assert sha(got["rgb"]) == tr["bundle%d_rgb_sha256" % K], K


def test_packet_and_bundle_on_gpu_morton_tree(ctx):
    tris = GOLDEN_SCENES["ico69632"]()
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, "morton63")
    nodes = bvh.download()
    cam = sq.pinhole_camera(tris, 320, 200, 60.0, 3.0)
    for K in (1, 8):
        got = ctx.trace_camera(bvh, cam, heatmap=sq.HEATMAP_PACKET, width=K, want_hits=False)
        assert np.array_equal(got["rgb"], ob.trace_packet(nodes, bvh.root, tris, cam, K))
    for K in (16, 64):
        got = ctx.trace_camera(bvh, cam, heatmap=sq.HEATMAP_BUNDLE, width=K, want_hits=False)
        assert np.array_equal(got["rgb"], ob.trace_bundle(nodes, bvh.root, tris, cam, K))


@pytest.mark.parametrize("kind", ["morton63", "recursive"])
def test_packets_equal_rays_at_full_size(ctx, kind):
    """The bench frame (1M soup, 1920x1080x4 jittered): the wave-packet
    kernel's per-pixel counts, heatmap bytes and per-ray hits equal the
    per-lane kernel's bit for bit (two and four lists per step), on a Morton
    and on a SAH tree."""
    say("packets vs rays", kind)
    tris = sq.gen_soup(1_000_000, 2, 0.01)
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, kind)
    cam = sq.pinhole_camera(tris, 1920, 1080, 60.0, 3.0, spp=4, jitter=1, seed=2)
    a = ctx.trace_camera(bvh, cam, kernel=sq.KERNEL_RAYS)
    b = ctx.trace_camera(bvh, cam, kernel=sq.KERNEL_PACKETS)
    c = ctx.trace_camera(bvh, cam, kernel=sq.KERNEL_RAYS, lists_per_step=4)
    for k in ("hit_id", "hit_t", "counts", "rgb"):
        assert np.array_equal(np.ascontiguousarray(a[k]).view(np.uint8), np.ascontiguousarray(b[k]).view(np.uint8)), k
        assert np.array_equal(np.ascontiguousarray(a[k]).view(np.uint8), np.ascontiguousarray(c[k]).view(np.uint8)), k


def test_wide_and_narrow_fetches_agree(ctx):
    """Traversal launches address node and triangle arrays below 4 GiB with
    32-bit offsets; a context created with SEQUOIA_TEST_WIDE_FETCH keeps the
    64-bit form every launch over larger arrays takes. Both give the same
    bytes: single camera (exhaustive and closest), ray lists and a batch."""
    import os
    tris = sq.gen_soup(200_000, 5, 0.01)
    cam = sq.pinhole_camera(tris, 320, 180, 60.0, 3.0, spp=4, jitter=1, seed=5)
    rays = sq.gen_secondary_rays(tris, 20000, seed=3, tmin=1e-4)
    os.environ["SEQUOIA_TEST_WIDE_FETCH"] = "1"
    try:
        wide = sq.Context(0)
    finally:
        del os.environ["SEQUOIA_TEST_WIDE_FETCH"]
    scenes = [sq.gen_icosphere(1280, 1, 0.1), sq.gen_soup(3000, 7, 0.01), tris]
    cams = [sq.pinhole_camera(t, 64, 48, 60.0, 3.0, spp=2, jitter=1, seed=9) for t in scenes]
    try:
        outs = []
        for c in (ctx, wide):
            s = c.upload(tris)
            b = c.build(s, "morton63")
            o = {"ex": c.trace_camera(b, cam), "cl": c.trace_camera(b, cam, semantics=sq.TRACE_CLOSEST),
                 "rays": c.trace_rays(b, rays)}
            b.free()
            s.free()
            bt = sq.Batch(c, scenes)
            bt.set_cameras(cams)
            bt.run("morton63")
            o["batch"] = [bt.download(i) for i in range(len(scenes))]
            bt.free()
            outs.append(o)
        n, w = outs
        for k in ("ex", "cl", "rays"):
            for f in ("hit_id", "hit_t", "counts"):
                assert np.array_equal(np.ascontiguousarray(n[k][f]).view(np.uint8),
                                      np.ascontiguousarray(w[k][f]).view(np.uint8)), (k, f)
        for i in range(len(scenes)):
            for f in ("hit_id", "hit_t", "counts", "rgb"):
                assert np.array_equal(np.ascontiguousarray(n["batch"][i][f]).view(np.uint8),
                                      np.ascontiguousarray(w["batch"][i][f]).view(np.uint8)), (i, f)
    finally:
        wide.close()
