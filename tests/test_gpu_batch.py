"""Batched build + trace: every scene's tree and outputs equal the
single-scene calls' (which equal the oracle's), eagerly and through the
captured graph, before and after autotuning, for batches that need one and
two segment sort passes."""
import numpy as np
import pytest

from oracle import binding as ob

import sequoia_amd as sq

from conftest import nodes_equal

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def mixed_scenes():
    """Uneven sizes, including fans (< 8 triangles) and equal-code ranges."""
    out = [
        sq.gen_soup(5, 11, 0.01),
        sq.gen_soup(8, 12, 0.01),
        sq.gen_icosphere(1280, 1, 0.1),
        sq.gen_soup(2000, 7, 0.01),
        sq.gen_clusters(3000, 16, 5),
        sq.gen_soup(3001, 13, 0.05),
        sq.gen_icosphere(20480, 3, 0.1),
        sq.gen_clusters(50000, 64, 21),
        sq.gen_soup(130000, 22, 0.01),
    ]
    # many copies of one triangle: every code equal (midpoint splits)
    out.append(np.repeat(sq.gen_soup(1, 3, 0.01), 700, axis=0))
    return out


def cameras_for(scenes):
    cams = []
    for i, t in enumerate(scenes):
        w, h = [(64, 48), (100, 60), (37, 23), (160, 90)][i % 4]
        cams.append(sq.pinhole_camera(t, w, h, 60.0, 3.0, spp=1 + (i % 3), jitter=i % 2, seed=100 + i))
    return cams


@pytest.fixture(scope="module")
def mixed(ctx):
    scenes = mixed_scenes()
    b = sq.Batch(ctx, scenes)
    yield scenes, b
    b.free()


def plain_tree(ctx, tris, kind, layout=sq.LAYOUT_REFERENCE):
    s = ctx.upload(tris)
    bv = ctx.build(s, kind, layout=layout)
    nodes = bv.download()
    depth = bv.info()[2]
    bv.free()
    s.free()
    return nodes, depth


@pytest.mark.parametrize("kind", ["morton63", "morton", "morton2"])
def test_batch_trees_equal_plain(ctx, mixed, kind):
    scenes, b = mixed
    say("batch build", kind, len(scenes))
    b.build(kind)
    for i, t in enumerate(scenes):
        got = b.bvh(i).download()
        want, depth = plain_tree(ctx, t, kind)
        assert len(got) == len(want)
        assert len(nodes_equal(got, want)) == 0, (kind, i, nodes_equal(got, want)[:5])
        assert b.bvh(i).info()[2] == depth, (kind, i)


def test_batch_fixed_layout_equals_plain(ctx, mixed):
    scenes, b = mixed
    b.build("morton63", layout=sq.LAYOUT_FIXED)
    for i, t in enumerate(scenes):
        want, _ = plain_tree(ctx, t, "morton63", layout=sq.LAYOUT_FIXED)
        assert len(nodes_equal(b.bvh(i).download(), want)) == 0, i
    b.build("morton63")


def test_batch_tree_matches_oracle(ctx, mixed):
    """One scene of the batch straight against the oracle's Morton tree."""
    scenes, b = mixed
    b.build("morton63")
    i = 4
    nodes, root, _, _ = ob.build_morton(scenes[i], 21)
    got = b.bvh(i).download()
    assert root == 0
    assert len(nodes_equal(got, nodes)) == 0


def plain_trace(ctx, tris, cam, kind, semantics):
    s = ctx.upload(tris)
    bv = ctx.build(s, kind)
    r = ctx.trace_camera(bv, cam, semantics=semantics)
    bv.free()
    s.free()
    return r


def assert_same(got, want, tag):
    for k in ("hit_id", "hit_t", "counts", "rgb"):
        g, w = got[k], want[k]
        assert np.array_equal(np.ascontiguousarray(g).view(np.uint8), np.ascontiguousarray(w).view(np.uint8)), (tag, k)


@pytest.mark.parametrize("semantics,kernel,lists", [(sq.TRACE_EXHAUSTIVE, sq.KERNEL_RAYS, 0),
                                                    (sq.TRACE_EXHAUSTIVE, sq.KERNEL_RAYS, 4),
                                                    (sq.TRACE_EXHAUSTIVE, sq.KERNEL_PACKETS, 0),
                                                    (sq.TRACE_CLOSEST, sq.KERNEL_AUTO, 0)])
def test_batch_trace_equals_plain(ctx, mixed, semantics, kernel, lists):
    scenes, b = mixed
    cams = cameras_for(scenes)
    say("batch trace", semantics, kernel, lists)
    b.build("morton63")
    b.set_cameras(cams, semantics=semantics, kernel=kernel, lists_per_step=lists)
    b.trace()
    for i, t in enumerate(scenes):
        assert_same(b.download(i), plain_trace(ctx, t, cams[i], "morton63", semantics), (semantics, i))


def test_batch_trace_matches_oracle(ctx, mixed):
    scenes, b = mixed
    cams = cameras_for(scenes)
    b.build("morton63")
    b.set_cameras(cams)
    b.trace()
    i = 3
    nodes = b.bvh(i).download()
    want = ob.trace_camera(nodes, 0, scenes[i], cams[i])
    got = b.download(i)
    assert np.array_equal(got["counts"], want["counts"])
    assert np.array_equal(got["hit_id"], want["hit_id"])


def test_batch_graph_replays(ctx, mixed):
    scenes, b = mixed
    cams = cameras_for(scenes)
    b.set_cameras(cams)
    b.build("morton63")
    b.trace()
    eager = [b.download(i) for i in range(len(scenes))]
    for rep in range(3):
        say("graph replay", rep)
        b.run("morton63")
        for i in range(len(scenes)):
            assert_same(b.download(i), eager[i], ("replay", rep, i))
        # eager work between replays
        b.build("morton")
        b.build("morton63")


def test_batch_autotune_keeps_results(ctx, mixed):
    scenes, b = mixed
    cams = cameras_for(scenes)
    b.set_cameras(cams)
    b.build("morton63")
    b.trace()
    before = [b.download(i) for i in range(len(scenes))]
    trees = [b.bvh(i).download() for i in range(len(scenes))]
    r = b.autotune("morton63", repeats=1)
    say("autotune", r)
    assert r["candidates_tried"] == 33 and r["kernel"] in (sq.KERNEL_RAYS, sq.KERNEL_PACKETS)
    assert 4 <= r["radix_digit_bits"] <= 8 and r["stack_slots"] in (8, 16, 32) and r["block_size"] in (64, 128)
    assert r["blocks_per_cu"] in (0, 8, 12, 24)
    assert r["lists_per_step"] in ((2, 4) if r["kernel"] == sq.KERNEL_RAYS else (0,))
    b.run("morton63")
    for i in range(len(scenes)):
        assert len(nodes_equal(b.bvh(i).download(), trees[i])) == 0
        assert_same(b.download(i), before[i], ("tuned", i))


def test_many_scenes_two_segment_passes(ctx):
    """300 scenes need two 8-bit segment passes (and 3 passes at 6 bits)."""
    rng = np.random.default_rng(5)
    scenes = [sq.gen_soup(int(n), 1000 + i, 0.02) for i, n in enumerate(rng.integers(1, 60, 300))]
    b = sq.Batch(ctx, scenes)
    for digit in (8, 6):
        say("many scenes, digit", digit)
        b.build("morton63", digit_bits=digit)
        for i in range(0, 300, 7):
            want, _ = plain_tree(ctx, scenes[i], "morton63")
            assert len(nodes_equal(b.bvh(i).download(), want)) == 0, (digit, i)
    b.free()


def test_batch_rejects_bad_input(ctx):
    with pytest.raises(sq.SqError):
        sq.Batch(ctx, [sq.gen_soup(10, 1, 0.01), np.zeros((0, 9), np.float32)])
    b = sq.Batch(ctx, [sq.gen_soup(100, 1, 0.01)])
    with pytest.raises(sq.SqError):
        b.build("grid")  # batches build with the Morton and binned-SAH (recursive / horizontal) builders
    with pytest.raises(sq.SqError):
        b.build("recursive", num_bins=65)
    with pytest.raises(sq.SqError):
        b.trace()
    b.free()


def test_batch_large_scenes(ctx):
    """Scenes of 0.9M-2M triangles in one batch (two segment digits' worth of
    triangles, code ranges split by the top phase over many levels) equal
    their single-scene builds and traces."""
    scenes = [sq.gen_icosphere(1837637, 6023, 0.1), sq.gen_soup(2_000_000, 31, 0.01),
              sq.gen_icosphere(918884, 6003, 0.1), sq.gen_soup(3000, 32, 0.02)]
    cams = [sq.pinhole_camera(t, 320, 180, 60.0, 3.0, spp=2, jitter=1, seed=40 + i) for i, t in enumerate(scenes)]
    b = sq.Batch(ctx, scenes)
    say("large batch build")
    b.build("morton63")
    b.set_cameras(cams)
    b.trace()
    for i, t in enumerate(scenes):
        want, depth = plain_tree(ctx, t, "morton63")
        assert len(nodes_equal(b.bvh(i).download(), want)) == 0, i
        assert b.bvh(i).info()[2] == depth, i
        assert_same(b.download(i), plain_trace(ctx, t, cams[i], "morton63", sq.TRACE_EXHAUSTIVE), ("large", i))
    b.free()


def plain_sah_tree(ctx, tris, kind, nb):
    s = ctx.upload(tris)
    bv = ctx.build(s, kind, num_bins=nb)
    nodes, root = bv.download(), bv.root
    bv.free()
    s.free()
    return nodes, root


@pytest.mark.parametrize("kind,nb", [("recursive", 8), ("recursive", 16), ("recursive", 32), ("horizontal", 8)])
def test_batch_sah_trees_equal_plain(ctx, mixed, kind, nb):
    """Binned SAH over every scene of the batch at once: each scene's Node1[]
    and root equal its single-scene build with the same bin count."""
    scenes, b = mixed
    say("batch sah", kind, nb)
    b.build(kind, num_bins=nb)
    for i, t in enumerate(scenes):
        want, root = plain_sah_tree(ctx, t, kind, nb)
        got = b.bvh(i)
        assert got.root == root, (i, got.root, root)
        nodes = got.download()
        assert len(nodes) == len(want)
        assert len(nodes_equal(nodes, want)) == 0, (kind, nb, i, nodes_equal(nodes, want)[:5])
    b.build("morton63")


def test_batch_sah_matches_oracle(ctx, mixed):
    scenes, b = mixed
    b.build("recursive", num_bins=16)
    for i in (2, 4, 7):
        want, root = ob.build_sah(scenes[i], 16)
        assert b.bvh(i).root == root
        assert len(nodes_equal(b.bvh(i).download(), want)) == 0, i
    b.build("morton63")


@pytest.mark.parametrize("semantics", [sq.TRACE_EXHAUSTIVE, sq.TRACE_CLOSEST])
def test_batch_sah_trace_equals_plain(ctx, mixed, semantics):
    """Traces of the batched SAH trees (eagerly and through run(): eager
    build, graph trace), then back to Morton trees in the same batch."""
    scenes, b = mixed
    cams = cameras_for(scenes)
    b.set_cameras(cams, semantics=semantics)
    for rep in range(2):
        if rep == 0:
            b.build("recursive", num_bins=16)
            b.trace()
        else:
            b.run("recursive", num_bins=16)
        for i, t in enumerate(scenes):
            s = ctx.upload(t)
            bv = ctx.build(s, "recursive", num_bins=16)
            want = ctx.trace_camera(bv, cams[i], semantics=semantics)
            bv.free()
            s.free()
            assert_same(b.download(i), want, ("sah", rep, i))
    b.run("morton63")
    for i, t in enumerate(scenes):
        assert_same(b.download(i), plain_trace(ctx, t, cams[i], "morton63", semantics), ("back to morton", i))


def test_batch_sah_autotune_bins(ctx, mixed):
    scenes, b = mixed
    cams = cameras_for(scenes)
    b.set_cameras(cams)
    r = b.autotune("recursive", repeats=1)
    say("sah autotune", r)
    assert r["num_bins"] in (8, 16, 32) and r["candidates_tried"] == 3 + 18
    b.run("recursive", num_bins=r["num_bins"])
    for i, t in enumerate(scenes):
        want, root = plain_sah_tree(ctx, t, "recursive", r["num_bins"])
        assert len(nodes_equal(b.bvh(i).download(), want)) == 0, i
    b.build("morton63")
