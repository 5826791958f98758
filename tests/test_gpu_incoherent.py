"""Incoherent workload: 10M triangles in 64 Gaussian clusters, grid and
binned-SAH trees bit-exact against the oracle, and cosine-weighted secondary
rays (tmin 1e-4) traced on them, checked against the oracle's traversal on a
sample of the rays. Also a 3M-triangle scene whose grid cells and SAH ranges
exceed the single-workgroup task size, for every partition kind."""
import numpy as np
import pytest

from conftest import nodes_equal, tree_depth
from oracle import binding as ob
from test_gpu_trace import check_hits

import sequoia_amd as sq

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("kind", ["recursive", "grid", "grid_sse"])
def test_large_tasks_bit_exact(ctx, kind):
    """Cells of ~47k triangles and SAH ranges up to 3M go through the
    chunked (many-workgroup) split; Hoare and stable partitions both."""
    say("large tasks", kind)
    tris = sq.gen_clusters(3_000_000, 64, 17)
    scene = ctx.upload(tris)
    bvh = ctx.build(scene, kind, num_bins=16)
    got = bvh.download()
    if kind == "recursive":
        want, root = ob.build_sah(tris, 16, ob.PART_HOARE, ob.SA_SCALAR)
    elif kind == "grid":
        want, root = ob.build_grid(tris, 4, 16, ob.PART_HOARE, ob.SA_SCALAR, ob.CELLBOX_SCENE)
    else:
        want, root = ob.build_grid(tris, 4, 16, ob.PART_STABLE, ob.SA_SSE, ob.CELLBOX_TIGHT)
    assert len(nodes_equal(got, want)) == 0
    # the reported depth bounds the traversal stack: never below the tree's
    assert bvh.info()[2] >= tree_depth(got, root)


@pytest.fixture(scope="module")
def clusters10m(ctx):
    say("generating 10M clusters")
    tris = sq.gen_clusters(10_000_000, 64, 3)
    scene = ctx.upload(tris)
    yield tris, scene
    scene.free()


@pytest.mark.parametrize("kind,nb", [("grid", 8), ("recursive", 8)])
def test_clusters_10m_tree_and_secondary_rays(ctx, clusters10m, kind, nb):
    tris, scene = clusters10m
    say("10M", kind, nb)
    bvh = ctx.build(scene, kind, num_bins=nb)
    got = bvh.download()
    if kind == "grid":
        want, root = ob.build_grid(tris, 4, nb, ob.PART_HOARE, ob.SA_SCALAR, ob.CELLBOX_SCENE)
    else:
        want, root = ob.build_sah(tris, nb, ob.PART_HOARE, ob.SA_SCALAR)
    assert bvh.root == root
    assert len(nodes_equal(got, want)) == 0
    assert bvh.info()[2] >= tree_depth(got, root)
    del want
    say("10M trees equal; secondary rays")
    rays = sq.gen_secondary_rays(tris, 1_000_000, seed=3, tmin=1e-4)
    cl = ctx.trace_rays(bvh, rays, semantics=sq.TRACE_CLOSEST)
    sample = np.random.default_rng(11).choice(len(rays), 1500, replace=False)
    ex = ctx.trace_rays(bvh, rays[sample])
    ref = ob.trace_rays(got, bvh.root, tris, rays[sample])
    assert np.array_equal(ex["counts"], ref["counts"])
    check_hits(ex["hit_id"], ex["hit_t"], ref["hit_id"], ref["hit_t"])
    check_hits(cl["hit_id"][sample], cl["hit_t"][sample], ref["hit_id"], ref["hit_t"])
    # secondary rays leave a surface: almost all of them hit something
    assert (cl["hit_id"] != sq.MISS).mean() > 0.5
    bvh.free()
