"""Precomputed triangles on the GPU: the device precompute writes the same
records as the CPU restatement, and traces with tri_test=1 give the
restatement's hits and t bit for bit (ray lists and cameras, exhaustive and
closest) with box-test counts unchanged from the reference test's."""
import numpy as np
import pytest

from oracle import binding as ob
from scenes import GOLDEN_SCENES
from test_gpu_trace import check_hits

import sequoia_amd as sq

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("name", ["ico1280", "clus3000", "soup2000", "ico69632"])
def test_plucker_precompute_and_trace(ctx, name):
    say("plucker", name)
    tris = GOLDEN_SCENES[name]()
    if name == "soup2000":  # add degenerate (collinear and repeated-vertex) triangles
        bad = np.array([[0, 0, 0, 1, 1, 1, 2, 2, 2], [0.5, 0.5, 0.5] * 3], np.float32)
        tris = np.concatenate([tris, bad])
    s = ctx.upload(tris)
    b = ctx.build(s, "morton63")
    got = ctx.plucker(b)
    want = ob.plucker_precompute(tris)
    assert got.tobytes() == want.tobytes()
    nodes = b.download()
    rays = sq.gen_secondary_rays(tris[:len(GOLDEN_SCENES[name]())], 20000, seed=7, tmin=1e-4)
    g = ctx.trace_rays(b, rays, tri_test=1)
    o = ob.trace_rays_plucker(nodes, b.root, want, rays)
    assert np.array_equal(g["counts"], o["counts"])
    assert np.array_equal(g["hit_id"], o["hit_id"])
    assert np.array_equal(g["hit_t"].view(np.uint32), o["hit_t"].view(np.uint32))
    mt = ctx.trace_rays(b, rays)
    assert np.array_equal(mt["counts"], g["counts"])
    cl = ctx.trace_rays(b, rays, semantics=sq.TRACE_CLOSEST, tri_test=1)
    check_hits(cl["hit_id"], cl["hit_t"], o["hit_id"], o["hit_t"])
    cam = sq.pinhole_camera(tris, 96, 64, 60.0, 3.0, spp=2, jitter=1, seed=3)
    gc = ctx.trace_camera(b, cam, tri_test=1)
    oc = ob.trace_rays_plucker(nodes, b.root, want, ob.camera_rays(cam))
    assert np.array_equal(gc["hit_id"], oc["hit_id"])
    assert np.array_equal(gc["counts"], oc["counts"].reshape(-1, 2).sum(1))
    b.free()
    s.free()


def test_plucker_errors(ctx):
    tris = GOLDEN_SCENES["ico1280"]()
    s = ctx.upload(tris)
    b = ctx.build(s, "morton63")
    rays = sq.gen_secondary_rays(tris, 100, seed=1, tmin=1e-4)
    with pytest.raises(sq.SqError):
        ctx.trace_rays(b, rays, tri_test=1)  # not precomputed yet
    ctx.plucker(b)
    with pytest.raises(sq.SqError):
        ctx.trace_rays(b, rays, tri_test=2)
    t2 = ctx.node2(b)
    with pytest.raises(sq.SqError):
        ctx.trace_rays(t2, rays, tri_test=1)
    t2.free()
    b.free()
    s.free()
