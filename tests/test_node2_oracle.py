"""The Node2 restatement (oracle/node2.cpp) pinned on the CPU: decoded child
boxes contain every triangle below them and are the tightest the 8-bit grid
allows, every reachable triangle lands in exactly one leaf group, and the
closest hits through a Node2 tree equal those through its Node1 tree (where
they differ, only a grazing hit the Node1 slab test rounded away may appear,
and it must be a valid hit of the reference triangle test)."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as ob
from scenes import GOLDEN_SCENES

import sequoia_amd as sq


def trees():
    ico = GOLDEN_SCENES["ico1280"]()
    clus = GOLDEN_SCENES["clus3000"]()
    soup = GOLDEN_SCENES["soup2000"]()
    out = []
    nodes, root = ob.build_sah(ico, 8)
    out.append(("ico1280-sah", ico, nodes, root))
    nodes, root, _, _ = ob.build_morton(clus, 21)
    out.append(("clus3000-morton63", clus, nodes, root))
    nodes, root, _, _ = ob.build_morton(clus, 21, layout=ob.LAYOUT_FIXED)
    out.append(("clus3000-morton63-fixed", clus, nodes, root))
    nodes, root = ob.build_grid(soup, 4, 8, ob.PART_STABLE, ob.SA_SSE, ob.CELLBOX_TIGHT)
    out.append(("soup2000-grid_sse", soup, nodes, root))
    return out


TREES = trees()


def tri_boxes(t):
    v = t.reshape(-1, 3, 3)
    return v.min(1), v.max(1)


def reachable_tris(nodes, root):
    seen, st = [], [(root, 1)]
    while st:
        f, c = st.pop()
        for i in range(f, f + c):
            n = nodes[i]
            if n["flags"] & 1:
                seen.extend(range(int(n["child"]), int(n["child"]) + int(n["count"])))
            else:
                st.append((int(n["child"]), int(n["count"])))
    return np.array(sorted(seen), np.uint32)


@pytest.mark.parametrize("name,tris,nodes,root", TREES, ids=[t[0] for t in TREES])
def test_node2_boxes_contain_subtrees(name, tris, nodes, root):
    n2, r2, tris2, perm = ob.node2_convert(nodes, root, tris)
    amn, amx, bmn, bmx = ob.node2_decode(n2)
    lo, hi = tri_boxes(tris2)
    sub = {}

    def box_of(i, side):
        w0, w1 = int(n2[i]["w0"]), int(n2[i]["w1"])
        child = (w0 if side == 0 else w1) & 0xFFFFFF
        leaf = (w1 >> 27) & 1 if side == 0 else (w1 >> 31) & 1
        cnt = (w1 >> 24) & 7 if side == 0 else (w1 >> 28) & 7
        if leaf:
            assert 1 <= cnt <= 7
            return lo[child:child + cnt].min(0), hi[child:child + cnt].max(0)
        return subtree(child)

    def subtree(i):
        if i not in sub:
            a, b = box_of(i, 0), box_of(i, 1)
            sub[i] = (np.minimum(a[0], b[0]), np.maximum(a[1], b[1]))
        return sub[i]

    for i in range(len(n2) - 1, -1, -1):
        for side, (dmn, dmx) in enumerate(((amn, amx), (bmn, bmx))):
            smn, smx = box_of(i, side)
            assert (dmn[i] <= smn).all() and (dmx[i] >= smx).all(), (name, i, side)
    # tightness: one grid step inwards no longer contains the Node1 box
    step = np.array([np.float32(2.0) ** (int(w >> 24) - 127) for w in n2["w0"]], np.float32)
    q = n2["q"].astype(np.int32)
    for k in range(3):
        for s, (dmn, dmx) in enumerate(((amn, amx), (bmn, bmx))):
            qlo, qhi = q[:, 4 * k + 2 * s], q[:, 4 * k + 2 * s + 1]
            up = (n2["c"][:, k] + ((qlo + 1 - 128).astype(np.float32) * step)).astype(np.float32)
            dn = (n2["c"][:, k] + ((qhi - 1 - 128).astype(np.float32) * step)).astype(np.float32)
            assert (dmn[:, k] <= up).all() and (dmx[:, k] >= dn).all()
    # every reachable triangle, each as often as the Node1 tree reaches it
    # (a reference-layout quirk can reach a subtree through two paths)
    assert np.array_equal(np.sort(perm), reachable_tris(nodes, root))


def test_node2_quantiser_edge_cases():
    # a scene of one triangle, a fan of seven, a degenerate (all-equal) scene
    base = sq.gen_soup(7, 3, 0.01)
    for tris in (base[:1], base, np.repeat(base[:1], 40, axis=0)):
        nodes, root = ob.build_sah(tris, 8)
        n2, r2, tris2, perm = ob.node2_convert(nodes, root, tris)
        rays = sq.gen_secondary_rays(tris, 200, seed=1, tmin=0.0)
        a = ob.trace_rays(nodes, root, tris, rays)
        b = ob.node2_trace(n2, r2, tris2, perm, rays)
        assert np.array_equal(a["hit_id"], b["hit_id"])


def ray_hits(tri, ray):
    t = C.c_float()
    o = ob.oracle()
    ok = o.sqo_tri_test(np.ascontiguousarray(tri, np.float32).ctypes.data, ray.ctypes.data, C.byref(t))
    return ok and ray["tmin"] < t.value < ray["tmax"], t.value


@pytest.mark.parametrize("name,tris,nodes,root", TREES, ids=[t[0] for t in TREES])
def test_node2_closest_hits_equal_node1(name, tris, nodes, root):
    n2, r2, tris2, perm = ob.node2_convert(nodes, root, tris)
    cam = sq.pinhole_camera(tris, 48, 32, 60.0, 3.0, spp=2, jitter=1, seed=4)
    rays = np.concatenate([ob.camera_rays(cam), sq.gen_secondary_rays(tris, 3000, seed=9, tmin=1e-4)])
    a = ob.trace_rays(nodes, root, tris, rays)
    b = ob.node2_trace(n2, r2, tris2, perm, rays)
    c = ob.node2_trace(n2, r2, tris2, perm, rays, closest=True)
    assert np.array_equal(b["hit_id"], c["hit_id"])
    assert np.array_equal(b["hit_t"].view(np.uint32), c["hit_t"].view(np.uint32))
    diff = np.nonzero(a["hit_id"] != b["hit_id"])[0]
    assert len(diff) <= max(2, len(rays) // 1000), (name, len(diff))
    for i in diff:
        assert b["hit_id"][i] != sq.MISS
        ok, t = ray_hits(tris[b["hit_id"][i]], rays[i:i + 1])
        assert ok and t <= a["hit_t"][i], (name, i)
    assert (b["counts"] >= 1).all()
