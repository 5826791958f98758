"""Benchmark of the BVH build + traversal hot path on one MI355X.

Workload (the plain GPU path): a 1,000,000-triangle fp32 soup (centroids
U[0,1)^3, Gaussian edges sigma 0.01, seed 2) built as a 63-bit Morton LBVH,
then 1920x1080 primary rays at 4 jittered samples per pixel (8,294,400 rays)
from a 60-degree pinhole camera at distance 3, traced with the reference's
exhaustive semantics (per-pixel node-visit heatmap + closest hit per ray).

One step = build (Morton codes, radix sort, emission, refit) + trace, inputs
resident in HBM. `value` is Mrays/s of the trace (HIP events on the library
stream around the traversal kernel, summed over the timed steps).

The same JSON line carries more sections: "sah32" (the same soup built with
32-bin binned SAH and the same rays traced on that tree), "closest" (the
headline frame's closest hits through the Morton tree and through its
quantised two-child Node2 conversion), "incoherent" (10M
clustered triangles built with the grid and binned-SAH builders, 8M
cosine-weighted secondary rays traced to their closest hits), "batched" (256
scenes built and traced in one captured graph against 256 sequential plain
calls), "served" (a Poisson request stream through the serving loop against
one plain call per request, throughput and p50/p99 latency at three rates),
"build_roofline" (per-phase times of the Morton63 build and the 32-bin SAH
build against their algorithmic HBM bytes), "config1" (the 69,632-triangle
displaced icosphere: reference CPU builds and trace beside the GPU's), and
"cpu_baseline" (the reference's own CPU build + traverse on a bounded sample,
with the GPU trace of the same tree beside it).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-sah] [--no-incoherent]
                    [--no-batched] [--no-served] [--no-build-roofline] [--no-config1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_TRI = 1_000_000
W, H, SPP = 1920, 1080, 4
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "trace_traffic_soup1m.json")
# the sources whose code the committed traffic counters measured
TRACE_SOURCES = ("sequoia_amd/csrc/trace.hip", "sequoia_amd/csrc/trace_common.h")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sources_sha256(paths=TRACE_SOURCES):
    import hashlib
    h = hashlib.sha256()
    for p in paths:
        with open(os.path.join(ROOT, p), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()


def cpu_baseline(sq, ctx, scene, tris, cam, gpu_rays_per_s, budget_s):
    """Times the reference build + traverse on a bounded sample (host cores),
    and the GPU trace of the full frame through that same tree."""
    from oracle import binding as ob
    threads = cpu_threads()
    kind = "reference" if ob.reference_available() else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        nodes, root = ob.ref_build(tris, "recursive")
    else:
        nodes, root = ob.build_sah(tris, 8)
    build_s = time.perf_counter() - t0
    rays = ob.camera_rays(cam)
    # bounded sample: every k-th ray, sized from a small probe to ~budget_s
    probe = rays[:: max(1, len(rays) // 20000)]
    t0 = time.perf_counter()
    if kind == "reference":
        ob.ref_traverse_rays(nodes, root, tris, probe)
    else:
        ob.trace_rays(nodes, root, tris, probe)
    probe_s = max(time.perf_counter() - t0, 1e-6)
    per_ray = probe_s / len(probe)
    n_sample = int(min(len(rays), max(len(probe), budget_s / per_ray)))
    stride = max(1, len(rays) // n_sample)
    sample = rays[::stride]
    t0 = time.perf_counter()
    if kind == "reference":
        ob.ref_traverse_rays(nodes, root, tris, sample)
    else:
        ob.trace_rays(nodes, root, tris, sample)
    trace_s = time.perf_counter() - t0
    mrays = len(sample) / trace_s / 1e6
    # the GPU on the very same tree (uploaded), full frame, device outputs
    same = ctx.bvh_from_host(scene, nodes, root)
    nr = cam.width * cam.height * cam.spp
    d_id, d_t = ctx.alloc(nr * 4), ctx.alloc(nr * 4)
    try:
        ms = []
        for r in range(4):
            ctx.trace_camera_device(same, cam, d_id, d_t, None, None)
            if r:
                ms.append(ctx.timings()["trace_ms"])
    finally:
        ctx.free(d_id)
        ctx.free(d_t)
        same.free()
    gpu_same = nr / (float(np.median(ms)) / 1e3) / 1e6
    return {
        "value": round(mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": kind,
        "cpu_model": cpu_model(), "host_threads_available": os.cpu_count(),
        "sample": "%d of %d config rays (every %d-th) traced exhaustively on the reference's 8-bin recursive SAH "
                  "tree of the same 1M soup, %.1f s; that build took %.2f s (%.2f Mtri/s)"
                  % (len(sample), len(rays), stride, trace_s, build_s, len(tris) / build_s / 1e6),
        "build_mtri_s": round(len(tris) / build_s / 1e6, 3),
        "gpu_over_cpu": round(gpu_rays_per_s / 1e6 / mrays, 1) if mrays > 0 else None,
        "gpu_same_tree_mrays_s": round(gpu_same, 1),
        "gpu_over_cpu_same_tree": round(gpu_same / mrays, 1) if mrays > 0 else None,
    }


def prim_levels(nodes, root):
    """Sum over leaves of (triangles in the leaf) x (depth of the leaf), the
    root at depth 0: the number of binning/partition levels every primitive
    passes through in a top-down SAH build."""
    frontier, depth, total = np.array([root], np.int64), 0, 0
    while len(frontier):
        leaf = (nodes["flags"][frontier] & 1) != 0
        total += depth * int(nodes["count"][frontier[leaf]].astype(np.int64).sum())
        inner = frontier[~leaf]
        cnt = nodes["count"][inner].astype(np.int64)
        first = nodes["child"][inner].astype(np.int64)
        frontier = np.repeat(first, cnt) + (np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt))
        depth += 1
    return total


def _rl(nbytes, ms):
    gbs = nbytes / (ms / 1e3) / 1e9
    return {"ms": round(ms, 4), "bytes": int(nbytes), "gb_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def build_roofline_section(sq, ctx, scene, reps=5):
    """Each Morton63 build phase (HIP events per phase, median of `reps`
    builds after one warm-up, outside the timed steps) against its
    algorithmic HBM bytes, and the 32-bin SAH build against 52 B per
    primitive per level over the levels its tree actually has."""
    log("build roofline: Morton63 phases")
    n = N_TRI
    keys = ("bounds_ms", "morton_ms", "sort_ms", "emit_ms", "refit_ms", "total_build_ms")
    ph = {k: [] for k in keys}
    ctx.set_phase_timing(True)
    try:
        for r in range(reps + 1):
            b = ctx.build(scene, "morton63")
            t = ctx.timings()
            b.free()
            if r:
                for k in keys:
                    ph[k].append(t[k])
        passes = int(t["sort_passes"])
    finally:
        ctx.set_phase_timing(False)
    med = {k: float(np.median(v)) for k, v in ph.items()}
    m = 2 * n - 1  # nodes written
    alg = {"bounds": 36 * n, "morton": (36 + 12) * n, "sort": (8 + 24 * passes) * n,
           "emit": 12 * n + 32 * m, "refit": 36 * n + 36 * m}
    out = {"workload": "1M-triangle soup (seed 2), morton63", "sort_passes": passes,
           "formulas": {"bounds": "36 B/tri triangle read", "morton": "36 B/tri read + 12 B/tri key+value write",
                        "sort": "8 B/key histogram read + %d passes x (12 B read + 12 B write)" % passes,
                        "emit": "12 B/tri sorted key+value read + 32 B per node written (2N-1 nodes)",
                        "refit": "36 B/tri read + 36 B per node (32 B box write + 4 B link read)"}}
    tot_b, tot_ms = 0, 0.0
    for k in ("bounds", "morton", "sort", "emit", "refit"):
        out[k] = _rl(alg[k], med[k + "_ms"])
        tot_b += alg[k]
        tot_ms += med[k + "_ms"]
    out["phases_sum"] = _rl(tot_b, tot_ms)
    out["build_ms_phase_timed"] = round(med["total_build_ms"], 4)
    log("build roofline: 32-bin SAH")
    ms = []
    for r in range(4):
        b = ctx.build(scene, "recursive", num_bins=32)
        if r:
            ms.append(ctx.timings()["total_build_ms"])
        if r < 3:
            b.free()
    nodes, root = b.download(), b.root
    b.free()
    pl = prim_levels(nodes, root)
    out["sah32"] = dict(_rl(52 * pl, float(np.median(ms))), prim_levels=int(pl), mean_levels=round(pl / n, 2),
                        formula="52 B per primitive per level (12 centroid + 24 box + 4 id binned; 4 id + 4 key + "
                                "4 id write partitioned) x levels summed over primitives")
    return out


C1_TRI = 69632


def config1_section(sq, ctx, reps=5):
    """The 69,632-triangle displaced icosphere (seed 1, amplitude 0.1): the
    reference's Morton (30-bit codes, 8-bit digits) and 16-bin SAH builds and
    the 1920x1080 60-degree primary-ray trace on host cores, beside the
    library's same builds and trace on the GPU (median of `reps` after one
    warm-up each)."""
    from oracle import binding as ob
    log("config 1: %d-triangle icosphere" % C1_TRI)
    tris = sq.gen_icosphere(C1_TRI, 1, 0.1)
    n = len(tris)
    cam = sq.pinhole_camera(tris, W, H, 60.0, 3.0)
    nr = W * H
    sc = ctx.upload(tris)
    d_id, d_t, d_cnt, d_rgb = ctx.alloc(nr * 4), ctx.alloc(nr * 4), ctx.alloc(nr * 4), ctx.alloc(nr * 3)
    gpu = {}
    try:
        for name, kind, kw in (("morton30", "morton", {}), ("sah16", "recursive", {"num_bins": 16})):
            ms = []
            for r in range(reps + 1):
                b = ctx.build(sc, kind, **kw)
                if r:
                    ms.append(ctx.timings()["total_build_ms"])
                if r < reps:
                    b.free()
            bm = float(np.median(ms))
            tms = []
            for r in range(reps + 1):
                ctx.trace_camera_device(b, cam, d_id, d_t, d_cnt, d_rgb)
                if r:
                    tms.append(ctx.timings()["trace_ms"])
            b.free()
            tm = float(np.median(tms))
            gpu[name] = {"build_ms": round(bm, 4), "build_mtri_s": round(n / (bm / 1e3) / 1e6, 1),
                         "trace_ms": round(tm, 4), "trace_mrays_s": round(nr / (tm / 1e3) / 1e6, 1)}
    finally:
        for p in (d_id, d_t, d_cnt, d_rgb):
            ctx.free(p)
        sc.free()
    out = {"workload": "69,632-triangle icosphere (3-octave value noise, amplitude 0.1, seed 1); 1920x1080 "
                       "primary rays, 60-degree pinhole at distance 3, exhaustive with heatmap",
           "triangles": n, "rays": nr, "gpu": gpu}
    if not ob.reference_available():
        out["cpu"] = None
        return out
    cpu = {"kind": "reference", "cores": cpu_threads(), "cpu_model": cpu_model()}
    mt = ob.ref_time_build(tris, "morton", reps + 1)[1:]
    cpu["morton30_build_ms"] = round(float(np.median(mt)), 3)
    cpu["morton30_build_mtri_s"] = round(n / (np.median(mt) / 1e3) / 1e6, 2)
    bins = 16 if ob.reference_available(16) else 8
    st = ob.ref_time_build(tris, "recursive", reps + 1, bins)[1:]
    cpu["sah_bins"] = bins
    cpu["sah_build_ms"] = round(float(np.median(st)), 3)
    cpu["sah_build_mtri_s"] = round(n / (np.median(st) / 1e3) / 1e6, 2)
    nodes, root = ob.ref_build(tris, "recursive", bins)
    rays = ob.camera_rays(cam)
    ob.ref_traverse_rays(nodes, root, tris, rays)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        ob.ref_traverse_rays(nodes, root, tris, rays)
        ts.append(time.perf_counter() - t0)
    tm = float(np.median(ts))
    cpu["trace_ms"] = round(tm * 1e3, 2)
    cpu["trace_mrays_s"] = round(nr / tm / 1e6, 2)
    cpu["trace_note"] = "the reference's traverse() per ray over the config camera's rays, OpenMP over rays, " \
                        "on its %d-bin SAH tree" % bins
    out["cpu"] = cpu
    g = gpu["sah16"]
    out["gpu_over_cpu"] = {"morton_build": round(cpu["morton30_build_ms"] / gpu["morton30"]["build_ms"], 1),
                           "sah_build": round(cpu["sah_build_ms"] / g["build_ms"], 1),
                           "trace_sah_tree": round(cpu["trace_ms"] / g["trace_ms"], 1)}
    return out


BATCH_SCENES = 256


def sah_section(sq, ctx, scene, cam, nrays, reps=3):
    """The headline soup built with 32-bin binned SAH (median of `reps`
    builds) and the headline rays traced on that tree."""
    log("sah32 build + trace")
    ms = []
    for _ in range(reps):
        b = ctx.build(scene, "recursive", num_bins=32)
        ms.append(ctx.timings()["total_build_ms"])
        b.free()
    b = ctx.build(scene, "recursive", num_bins=32)
    depth = b.info()[2]
    ctx.trace_camera(b, cam, want_hits=False)  # warm
    tms = []
    for _ in range(reps):
        ctx.trace_camera(b, cam, want_hits=False)
        tms.append(ctx.timings()["trace_ms"])
    b.free()
    bms, tm = float(np.median(ms)), float(np.median(tms))
    return {"build_ms": round(bms, 3), "build_mtri_s": round(N_TRI / (bms / 1e3) / 1e6, 1), "tree_depth": depth,
            "trace_ms": round(tm, 3), "mrays_s": round(nrays / (tm / 1e3) / 1e6, 1)}


def closest_section(sq, ctx, scene, cam, nrays, reps=3):
    """Closest hits of the headline frame through the Morton63 tree (Node1),
    its quantised two-child conversion (Node2) and its wide conversions
    (Node<4>, Node<8>), device outputs; each conversion's time is reported
    with it, and the wide trees' exhaustive frame (their own box-test counts)
    beside it."""
    log("closest hits: Node1, Node2, Node<4>, Node<8>")
    b = ctx.build(scene, "morton63")
    trees, conv = [(b, "node1")], {}
    d_id, d_t = ctx.alloc(nrays * 4), ctx.alloc(nrays * 4)
    out = {"workload": "headline frame, closest-hit semantics"}
    try:
        for name, make in (("node2", lambda: ctx.node2(b)), ("node4", lambda: ctx.nodek(b, 4)),
                           ("node8", lambda: ctx.nodek(b, 8))):
            ctx.sync()
            t0 = time.perf_counter()
            t = make()
            ctx.sync()
            conv[name] = (time.perf_counter() - t0) * 1e3
            trees.append((t, name))

        def timed(tree, sem):
            ms = []
            for r in range(reps + 1):
                ctx.trace_camera_device(tree, cam, d_id, d_t, None, None, semantics=sem)
                if r:
                    ms.append(ctx.timings()["trace_ms"])
            return float(np.median(ms))

        for tree, name in trees:
            m = timed(tree, sq.TRACE_CLOSEST)
            out[name + "_trace_ms"] = round(m, 3)
            out[name + "_mrays_s"] = round(nrays / (m / 1e3) / 1e6, 1)
            if name != "node1":
                out[name + "_convert_ms"] = round(conv[name], 2)
                out[name + "_nodes"] = tree.info()["nodes"]
            if name in ("node4", "node8"):
                m = timed(tree, sq.TRACE_EXHAUSTIVE)
                out[name + "_exhaustive_trace_ms"] = round(m, 3)
    finally:
        ctx.free(d_id)
        ctx.free(d_t)
        for tree, name in trees[1:]:
            tree.free()
        b.free()
    return out


INC_TRI, INC_RAYS, INC_SUB = 10_000_000, 8_000_000, 100_000


def incoherent_section(sq, ctx, reps=3):
    """10M triangles in 64 Gaussian clusters (seed 3): grid and 8-bin binned
    SAH builds (median of `reps`), then 8M secondary rays (origins on random
    triangles, cosine-weighted about the normal, tmin 1e-4) traced to their
    closest hits on the SAH tree, and a 100k-ray subset with the exhaustive
    reference semantics."""
    log("incoherent: generating %d clustered triangles, %d secondary rays" % (INC_TRI, INC_RAYS))
    tris = sq.gen_clusters(INC_TRI, 64, 3)
    rays = sq.gen_secondary_rays(tris, INC_RAYS, seed=3, tmin=1e-4)
    sc = ctx.upload(tris)
    out = {"workload": "10M clustered triangles (64 clusters, seed 3), 8M secondary rays, tmin 1e-4"}
    try:
        for kind in ("grid", "recursive"):
            ms = []
            for _ in range(reps):
                b = ctx.build(sc, kind, num_bins=8)
                ms.append(ctx.timings()["total_build_ms"])
                b.free()
            bms = float(np.median(ms))
            out["build_%s_ms" % kind] = round(bms, 2)
            out["build_%s_mtri_s" % kind] = round(INC_TRI / (bms / 1e3) / 1e6, 1)
        b = ctx.build(sc, "recursive", num_bins=8)
        log("incoherent: closest-hit trace")
        r = ctx.trace_rays(b, rays, semantics=sq.TRACE_CLOSEST, count_work=True)
        t = ctx.timings()
        out["closest_trace_ms"] = round(t["trace_ms"], 1)
        out["closest_mrays_s"] = round(INC_RAYS / (t["trace_ms"] / 1e3) / 1e6, 3)
        out["closest_node_tests_per_ray"] = round(t["node_tests"] / INC_RAYS, 1)
        out["closest_tri_tests_per_ray"] = round(t["tri_tests"] / INC_RAYS, 1)
        alg = 32.0 * t["node_tests"] + 36.0 * t["tri_tests"]
        out["closest_roofline"] = dict(_rl(alg, t["trace_ms"]), kernel="k_trace<true, false, 1, 1, true>",
                                       formula="32 B per node test + 36 B per triangle test (counted on the device "
                                               "in the same launch)")
        out["hit_fraction"] = round(float((r["hit_id"] != sq.MISS).mean()), 4)
        # the same rays through the tree's Node<4> conversion (hits compared)
        ctx.sync()
        t0 = time.perf_counter()
        t4 = ctx.nodek(b, 4)
        ctx.sync()
        out["node4_convert_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
        log("incoherent: closest-hit trace, Node<4>")
        r4 = ctx.trace_rays(t4, rays, semantics=sq.TRACE_CLOSEST, count_work=True)
        t = ctx.timings()
        out["node4_closest_trace_ms"] = round(t["trace_ms"], 1)
        out["node4_closest_mrays_s"] = round(INC_RAYS / (t["trace_ms"] / 1e3) / 1e6, 3)
        out["node4_box_tests_per_ray"] = round(t["node_tests"] / INC_RAYS, 1)
        out["node4_tri_tests_per_ray"] = round(t["tri_tests"] / INC_RAYS, 1)
        alg4 = 32.0 * t["node_tests"] + 36.0 * t["tri_tests"]
        out["node4_closest_roofline"] = dict(_rl(alg4, t["trace_ms"]), kernel="k_trace<true, false, 4, 1>",
                                             formula="32 B per box test (a Node<4> is 4 boxes in 128 B) + 36 B "
                                                     "per triangle test")
        out["node4_hits_differ_from_node1"] = int((r4["hit_id"] != r["hit_id"]).sum())
        t4.free()
        sub = rays[:INC_SUB]
        ctx.trace_rays(b, sub, count_work=True)
        t = ctx.timings()
        out["exhaustive_subset_rays"] = INC_SUB
        out["exhaustive_mrays_s"] = round(INC_SUB / (t["trace_ms"] / 1e3) / 1e6, 3)
        out["exhaustive_node_tests_per_ray"] = round(t["node_tests"] / INC_SUB, 1)
        b.free()
    finally:
        sc.free()
    return out


def batched_workload(sq):
    """256 icosphere scenes (3-octave value noise, amplitude 0.1, seeds 4000+i)
    with triangle counts log-uniform in [1k, 500k] and 4-sample jittered
    16:9 cameras whose ray counts are log-uniform in [64k, 2M] (draws: numpy
    default_rng(4))."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(4)
    ntri = np.exp(rng.uniform(np.log(1e3), np.log(5e5), BATCH_SCENES)).astype(np.int64)
    nray = np.exp(rng.uniform(np.log(65536.0), np.log(2097152.0), BATCH_SCENES)).astype(np.int64)
    with ThreadPoolExecutor(16) as ex:
        scenes = list(ex.map(lambda i: sq.gen_icosphere(int(ntri[i]), 4000 + i, 0.1), range(BATCH_SCENES)))
    cams = []
    for i, t in enumerate(scenes):
        npix = nray[i] / 4.0
        h = max(8, int(round(np.sqrt(npix * 9.0 / 16.0))))
        w = max(8, int(round(npix / h)))
        cams.append(sq.pinhole_camera(t, w, h, 60.0, 3.0, spp=4, jitter=1, seed=4000 + i))
    return scenes, cams


def batched_section(sq, ctx, steps, warmup):
    """Batched (one captured graph: build all + trace all) against 256
    sequential plain calls (sq_build + sq_trace_camera per scene), both with
    triangles resident in HBM and outputs written to device memory."""
    log("batched: generating %d scenes" % BATCH_SCENES)
    scenes, cams = batched_workload(sq)
    ntri = sum(len(t) for t in scenes)
    nrays = sum(c.width * c.height * c.spp for c in cams)
    out_mask = sq.OUT_HIT_ID | sq.OUT_HIT_T | sq.OUT_COUNTS | sq.OUT_RGB
    batch = sq.Batch(ctx, scenes)
    batch.set_cameras(cams, outputs=out_mask)
    log("batched: autotune")
    tune = batch.autotune("morton63", repeats=1)
    for _ in range(warmup):
        batch.run("morton63")
    ctx.sync()
    runs, t0 = [], time.perf_counter()
    for _ in range(steps):
        batch.run("morton63")
        runs.append(batch.times()["run_ms"])
    batch_wall = (time.perf_counter() - t0) / steps
    batch.build("morton63")
    batch.trace()
    eager = batch.times()
    # node and triangle tests of the batched trace (a separate counting pass)
    tuned = dict(outputs=out_mask, block_size=tune["block_size"], stack_slots=tune["stack_slots"],
                 kernel=tune["kernel"], lists_per_step=tune["lists_per_step"], blocks_per_cu=tune["blocks_per_cu"])
    batch.set_cameras(cams, count_work=True, **tuned)
    batch.trace()
    work = ctx.timings()
    batch.set_cameras(cams, **tuned)
    batch.trace()  # outputs again (set_cameras allocates fresh output buffers)
    batch_alg = 32.0 * work["node_tests"] + 36.0 * work["tri_tests"]

    # plain path: one build + one trace call per scene
    log("batched: plain per-scene calls")
    resident = [ctx.upload(t) for t in scenes]
    maxr = max(c.width * c.height * c.spp for c in cams)
    maxp = max(c.width * c.height for c in cams)
    d_id, d_t, d_cnt, d_rgb = ctx.alloc(maxr * 4), ctx.alloc(maxr * 4), ctx.alloc(maxp * 4), ctx.alloc(maxp * 3)

    def plain_all(check=None):
        for i, sc in enumerate(resident):
            bvh = ctx.build(sc, "morton63")
            ctx.trace_camera_device(bvh, cams[i], d_id, d_t, d_cnt, d_rgb)
            if check is not None and i in check:
                npix = cams[i].width * cams[i].height
                check[i] = ctx.to_host(d_cnt, npix * 4, np.uint32)
            bvh.free()

    check = {i: None for i in range(BATCH_SCENES)}
    plain_all(check)
    parity = all(np.array_equal(check[i], batch.download(i, hits=False, rgb=False)["counts"]) for i in check)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain_all()
    ctx.sync()
    plain_wall = (time.perf_counter() - t0) / steps
    # binned SAH: the same scenes built with the recursive builder (bin count
    # autotuned among 8/16/32) and traced, against the plain per-scene calls
    log("batched: binned SAH")
    batch.set_cameras(cams, outputs=out_mask)
    stune = batch.autotune("recursive", repeats=1)
    nb = stune["num_bins"]
    batch.run("recursive", num_bins=nb)
    sruns, t0 = [], time.perf_counter()
    for _ in range(steps):
        batch.run("recursive", num_bins=nb)
        sruns.append(batch.times()["run_ms"])
    sah_wall = (time.perf_counter() - t0) / steps

    def plain_sah(check=None):
        for i, sc in enumerate(resident):
            bvh = ctx.build(sc, "recursive", num_bins=nb)
            ctx.trace_camera_device(bvh, cams[i], d_id, d_t, d_cnt, d_rgb)
            if check is not None and i in check:
                check[i] = ctx.to_host(d_cnt, cams[i].width * cams[i].height * 4, np.uint32)
            bvh.free()

    scheck = {i: None for i in range(BATCH_SCENES)}
    plain_sah(scheck)
    sah_parity = all(np.array_equal(scheck[i], batch.download(i, hits=False, rgb=False)["counts"]) for i in scheck)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain_sah()
    ctx.sync()
    plain_sah_wall = (time.perf_counter() - t0) / steps
    for p in (d_id, d_t, d_cnt, d_rgb):
        ctx.free(p)
    for sc in resident:
        sc.free()
    batch.free()
    run_ms = float(np.median(runs))
    return {
        "workload": "256 icosphere scenes (value-noise amplitude 0.1), triangles log-uniform [1k, 500k], "
                    "4-spp jittered cameras with rays log-uniform [64k, 2M]; morton63 build + exhaustive trace",
        "scenes": BATCH_SCENES, "triangles": int(ntri), "rays": int(nrays),
        "batched_ms": round(run_ms, 2), "batched_wall_ms": round(batch_wall * 1e3, 2),
        "batched_mrays_s": round(nrays / (batch_wall) / 1e6, 1),
        "batched_build_ms": round(eager["build_ms"], 2), "batched_trace_ms": round(eager["trace_ms"], 2),
        "plain_wall_ms": round(plain_wall * 1e3, 2),
        "plain_mrays_s": round(nrays / plain_wall / 1e6, 1),
        "speedup_vs_plain": round(plain_wall / batch_wall, 2),
        "trace_roofline": dict(_rl(batch_alg, eager["trace_ms"]), kernel="k_trace<false, true, 1, 2, true>",
                               node_tests=int(work["node_tests"]), tri_tests=int(work["tri_tests"]),
                               formula="32 B per node test + 36 B per triangle test over all 256 scenes, "
                                       "against the eager batched trace time"),
        "counts_equal_plain": bool(parity), "checked_scenes": len(check),
        "autotune": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tune.items()},
        "sah": {
            "workload": "the same 256 scenes built with binned SAH (recursive builder, bin count autotuned among "
                        "8/16/32; eager build, graph trace) against 256 plain sq_build + sq_trace_camera calls with "
                        "that bin count",
            "num_bins": nb, "batched_ms": round(float(np.median(sruns)), 2), "batched_wall_ms": round(sah_wall * 1e3, 2),
            "batched_mrays_s": round(nrays / sah_wall / 1e6, 1), "plain_wall_ms": round(plain_sah_wall * 1e3, 2),
            "plain_mrays_s": round(nrays / plain_sah_wall / 1e6, 1),
            "speedup_vs_plain": round(plain_sah_wall / sah_wall, 2), "counts_equal_plain": bool(sah_parity),
            "autotune": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in stune.items()},
        },
    }


SERVED_REQUESTS = 240
SERVED_RATES = (100.0, 300.0, 1000.0, 10000.0)


def served_workload(sq):
    """Request mix (numpy default_rng(5)): 70% build+trace of a new icosphere
    scene (triangles log-uniform [1k, 2M], value-noise amplitude 0.1), 30%
    trace-only on an already built scene picked Zipf(s=1.1) by build order;
    rays log-uniform [4k, 2M], half primary (16:9 jittered camera, 1-4 spp),
    half cosine-weighted secondary rays from the scene surface (tmin 1e-4)."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(5)
    plan, built = [], []
    for i in range(SERVED_REQUESTS):
        nray = int(np.exp(rng.uniform(np.log(4096.0), np.log(2097152.0))))
        primary = rng.random() < 0.5
        if built and rng.random() < 0.3:
            ranks = np.arange(1, len(built) + 1, dtype=np.float64)
            p = ranks ** -1.1
            key = built[int(rng.choice(len(built), p=p / p.sum()))]
            plan.append((key, None, nray, primary, 7000 + i))
        else:
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
            key = len(built)
            built.append(key)
            plan.append((key, n, nray, primary, 7000 + i))
    sizes = {k: n for (k, n, _, _, _) in plan if n}
    with ThreadPoolExecutor(16) as ex:
        scene = dict(zip(sizes, ex.map(lambda k: sq.gen_icosphere(sizes[k], 6000 + k, 0.1), list(sizes))))
    reqs = []
    for key, n, nray, primary, seed in plan:
        tris = scene[key]
        if primary:
            spp = 1 + seed % 4
            npix = max(64, nray // spp)
            h = max(8, int(round(np.sqrt(npix * 9.0 / 16.0))))
            w = max(8, int(round(npix / h)))
            req = dict(key=key, build=n is not None, camera=sq.pinhole_camera(tris, w, h, 60.0, 3.0, spp=spp, jitter=1,
                                                                                 seed=seed), rays=None)
        else:
            req = dict(key=key, build=n is not None, camera=None,
                       rays=sq.gen_secondary_rays(tris, nray, seed=seed, tmin=1e-4))
        reqs.append(req)
    return reqs, scene


def _arrivals(rate, n, seed):
    return np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, n))


def _device_outputs(ctx, reqs):
    """One set of device output buffers per request (requests in flight
    together must not share them)."""
    bufs = []
    for r in reqs:
        if r["rays"] is not None:
            nray = npix = len(r["rays"])
        else:
            c = r["camera"]
            npix = c.width * c.height
            nray = npix * c.spp
        bufs.append({"hit_id": ctx.alloc(4 * nray), "hit_t": ctx.alloc(4 * nray), "counts": ctx.alloc(4 * npix),
                     "rgb": ctx.alloc(3 * npix) if r["rays"] is None else None})
    return bufs


def _served_run(sq, reqs, scene, arrivals, srv, device_out=None):
    """One pass of the request stream through the (warm, long-lived) server
    `srv`, as the plain path runs on a warm context; device_out: per-request
    device output buffers (else host outputs)."""
    import threading
    srv.reset_stats()
    ids = [None] * len(reqs)

    def client():
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            d = arrivals[i] - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids[i] = srv.submit(r["key"], tris=scene[r["key"]] if r["build"] else None, camera=r["camera"],
                                rays=r["rays"], device_out=device_out[i] if device_out else None)

    th = threading.Thread(target=client)
    t0 = time.perf_counter()
    th.start()
    th.join()
    srv.drain()
    wall = time.perf_counter() - t0
    lat, failed = [], []
    for k, i in enumerate(ids):
        st, _ = srv.wait(i)
        if st["status"] != 0:
            failed.append((k, st["status"], st["fault"]))
            continue
        lat.append(st["latency_ms"])
    stats = srv.stats()
    stats["failed_requests"] = failed
    return np.array(lat), wall, stats


def _plain_run(sq, ctx, reqs, scene, arrivals, device_out=None, ray_buf=None):
    """One request at a time in arrival order with the plain calls (upload +
    sq_build morton63 for new scenes, sq_trace_camera / sq_trace_rays with
    host outputs, or with device outputs when device_out is given: ray lists
    are then copied to ray_buf first); built trees stay resident for
    trace-only requests."""
    cache, lat = {}, []
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        now = time.perf_counter() - t0
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        if r["build"]:
            sc = ctx.upload(scene[r["key"]])
            cache[r["key"]] = (sc, ctx.build(sc, "morton63"))
        bvh = cache[r["key"]][1]
        try:
            if device_out is not None:
                b = device_out[i]
                if r["rays"] is not None:
                    ctx.to_device(ray_buf, r["rays"])
                    ctx.trace_rays_device(bvh, ray_buf, len(r["rays"]), b["hit_id"], b["hit_t"], b["counts"])
                else:
                    ctx.trace_camera_device(bvh, r["camera"], b["hit_id"], b["hit_t"], b["counts"], b["rgb"])
            elif r["rays"] is not None:
                ctx.trace_rays(bvh, r["rays"])
            else:
                ctx.trace_camera(bvh, r["camera"])
        except sq.SqError as e:
            log("plain: request %d failed: %s (tree depth %d)" % (i, e, bvh.info()[2]))
            continue
        lat.append((time.perf_counter() - t0 - arrivals[i]) * 1e3)
    wall = time.perf_counter() - t0
    for sc, b in cache.values():
        b.free()
        sc.free()
    return np.array(lat), wall


def served_section(sq, ctx):
    log("served: generating %d requests" % SERVED_REQUESTS)
    reqs, scene = served_workload(sq)
    nrays = sum((r["rays"].shape[0] if r["rays"] is not None else r["camera"].width * r["camera"].height *
                 r["camera"].spp) for r in reqs)
    nbuild = sum(1 for r in reqs if r["build"])
    # warm both paths on the whole stream once (untimed, back to back): the
    # host allocator then serves the output arrays of the timed runs from
    # memory it already holds, whichever path runs first
    srv = sq.Server(0)
    _served_run(sq, reqs, scene, np.zeros(len(reqs)), srv)
    _plain_run(sq, ctx, reqs, scene, np.zeros(len(reqs)))
    def point(rate, sl, sw, st, pl, pw):
        return {
            "offered_req_s": rate,
            "served": {"req_s": round(len(reqs) / sw, 1), "mrays_s": round(nrays / sw / 1e6, 1),
                       "p50_ms": round(float(np.percentile(sl, 50)), 2),
                       "p99_ms": round(float(np.percentile(sl, 99)), 2),
                       "waves": int(st["waves"]), "mean_wave_requests": round(st["mean_wave_requests"], 2),
                       "failed": len(st["failed_requests"]), "host_stage_in_ms": round(st["stage_in_ms"], 1),
                       "host_stage_out_ms": round(st["stage_out_ms"], 1), "wave_ms_total": round(st["wave_ms"], 1)},
            "plain": {"req_s": round(len(reqs) / pw, 1), "mrays_s": round(nrays / pw / 1e6, 1),
                      "p50_ms": round(float(np.percentile(pl, 50)), 2),
                      "p99_ms": round(float(np.percentile(pl, 99)), 2)},
        }

    # device outputs: both paths trace into per-request device buffers
    bufs = _device_outputs(ctx, reqs)
    ray_buf = ctx.alloc(32 * max(len(r["rays"]) for r in reqs if r["rays"] is not None))
    points_dev = []
    try:
        _served_run(sq, reqs, scene, np.zeros(len(reqs)), srv, bufs)
        _plain_run(sq, ctx, reqs, scene, np.zeros(len(reqs)), bufs, ray_buf)
        for rate in SERVED_RATES:
            arr = _arrivals(rate, len(reqs), 55)
            log("served: rate %.0f/s, device outputs" % rate)
            sl, sw, st = _served_run(sq, reqs, scene, arr, srv, bufs)
            pl, pw = _plain_run(sq, ctx, reqs, scene, arr, bufs, ray_buf)
            points_dev.append(point(rate, sl, sw, st, pl, pw))
    finally:
        for b in bufs:
            for p in b.values():
                if p is not None:
                    ctx.free(p)
        ctx.free(ray_buf)
    points = []
    for rate in SERVED_RATES:
        arr = _arrivals(rate, len(reqs), 55)
        log("served: rate %.0f/s" % rate)
        sl, sw, st = _served_run(sq, reqs, scene, arr, srv)
        for k, code, fault in st["failed_requests"]:
            r = reqs[k]
            log("served: request %d failed (status %d, fault %d): build=%s n=%d rays=%s" % (
                k, code, fault, r["build"], len(scene[r["key"]]),
                len(r["rays"]) if r["rays"] is not None else "camera"))
        pl, pw = _plain_run(sq, ctx, reqs, scene, arr)
        points.append(point(rate, sl, sw, st, pl, pw))
    srv.close()
    return {
        "workload": "%d requests: %d build+trace of icosphere scenes (log-uniform [1k, 2M] triangles), %d "
                    "trace-only Zipf(1.1) repeats; rays log-uniform [4k, 2M], half primary camera, half cosine "
                    "secondary lists; Poisson arrivals; host inputs and outputs (PCIe included); budget 200 GB"
                    % (len(reqs), nbuild, len(reqs) - nbuild),
        "requests": len(reqs), "rays": int(nrays), "triangles_built": int(sum(len(scene[k]) for k in scene)),
        "points": points,
        "points_device_outputs": points_dev,
        "modes": "points: host inputs and host outputs (PCIe both ways, outputs in freshly allocated host "
                 "arrays); points_device_outputs: host inputs, outputs written in place to device buffers "
                 "(both paths)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--semantics", default="exhaustive", choices=["exhaustive", "closest"])
    ap.add_argument("--no-sah", action="store_true", help="skip the 32-bin SAH build + trace of the headline soup")
    ap.add_argument("--no-incoherent", action="store_true", help="skip the 10M-triangle secondary-ray section")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-scene batched-vs-plain comparison")
    ap.add_argument("--batched-steps", type=int, default=3)
    ap.add_argument("--no-served", action="store_true", help="skip the Poisson-stream served-vs-plain comparison")
    ap.add_argument("--no-build-roofline", action="store_true", help="skip the per-phase build roofline")
    ap.add_argument("--no-config1", action="store_true", help="skip the 69,632-triangle CPU-vs-GPU comparison")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (byte-counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import sequoia_amd as sq
    log("generating %d-triangle soup" % N_TRI)
    tris = sq.gen_soup(N_TRI, 2, 0.01)
    cam = sq.pinhole_camera(tris, W, H, 60.0, 3.0, spp=SPP, jitter=1, seed=2)
    nrays = W * H * SPP
    ctx = sq.Context(0)
    scene = ctx.upload(tris)
    d_id = ctx.alloc(nrays * 4)
    d_t = ctx.alloc(nrays * 4)
    d_cnt = ctx.alloc(W * H * 4)
    d_rgb = ctx.alloc(W * H * 3)
    sem = sq.TRACE_CLOSEST if a.semantics == "closest" else sq.TRACE_EXHAUSTIVE
    if a.copy_calibration:
        nb = a.copy_calibration << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy(dst, src, nb)
        ctx.free(src)
        ctx.free(dst)

    def step(count_work=False):
        bvh = ctx.build(scene, "morton63")
        bt = ctx.timings()
        ctx.trace_camera_device(bvh, cam, d_id, d_t, d_cnt, d_rgb, semantics=sem, count_work=count_work)
        tt = ctx.timings()
        bvh.free()
        return bt, tt

    for i in range(a.warmup):
        log("warmup", i)
        step()
    # work counters for the roofline (separate pass, outside the timed region)
    _, tw = step(count_work=True)
    node_tests, tri_tests = tw["node_tests"], tw["tri_tests"]

    ctx.sync()
    t0 = time.perf_counter()
    builds, traces = [], []
    for i in range(a.steps):
        bt, tt = step()
        builds.append(bt["total_build_ms"])
        traces.append(tt["trace_ms"])
    ctx.sync()
    wall = time.perf_counter() - t0
    trace_ms = float(np.mean(traces))
    build_ms = float(np.mean(builds))
    rays_per_s = nrays / (trace_ms / 1e3)
    alg_bytes = 32.0 * node_tests + 36.0 * tri_tests
    achieved = alg_bytes / (trace_ms / 1e3) / 1e9

    heat = ctx.to_host(d_cnt, W * H * 4, np.uint32)
    # traffic below L2 comes from the committed counter passes; it is used only
    # when they measured the trace sources of this tree
    traffic, traffic_note = None, "no counter profile"
    if a.semantics == "exhaustive" and os.path.exists(TRAFFIC_PROFILE):
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
        if prof.get("sources_sha256") == sources_sha256():
            traffic = prof.get("traffic_bytes_per_launch")
            traffic_note = "FETCH_SIZE + WRITE_SIZE passes on these trace sources (%s)" % TRAFFIC_PROFILE[len(ROOT) + 1:]
        else:
            traffic_note = "stale: the counter profile measured other trace sources"
    line = {
        "metric": "Mrays/s & build Mtri/s on 1 GPU; served p50/p99 latency, batched vs plain",
        "value": round(rays_per_s / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain GPU path: 1M-triangle soup (U[0,1)^3 centroids, N(0,0.01^2) edges, seed 2), "
                        "63-bit Morton LBVH, 1920x1080x4 jittered primary rays (8,294,400), 60-deg pinhole at "
                        "distance 3, exhaustive reference semantics with per-pixel heatmap",
            "triangles": N_TRI, "rays": nrays, "builder": "morton63", "semantics": a.semantics,
        },
        "build_mtri_s": round(N_TRI / (build_ms / 1e3) / 1e6, 2),
        "build_ms": round(build_ms, 3),
        "trace_ms": round(trace_ms, 3),
        "node_tests_per_ray": round(node_tests / nrays, 2),
        "tri_tests_per_ray": round(tri_tests / nrays, 2),
        "heatmap_mean_tests_per_pixel": round(float(heat.mean()) / SPP, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": "k_trace<false, false, 1, 2, true>", "algorithmic_bytes_per_launch": alg_bytes},
    }
    # the secondary sections never take the headline line down with them
    if not a.no_build_roofline:
        try:
            line["build_roofline"] = build_roofline_section(sq, ctx, scene)
        except Exception as e:  # noqa: BLE001
            line["build_roofline"] = {"error": repr(e)}
    if not a.no_config1:
        try:
            line["config1"] = config1_section(sq, ctx)
        except Exception as e:  # noqa: BLE001
            line["config1"] = {"error": repr(e)}
    if not a.no_sah:
        try:
            line["sah32"] = sah_section(sq, ctx, scene, cam, nrays)
        except Exception as e:  # noqa: BLE001
            line["sah32"] = {"error": repr(e)}
    if not a.no_sah:
        try:
            line["closest"] = closest_section(sq, ctx, scene, cam, nrays)
        except Exception as e:  # noqa: BLE001
            line["closest"] = {"error": repr(e)}
    if not a.no_incoherent:
        try:
            line["incoherent"] = incoherent_section(sq, ctx)
        except Exception as e:  # noqa: BLE001
            line["incoherent"] = {"error": repr(e)}
    if not a.no_batched:
        try:
            line["batched"] = batched_section(sq, ctx, a.batched_steps, 1)
        except Exception as e:  # noqa: BLE001
            line["batched"] = {"error": repr(e)}
    if not a.no_served:
        try:
            line["served"] = served_section(sq, ctx)
        except Exception as e:  # noqa: BLE001
            line["served"] = {"error": repr(e)}
    if not a.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(sq, ctx, scene, tris, cam, rays_per_s, a.cpu_budget)
    for p in (d_id, d_t, d_cnt, d_rgb):
        ctx.free(p)
    scene.free()
    ctx.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
