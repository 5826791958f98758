// batch.hip — many scenes built and traced together.
//
// The reference builds and traces one Hierarchy per call
// (build_hierarchy_morton, builder.cpp:2474-2508; trace, tracer.cpp:195-230).
// A batch concatenates the triangles of many scenes and runs each stage once
// for all of them:
//   bounds   one workgroup per 4096-triangle tile reduces the tile's part of
//            every scene it overlaps into that scene's ordered-uint box,
//   codes    each triangle's code against its own scene's box (the same
//            device function as the single-scene path),
//   sort     LSD passes on the codes, then segment passes on the scene of
//            each value: order (scene, code, index), codes carried along,
//   emit     the LBVH kernels with one descriptor per scene,
//   trace    one persistent launch over a job table (one job per scene);
//            waves steal 8x8-pixel tiles of any scene from one counter.
// No stage synchronises with the host, so build + trace is captured once
// as a HIP graph and replayed. A scene's tree and outputs equal those of the
// single-scene calls: bounds are exact min/max, the code function is shared,
// the order within a scene is (code, index) in both, emission is per scene
// with scene-local slots and triangle indices, traversal is per ray.
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <vector>

#include "trace_common.h"

namespace sq {

constexpr uint32_t kBbTile = 4096;
constexpr uint32_t kBbThreads = 256;

__global__ void k_batch_bounds_init(uint32_t* ord, uint32_t nscenes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 6 * nscenes; i += gridDim.x * blockDim.x)
        ord[i] = (i % 6) < 3 ? 0xFFFFFFFFu : 0u;
}

// Largest s in [0, nscenes) with off[s] <= t.
__device__ __forceinline__ uint32_t scene_of(const uint32_t* __restrict__ off, uint32_t nscenes, uint32_t t) {
    uint32_t lo = 0, hi = nscenes;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kBbThreads) void k_batch_bounds(const sq_triangle* __restrict__ tris,
                                                             const uint32_t* __restrict__ off, uint32_t nscenes,
                                                             uint32_t total, uint32_t* __restrict__ ord) {
    __shared__ uint32_t s_red[6][kBbThreads / 64];
    const uint32_t t0 = blockIdx.x * kBbTile;
    const uint32_t t1 = (total - t0) < kBbTile ? total : t0 + kBbTile;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t s = scene_of(off, nscenes, t0); s < nscenes && off[s] < t1; s++) {
        const uint32_t b = off[s] > t0 ? off[s] : t0;
        const uint32_t e = off[s + 1] < t1 ? off[s + 1] : t1;
        uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
        uint32_t hi[3] = {0u, 0u, 0u};
        for (uint32_t i = b + threadIdx.x; i < e; i += kBbThreads) {
            float v[9];
            sqd::load_tri(tris, i, v);
            sqd::Box3 bx = sqd::tri_box(v);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                uint32_t x = sqd::f2ord(bx.mn[k]), y = sqd::f2ord(bx.mx[k]);
                lo[k] = x < lo[k] ? x : lo[k];
                hi[k] = y > hi[k] ? y : hi[k];
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                uint32_t x = __shfl_xor(lo[k], o, 64), y = __shfl_xor(hi[k], o, 64);
                lo[k] = x < lo[k] ? x : lo[k];
                hi[k] = y > hi[k] ? y : hi[k];
            }
        }
        if (lane == 0) {
            for (int k = 0; k < 3; k++) { s_red[k][wave] = lo[k]; s_red[3 + k][wave] = hi[k]; }
        }
        __syncthreads();
        if (threadIdx.x < 6) {
            const int k = threadIdx.x;
            uint32_t r = s_red[k][0];
            for (int w = 1; w < kBbThreads / 64; w++)
                r = (k < 3) ? (s_red[k][w] < r ? s_red[k][w] : r) : (s_red[k][w] > r ? s_red[k][w] : r);
            if (k < 3) atomicMin(&ord[6 * s + k], r);
            else atomicMax(&ord[6 * s + k], r);
        }
        __syncthreads();
    }
}

template <bool KEY64>
__global__ __launch_bounds__(256) void k_batch_morton(const sq_triangle* __restrict__ tris,
                                                      const uint32_t* __restrict__ off, uint32_t nscenes,
                                                      uint32_t total, uint32_t bits, const uint32_t* __restrict__ ord,
                                                      void* __restrict__ keys_out, uint32_t* __restrict__ vals_out) {
    const uint32_t first = blockIdx.x * blockDim.x;
    const uint32_t i = first + threadIdx.x;
    if (i >= total) return;
    uint32_t s = scene_of(off, nscenes, first);
    while (off[s + 1] <= i) s++;
    float smn[3], ext[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        smn[k] = sqd::ord2f(ord[6 * s + k]);
        ext[k] = sqd::ord2f(ord[6 * s + 3 + k]) - smn[k];
    }
    const float scale = (float)((1u << bits) - 1u);
    float v[9];
    sqd::load_tri(tris, i, v);
    if (KEY64) ((uint64_t*)keys_out)[i] = sqd::morton_code<true>(v, smn, ext, scale);
    else ((uint32_t*)keys_out)[i] = (uint32_t)sqd::morton_code<false>(v, smn, ext, scale);
    vals_out[i] = i;
}

int enqueue_batch_build(sq_ctx* c, hipStream_t stream, const BatchBuild& j) {
    const bool key64 = j.bits > 10;
    // the emission (with refit) writes every node slot of every scene, unused
    // ones as zeros, so the node array takes no fill
    if (c->poison_nodes) SQ_TRY(launch_fill_u32_on(stream, c->num_cus, j.nodes, 0xFFFFFFFFu, (size_t)2 * j.total * 8));
    const uint32_t ib = (6 * j.nscenes + 255) / 256;
    hipLaunchKernelGGL(k_batch_bounds_init, dim3(ib), dim3(256), 0, stream, j.d_ord, j.nscenes);
    const uint32_t tiles = (j.total + kBbTile - 1) / kBbTile;
    hipLaunchKernelGGL(k_batch_bounds, dim3(tiles), dim3(kBbThreads), 0, stream, j.tris, j.d_off, j.nscenes, j.total,
                       j.d_ord);
    const uint32_t mb = (j.total + 255) / 256;
    if (key64)
        hipLaunchKernelGGL(k_batch_morton<true>, dim3(mb), dim3(256), 0, stream, j.tris, j.d_off, j.nscenes, j.total,
                           j.bits, j.d_ord, j.keys[0], j.vals[0]);
    else
        hipLaunchKernelGGL(k_batch_morton<false>, dim3(mb), dim3(256), 0, stream, j.tris, j.d_off, j.nscenes, j.total,
                           j.bits, j.d_ord, j.keys[0], j.vals[0]);
    SQ_CHECK_HIP(hipGetLastError());
    SortBufs sb{stream, {j.keys[0], j.keys[1]}, {j.vals[0], j.vals[1]}, j.hist, j.os, j.os_err, c->num_cus};
    int slot = 0;
    uint32_t passes = 0;
    SQ_TRY(radix_sort_bufs(sb, j.total, 3 * j.bits, j.digit, key64, j.d_off, j.nscenes, 0, &slot, &passes));
    SQ_TRY(lbvh_emit_batch(c, stream, j.tris, j.total, j.keys[slot], j.vals[slot], key64, j.d_desc, j.nscenes,
                           j.layout, j.threshold, j.nodes, j.ws, j.d_depth));
    return SQ_OK;
}

uint32_t morton_depth_bound(uint32_t bits, uint32_t n) {
    uint32_t lg = 0;
    while ((1ull << lg) < n) lg++;
    return 3 * bits + lg + 3;
}

}  // namespace sq

using namespace sq;

struct sq_batch {
    sq_ctx* ctx = nullptr;
    uint32_t nscenes = 0, total = 0, max_n = 0;
    std::vector<uint32_t> counts, off;
    std::vector<SceneDesc> desc;
    sq_triangle* d_tris = nullptr;
    bool own_tris = false;
    std::vector<void*> allocs;  // device allocations owned by the batch
    uint32_t* d_off = nullptr;
    SceneDesc* d_desc = nullptr;
    uint32_t* d_ord = nullptr;
    void* d_keys[2] = {nullptr, nullptr};
    uint32_t* d_vals[2] = {nullptr, nullptr};
    uint32_t* d_hist = nullptr;
    size_t hist_bytes = 0;
    uint32_t* d_os = nullptr;        // single-pass sort status area (sort_onesweep_bytes)
    size_t os_bytes = 0;
    uint32_t* h_sort_err = nullptr;  // pinned: set by a failed look-back wait of the last build
    void* d_ws = nullptr;
    size_t ws_bytes = 0;
    uint32_t* d_depth = nullptr;
    sq_node* d_nodes = nullptr;
    std::vector<sq_scene> scenes;
    std::vector<sq_bvh> bvhs;
    // build configuration of the last build / graph
    sq_builder kind = SQ_BUILDER_MORTON63;
    uint32_t bits = 21, digit = 8, layout = SQ_LAYOUT_REFERENCE, threshold = 0, num_bins = 8;
    bool sah = false;          // binned-SAH trees (built eagerly, outside the graph)
    uint32_t sah_depth = 0;    // depth bound of the last SAH build
    bool built = false;
    // trace configuration
    bool cams_set = false;
    std::vector<sq_camera> cams;
    sq_trace_params tparams{};
    uint32_t outputs = 0;
    std::vector<TraceArgs> jobs;
    std::vector<uint32_t> job_tiles;
    std::vector<size_t> ray_off, pix_off;
    TraceArgs* d_jobs = nullptr;
    uint32_t* d_job_tiles = nullptr;
    uint32_t* d_hit_id = nullptr;
    float* d_hit_t = nullptr;
    uint32_t* d_counts = nullptr;
    uint8_t* d_rgb = nullptr;
    size_t total_rays = 0, total_pix = 0;
    uint32_t* d_spill = nullptr;
    size_t spill_bytes = 0;
    uint32_t* d_work = nullptr;
    unsigned long long* d_stats = nullptr;
    uint32_t* d_job_fail = nullptr;  // fault bits per scene of the last trace
    uint32_t pk_grid = 0;            // workgroups of a wave-packet trace launch
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    float build_ms = 0.f, trace_ms = 0.f, run_ms = 0.f;
};

namespace {

// Host-to-device copy ordered on the context stream, the stream of every
// kernel that reads the data (then waits, so the host source may go away).
int upload(sq_ctx* c, void* dst, const void* src, size_t bytes) {
    SQ_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

int balloc(sq_batch* b, size_t bytes, void** out) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
    if (e != hipSuccess) SQ_RET(e == hipErrorOutOfMemory ? SQ_ERR_NOMEM : SQ_ERR_HIP, "batch alloc %zu B: %s", bytes,
                                hipGetErrorString(e));
    b->allocs.push_back(p);
    *out = p;
    return SQ_OK;
}

template <typename T>
int balloc_t(sq_batch* b, size_t count, T** out) {
    void* p = nullptr;
    SQ_TRY(balloc(b, count * sizeof(T), &p));
    *out = (T*)p;
    return SQ_OK;
}

void bfree(sq_batch* b, void* p) {
    if (!p) return;
    auto it = std::find(b->allocs.begin(), b->allocs.end(), p);
    if (it != b->allocs.end()) {
        (void)hipFree(p);
        b->allocs.erase(it);
    }
}

void drop_graph(sq_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

bool sah_kind(sq_builder kind) { return kind == SQ_BUILDER_RECURSIVE || kind == SQ_BUILDER_HORIZONTAL; }

int morton_bits(sq_builder kind, const sq_build_params* p, uint32_t* bits) {
    uint32_t d = kind == SQ_BUILDER_MORTON ? 10 : kind == SQ_BUILDER_MORTON2 ? 20 : kind == SQ_BUILDER_MORTON63 ? 21 : 0;
    if (!d) SQ_RET(SQ_ERR_UNSUPPORTED, "batches build with the Morton and binned-SAH builders (got %d)", (int)kind);
    *bits = p && p->morton_bits_per_axis ? p->morton_bits_per_axis : d;
    if (*bits < 1 || *bits > 21) SQ_RET(SQ_ERR_INVALID, "morton bits per axis %u outside 1..21", *bits);
    return SQ_OK;
}

int upload(sq_ctx* c, void* dst, const void* src, size_t bytes);

// Roots of the chosen builder (Morton trees at slot 0, binned-SAH trees at
// 2n - 1), written into the trace job table too when cameras are set.
int set_roots(sq_batch* b) {
    bool changed = false;
    for (uint32_t s = 0; s < b->nscenes; s++) {
        const uint32_t r = b->sah ? 2 * b->counts[s] - 1 : 0;
        b->bvhs[s].root = r;
        if (b->cams_set && b->jobs[s].root != r) {
            b->jobs[s].root = r;
            changed = true;
        }
    }
    if (changed) SQ_TRY(upload(b->ctx, b->d_jobs, b->jobs.data(), b->jobs.size() * sizeof(TraceArgs)));
    return SQ_OK;
}

int configure_build_kind(sq_batch* b, sq_builder kind, const sq_build_params* p);

int configure_build(sq_batch* b, sq_builder kind, const sq_build_params* p) {
    SQ_TRY(configure_build_kind(b, kind, p));
    return set_roots(b);
}

int configure_build_kind(sq_batch* b, sq_builder kind, const sq_build_params* p) {
    if (sah_kind(kind)) {
        // binned SAH (recursive; the horizontal builder makes the same tree)
        const uint32_t nb = p && p->num_bins ? p->num_bins : 8;
        if (nb < 2 || nb > 64) SQ_RET(SQ_ERR_INVALID, "num_bins %u outside 2..64", nb);
        const uint32_t layout = p ? p->layout : SQ_LAYOUT_REFERENCE;
        if (layout > SQ_LAYOUT_FIXED) SQ_RET(SQ_ERR_INVALID, "layout %u", layout);
        if (!b->sah || nb != b->num_bins || layout != b->layout || kind != b->kind) drop_graph(b);
        b->kind = kind;
        b->sah = true;
        b->num_bins = nb;
        b->layout = layout;
        return SQ_OK;
    }
    if (b->sah) drop_graph(b);
    b->sah = false;
    uint32_t bits = 0;
    SQ_TRY(morton_bits(kind, p, &bits));
    uint32_t digit = p && p->radix_digit_bits ? p->radix_digit_bits : b->digit;
    if (digit < 1 || digit > 8) SQ_RET(SQ_ERR_INVALID, "radix digit bits %u outside 1..8", digit);
    uint32_t layout = p ? p->layout : SQ_LAYOUT_REFERENCE;
    if (layout > SQ_LAYOUT_FIXED) SQ_RET(SQ_ERR_INVALID, "layout %u", layout);
    if (p && p->no_refit) SQ_RET(SQ_ERR_UNSUPPORTED, "batched builds always refit");
    uint32_t thr = p && p->subtree_threshold ? p->subtree_threshold : b->threshold;
    const bool changed = !b->built || kind != b->kind || bits != b->bits || digit != b->digit ||
                         layout != b->layout || thr != b->threshold;
    if (changed) drop_graph(b);
    if (changed || !b->d_ws) {
        // descriptors and emission workspace depend on the threshold
        lbvh_fill_descs(b->desc.data(), b->counts.data(), b->nscenes, thr);
        SQ_TRY(upload(b->ctx, b->d_desc, b->desc.data(), b->desc.size() * sizeof(SceneDesc)));
        size_t ws = lbvh_batch_workspace(b->total, b->nscenes, thr);
        if (ws > b->ws_bytes) {
            bfree(b, b->d_ws);
            b->d_ws = nullptr;
            SQ_TRY(balloc(b, ws, &b->d_ws));
            b->ws_bytes = ws;
        }
        size_t hb = sort_hist_bytes(b->total, digit);
        if (hb > b->hist_bytes) {
            bfree(b, b->d_hist);
            b->d_hist = nullptr;
            SQ_TRY(balloc_t(b, hb / 4 + 1, &b->d_hist));
            b->hist_bytes = hb;
        }
        // key passes as single-pass kernels while prefix counts fit their 30-bit status words
        if (b->total < (1u << 30)) {
            const size_t ob = sort_onesweep_bytes(b->total, 3 * bits, digit);
            if (ob > b->os_bytes) {
                bfree(b, b->d_os);
                b->d_os = nullptr;
                SQ_TRY(balloc_t(b, ob / 4 + 1, &b->d_os));
                b->os_bytes = ob;
            }
            if (!b->h_sort_err) {
                void* hp = nullptr;
                SQ_CHECK_HIP(hipHostMalloc(&hp, 64, hipHostMallocDefault));
                b->h_sort_err = (uint32_t*)hp;
                *b->h_sort_err = 0;
            }
        }
    }
    const size_t kb = bits > 10 ? 8 : 4;
    if (!b->d_keys[0]) {
        // sized for 64-bit keys so that any builder fits
        for (int i = 0; i < 2; i++) {
            SQ_TRY(balloc(b, 8ull * b->total, &b->d_keys[i]));
            SQ_TRY(balloc_t(b, b->total, &b->d_vals[i]));
        }
    }
    (void)kb;
    b->kind = kind;
    b->bits = bits;
    b->digit = digit;
    b->layout = layout;
    b->threshold = thr;
    return SQ_OK;
}

// Enqueues the whole build on the context stream (no host synchronisation).
int enqueue_build(sq_ctx* c, sq_batch* b) {
    BatchBuild j;
    j.tris = b->d_tris;
    j.total = b->total;
    j.nscenes = b->nscenes;
    j.d_off = b->d_off;
    j.d_desc = b->d_desc;
    j.d_ord = b->d_ord;
    j.keys[0] = b->d_keys[0];
    j.keys[1] = b->d_keys[1];
    j.vals[0] = b->d_vals[0];
    j.vals[1] = b->d_vals[1];
    j.hist = b->d_hist;
    j.ws = b->d_ws;
    j.d_depth = b->d_depth;
    j.nodes = b->d_nodes;
    j.bits = b->bits;
    j.digit = b->digit;
    j.layout = b->layout;
    j.threshold = b->threshold;
    if (b->d_os && b->os_bytes >= sort_onesweep_bytes(b->total, 3 * b->bits, b->digit)) {
        j.os = b->d_os;
        j.os_err = b->h_sort_err;
    }
    return enqueue_batch_build(c, c->stream, j);
}

int plan_trace(sq_ctx* c, sq_batch* b, TracePlan* L) {
    if (!b->pk_grid) b->pk_grid = trace_packet_grid(c, true);  // outside any graph capture
    uint32_t depth = morton_depth_bound(b->bits, b->max_n);
    if (b->sah && b->sah_depth > depth) depth = b->sah_depth;
    return trace_plan(c, &b->tparams, depth, false, L);
}

// A binned-SAH build of every scene (eager), with the scenes' roots and depths.
int build_sah(sq_ctx* c, sq_batch* b) {
    SQ_TRY(launch_fill_u32_on(c->stream, c->num_cus, b->d_nodes, 0u, (size_t)2 * b->total * 8));
    uint32_t depth = 0;
    SQ_TRY(sah_build_batch(c, b->d_tris, b->total, b->d_off, b->nscenes, b->num_bins, b->layout, b->d_nodes, &depth));
    b->sah_depth = depth;
    for (uint32_t s = 0; s < b->nscenes; s++) b->bvhs[s].depth = depth;
    return SQ_OK;
}

// Enqueues the batched trace (cameras, outputs and spill already set up).
int enqueue_trace(sq_ctx* c, sq_batch* b, const TracePlan& L) {
    TraceArgs a{};
    if (use_packets(&b->tparams)) {
        a.work = b->d_work;
        a.stats = b->tparams.count_work ? b->d_stats : nullptr;
        a.num_tiles = b->job_tiles.back();
        a.jobs = b->d_jobs;
        a.job_tiles = b->d_job_tiles;
        a.njobs = b->nscenes;
        a.job_fail = b->d_job_fail;
        SQ_TRY(launch_fill_u32_on(c->stream, c->num_cus, b->d_job_fail, 0u, b->nscenes));
        return trace_launch_packets(c->stream, b->pk_grid, a, true);
    }
    a.slots = L.slots;
    a.spill = b->d_spill;
    a.spill_cap = L.spill_cap;
    a.work = b->d_work;
    a.stats = b->tparams.count_work ? b->d_stats : nullptr;
    a.num_tiles = b->job_tiles.back();
    a.jobs = b->d_jobs;
    a.job_tiles = b->d_job_tiles;
    a.njobs = b->nscenes;
    a.job_fail = b->d_job_fail;
    SQ_TRY(launch_fill_u32_on(c->stream, c->num_cus, b->d_job_fail, 0u, b->nscenes));
    bool narrow = !c->wide_fetch;
    for (const TraceArgs& j : b->jobs) narrow = narrow && narrow_ok(j.node_count, j.tri_count);
    return trace_launch_plan(c->stream, L, a, true, narrow);
}

int ensure_spill(sq_batch* b, const TracePlan& L) {
    const size_t need = L.spill_words * 4;
    if (need > b->spill_bytes) {
        bfree(b, b->d_spill);
        b->d_spill = nullptr;
        SQ_TRY(balloc_t(b, L.spill_words, &b->d_spill));
        b->spill_bytes = need;
        drop_graph(b);
    }
    return SQ_OK;
}

int finish_depths(sq_ctx* c, sq_batch* b) {
    std::vector<uint32_t> depth(b->nscenes);
    SQ_CHECK_HIP(hipMemcpyAsync(depth.data(), b->d_depth, 4ull * b->nscenes, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (b->h_sort_err && *b->h_sort_err) {
        *b->h_sort_err = 0;
        SQ_RET(SQ_ERR_HIP, "batch build: a radix-sort look-back wait exceeded its spin limit");
    }
    for (uint32_t s = 0; s < b->nscenes; s++) b->bvhs[s].depth = depth[s];
    for (uint32_t s = 0; s < b->nscenes; s++)
        if (depth[s] == 0) SQ_RET(SQ_ERR_OVERFLOW, "batch build: emission workspace overflow (scene %u)", s);
    return SQ_OK;
}

int check_trace(sq_ctx* c, sq_batch* b) {
    uint32_t h[4] = {0, 0, 0, 0};
    unsigned long long st[2] = {0, 0};
    SQ_CHECK_HIP(hipMemcpyAsync(h, b->d_work, 16, hipMemcpyDeviceToHost, c->stream));
    if (b->tparams.count_work) SQ_CHECK_HIP(hipMemcpyAsync(st, b->d_stats, 16, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (b->tparams.count_work) {
        c->last.node_tests = st[0];
        c->last.tri_tests = st[1];
    }
    if (h[1]) {
        std::vector<uint32_t> f(b->nscenes);
        SQ_CHECK_HIP(hipMemcpyAsync(f.data(), b->d_job_fail, 4ull * b->nscenes, hipMemcpyDeviceToHost, c->stream));
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
        uint32_t first = 0;
        while (first < b->nscenes && !f[first]) first++;
        char what[64];
        snprintf(what, sizeof(what), "batch trace (first failing scene %u)", first);
        return trace_fault(h[1], what);
    }
    return SQ_OK;
}

float elapsed(sq_batch* b) {
    float ms = 0.f;
    (void)hipEventSynchronize(b->ev[1]);
    (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
    return ms;
}

}  // namespace

extern "C" {

int sq_batch_create(sq_ctx* c, const sq_triangle* tris, const uint32_t* counts, uint32_t nscenes, int on_device,
                    sq_batch** out) {
    if (!c || !tris || !counts || !out || nscenes == 0) SQ_RET(SQ_ERR_INVALID, "sq_batch_create: bad argument");
    *out = nullptr;
    if (nscenes > 2048) SQ_RET(SQ_ERR_INVALID, "sq_batch_create: %u scenes (at most 2048)", nscenes);
    uint64_t total = 0;
    uint32_t max_n = 0;
    for (uint32_t s = 0; s < nscenes; s++) {
        if (counts[s] == 0) SQ_RET(SQ_ERR_INVALID, "sq_batch_create: scene %u is empty", s);
        total += counts[s];
        max_n = counts[s] > max_n ? counts[s] : max_n;
    }
    if (total * 2 + 1024 >= (1ull << 31)) SQ_RET(SQ_ERR_INVALID, "sq_batch_create: %llu triangles is too many",
                                                 (unsigned long long)total);
    SQ_CHECK_HIP(hipSetDevice(c->device));
    sq_batch* b = new sq_batch();
    b->ctx = c;
    b->nscenes = nscenes;
    b->total = (uint32_t)total;
    b->max_n = max_n;
    b->counts.assign(counts, counts + nscenes);
    b->off.resize(nscenes + 1);
    b->off[0] = 0;
    for (uint32_t s = 0; s < nscenes; s++) b->off[s + 1] = b->off[s] + counts[s];
    b->desc.resize(nscenes);
    int rc = SQ_OK;
    auto fail = [&](int r) { sq_batch_free(b); return r; };
    if (on_device) {
        b->d_tris = const_cast<sq_triangle*>(tris);
    } else {
        if ((rc = balloc_t(b, total, &b->d_tris)) != SQ_OK) return fail(rc);
        b->own_tris = true;
        if ((rc = upload(c, b->d_tris, tris, total * sizeof(sq_triangle))) != SQ_OK) return fail(rc);
    }
    if ((rc = balloc_t(b, nscenes + 1, &b->d_off)) != SQ_OK) return fail(rc);
    if ((rc = balloc_t(b, nscenes, &b->d_desc)) != SQ_OK) return fail(rc);
    if ((rc = balloc_t(b, 6ull * nscenes, &b->d_ord)) != SQ_OK) return fail(rc);
    if ((rc = balloc_t(b, nscenes, &b->d_depth)) != SQ_OK) return fail(rc);
    // Morton trees: every scene's first child list is at its slot 1 and
    // scenes start at even slots, so one slot of offset aligns the lists.
    if ((rc = balloc_t(b, 2ull * total + 1, &b->d_nodes)) != SQ_OK) return fail(rc);
    b->d_nodes += 1;
    if ((rc = balloc_t(b, 16, &b->d_work)) != SQ_OK) return fail(rc);
    if ((rc = balloc_t(b, 4, &b->d_stats)) != SQ_OK) return fail(rc);
    if ((rc = balloc_t(b, nscenes, &b->d_job_fail)) != SQ_OK) return fail(rc);
    if ((rc = upload(c, b->d_off, b->off.data(), 4ull * (nscenes + 1))) != SQ_OK) return fail(rc);
    b->scenes.resize(nscenes);
    b->bvhs.resize(nscenes);
    for (uint32_t s = 0; s < nscenes; s++) {
        sq_scene& sc = b->scenes[s];
        sc.ctx = c;
        sc.d_tris = b->d_tris + b->off[s];
        sc.n = counts[s];
        sc.owned = false;
        sq_bvh& bv = b->bvhs[s];
        bv.ctx = c;
        bv.scene = &b->scenes[s];
        bv.d_nodes = b->d_nodes + 2ull * b->off[s];
        bv.capacity = 2 * counts[s];
        bv.root = 0;
        bv.view = true;
    }
    (void)hipEventCreateWithFlags(&b->ev[0], hipEventReleaseToDevice);  // timing only
    (void)hipEventCreateWithFlags(&b->ev[1], hipEventReleaseToDevice);
    *out = b;
    return SQ_OK;
}

void sq_batch_free(sq_batch* b) {
    if (!b) return;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    drop_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    for (auto& ev : b->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->h_sort_err) (void)hipHostFree(b->h_sort_err);
    delete b;
}

uint32_t sq_batch_count(const sq_batch* b) { return b ? b->nscenes : 0; }

int sq_batch_build(sq_ctx* c, sq_batch* b, sq_builder kind, const sq_build_params* p) {
    if (!c || !b || b->ctx != c) SQ_RET(SQ_ERR_INVALID, "sq_batch_build: bad argument");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    SQ_TRY(configure_build(b, kind, p));
    SQ_CHECK_HIP(hipEventRecord(b->ev[0], c->stream));
    if (b->sah) {
        SQ_TRY(build_sah(c, b));
    } else {
        SQ_TRY(enqueue_build(c, b));
    }
    SQ_CHECK_HIP(hipEventRecord(b->ev[1], c->stream));
    if (!b->sah) SQ_TRY(finish_depths(c, b));
    b->build_ms = elapsed(b);
    b->built = true;
    return SQ_OK;
}

int sq_batch_bvh(sq_batch* b, uint32_t scene, const sq_bvh** out) {
    if (!b || !out || scene >= b->nscenes) SQ_RET(SQ_ERR_INVALID, "sq_batch_bvh: bad argument");
    if (!b->built) SQ_RET(SQ_ERR_INVALID, "sq_batch_bvh: batch not built");
    *out = &b->bvhs[scene];
    return SQ_OK;
}

int sq_batch_set_cameras(sq_ctx* c, sq_batch* b, const sq_camera* cams, const sq_trace_params* p, uint32_t outputs) {
    if (!c || !b || !cams || b->ctx != c) SQ_RET(SQ_ERR_INVALID, "sq_batch_set_cameras: bad argument");
    if (p && p->heatmap != SQ_HEATMAP_RAY) SQ_RET(SQ_ERR_UNSUPPORTED, "batches trace per-ray heatmaps only");
    if (p && p->kernel > 2) SQ_RET(SQ_ERR_INVALID, "sq_batch_set_cameras: kernel %u", p->kernel);
    if (p && p->tri_test) SQ_RET(SQ_ERR_UNSUPPORTED, "batches use the reference triangle test");
    if (outputs & ~0xFu) SQ_RET(SQ_ERR_INVALID, "sq_batch_set_cameras: output mask %#x", outputs);
    SQ_CHECK_HIP(hipSetDevice(c->device));
    sq_trace_params tp{};
    if (p) tp = *p;
    TracePlan probe;
    SQ_TRY(trace_plan(c, &tp, 64, false, &probe));  // validates block size
    size_t rays = 0, pix = 0;
    std::vector<size_t> ray_off(b->nscenes + 1), pix_off(b->nscenes + 1);
    std::vector<uint32_t> job_tiles(b->nscenes + 1);
    uint64_t tiles = 0;
    // 8x8-pixel tiles: with hundreds of jobs in one launch the tiles are
    // plentiful, and 4x4 ones only add per-tile work (256 scenes: 29.9 ms
    // trace with 4x4 tiles against 23.4 ms).
    const bool quad = false;
    for (uint32_t s = 0; s < b->nscenes; s++) {
        const sq_camera& k = cams[s];
        if (k.width == 0 || k.height == 0 || k.spp == 0) SQ_RET(SQ_ERR_INVALID, "camera %u: empty image", s);
        ray_off[s] = rays;
        pix_off[s] = pix;
        job_tiles[s] = (uint32_t)tiles;
        pix += (size_t)k.width * k.height;
        rays += (size_t)k.width * k.height * k.spp;
        TraceArgs g{};
        tiles += camera_tiling(k.width, k.height, k.spp, quad, g);
        if (tiles >= (1ull << 32)) break;
    }
    ray_off[b->nscenes] = rays;
    pix_off[b->nscenes] = pix;
    if (tiles >= (1ull << 32)) SQ_RET(SQ_ERR_INVALID, "sq_batch_set_cameras: too many tiles");
    job_tiles[b->nscenes] = (uint32_t)tiles;
    drop_graph(b);
    for (void** pp : {(void**)&b->d_hit_id, (void**)&b->d_hit_t, (void**)&b->d_counts, (void**)&b->d_rgb,
                      (void**)&b->d_jobs, (void**)&b->d_job_tiles}) {
        bfree(b, *pp);
        *pp = nullptr;
    }
    b->cams_set = false;
    if (outputs & SQ_OUT_HIT_ID) SQ_TRY(balloc_t(b, rays, &b->d_hit_id));
    if (outputs & SQ_OUT_HIT_T) SQ_TRY(balloc_t(b, rays, &b->d_hit_t));
    if (outputs & SQ_OUT_COUNTS) SQ_TRY(balloc_t(b, pix, &b->d_counts));
    if (outputs & SQ_OUT_RGB) SQ_TRY(balloc_t(b, pix * 3, &b->d_rgb));
    std::vector<TraceArgs> jobs(b->nscenes);
    for (uint32_t s = 0; s < b->nscenes; s++) {
        const sq_camera& k = cams[s];
        TraceArgs& a = jobs[s];
        a = TraceArgs{};
        a.nodes = b->bvhs[s].d_nodes;
        a.tris = b->scenes[s].d_tris;
        a.root = b->bvhs[s].root;
        a.node_count = b->bvhs[s].capacity;
        a.tri_count = b->scenes[s].n;
        a.camera = 1;
        a.ox = k.origin.x; a.oy = k.origin.y; a.oz = k.origin.z;
        a.tanh = k.tan_half_fov;
        a.w = k.width; a.h = k.height; a.spp = k.spp; a.jitter = k.jitter; a.seed = k.seed;
        a.tmin = k.tmin; a.tmax = k.tmax;
        camera_tiling(k.width, k.height, k.spp, quad, a);
        a.hit_id = b->d_hit_id ? b->d_hit_id + ray_off[s] : nullptr;
        a.hit_t = b->d_hit_t ? b->d_hit_t + ray_off[s] : nullptr;
        a.counts = b->d_counts ? b->d_counts + pix_off[s] : nullptr;
        a.rgb = b->d_rgb ? b->d_rgb + 3 * pix_off[s] : nullptr;
    }
    SQ_TRY(balloc_t(b, jobs.size(), &b->d_jobs));
    SQ_TRY(balloc_t(b, job_tiles.size(), &b->d_job_tiles));
    SQ_TRY(upload(c, b->d_jobs, jobs.data(), jobs.size() * sizeof(TraceArgs)));
    SQ_TRY(upload(c, b->d_job_tiles, job_tiles.data(), job_tiles.size() * 4));
    b->jobs = std::move(jobs);
    b->job_tiles = std::move(job_tiles);
    b->ray_off = std::move(ray_off);
    b->pix_off = std::move(pix_off);
    b->cams.assign(cams, cams + b->nscenes);
    b->tparams = tp;
    b->outputs = outputs;
    b->total_rays = rays;
    b->total_pix = pix;
    b->cams_set = true;
    return SQ_OK;
}

int sq_batch_trace(sq_ctx* c, sq_batch* b) {
    if (!c || !b || b->ctx != c) SQ_RET(SQ_ERR_INVALID, "sq_batch_trace: bad argument");
    if (!b->built || !b->cams_set) SQ_RET(SQ_ERR_INVALID, "sq_batch_trace: build and set cameras first");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    TracePlan L;
    SQ_TRY(plan_trace(c, b, &L));
    SQ_TRY(ensure_spill(b, L));
    SQ_CHECK_HIP(hipEventRecord(b->ev[0], c->stream));
    SQ_TRY(enqueue_trace(c, b, L));
    SQ_CHECK_HIP(hipEventRecord(b->ev[1], c->stream));
    SQ_TRY(check_trace(c, b));
    b->trace_ms = elapsed(b);
    return SQ_OK;
}

int sq_batch_run(sq_ctx* c, sq_batch* b, sq_builder kind, const sq_build_params* p) {
    if (!c || !b || b->ctx != c) SQ_RET(SQ_ERR_INVALID, "sq_batch_run: bad argument");
    if (!b->cams_set) SQ_RET(SQ_ERR_INVALID, "sq_batch_run: set cameras first");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    SQ_TRY(configure_build(b, kind, p));
    TracePlan L;
    SQ_TRY(plan_trace(c, b, &L));
    SQ_TRY(ensure_spill(b, L));
    if (b->sah) {
        // binned SAH waits on the host per level: the build runs eagerly and
        // the graph holds the trace alone
        SQ_CHECK_HIP(hipEventRecord(b->ev[0], c->stream));
        SQ_TRY(build_sah(c, b));
        b->built = true;
        SQ_TRY(plan_trace(c, b, &L));
        SQ_TRY(ensure_spill(b, L));
    }
    if (!b->graph_valid) {
        drop_graph(b);
        SQ_CHECK_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = b->sah ? SQ_OK : enqueue_build(c, b);
        if (rc == SQ_OK) rc = enqueue_trace(c, b, L);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != SQ_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) SQ_RET(SQ_ERR_HIP, "batch graph capture: %s", hipGetErrorString(e));
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            drop_graph(b);
            SQ_RET(SQ_ERR_HIP, "batch graph instantiate: %s", hipGetErrorString(e));
        }
        b->graph_valid = true;
    }
    if (!b->sah) SQ_CHECK_HIP(hipEventRecord(b->ev[0], c->stream));
    SQ_CHECK_HIP(hipGraphLaunch(b->exec, c->stream));
    SQ_CHECK_HIP(hipEventRecord(b->ev[1], c->stream));
    b->built = true;
    if (!b->sah) SQ_TRY(finish_depths(c, b));
    SQ_TRY(check_trace(c, b));
    b->run_ms = elapsed(b);
    return SQ_OK;
}

int sq_batch_outputs(const sq_batch* b, uint32_t s, sq_trace_out* o) {
    if (!b || !o || s >= b->nscenes || !b->cams_set) SQ_RET(SQ_ERR_INVALID, "sq_batch_outputs: bad argument");
    o->hit_id = b->d_hit_id ? b->d_hit_id + b->ray_off[s] : nullptr;
    o->hit_t = b->d_hit_t ? b->d_hit_t + b->ray_off[s] : nullptr;
    o->counts = b->d_counts ? b->d_counts + b->pix_off[s] : nullptr;
    o->rgb = b->d_rgb ? b->d_rgb + 3 * b->pix_off[s] : nullptr;
    return SQ_OK;
}

int sq_batch_download(sq_ctx* c, const sq_batch* b, uint32_t s, const sq_trace_out* h) {
    if (!c || !b || !h || s >= b->nscenes || !b->cams_set) SQ_RET(SQ_ERR_INVALID, "sq_batch_download: bad argument");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    const size_t nr = b->ray_off[s + 1] - b->ray_off[s], np = b->pix_off[s + 1] - b->pix_off[s];
    sq_trace_out d;
    SQ_TRY(sq_batch_outputs(b, s, &d));
    auto get = [&](void* dst, const void* src, size_t bytes, const char* what) -> int {
        if (!dst) return SQ_OK;
        if (!src) SQ_RET(SQ_ERR_INVALID, "sq_batch_download: %s was not requested", what);
        SQ_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        return SQ_OK;
    };
    SQ_TRY(get(h->hit_id, d.hit_id, nr * 4, "hit_id"));
    SQ_TRY(get(h->hit_t, d.hit_t, nr * 4, "hit_t"));
    SQ_TRY(get(h->counts, d.counts, np * 4, "counts"));
    SQ_TRY(get(h->rgb, d.rgb, np * 3, "rgb"));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

int sq_batch_times(const sq_batch* b, float* build_ms, float* trace_ms, float* run_ms) {
    if (!b) SQ_RET(SQ_ERR_INVALID, "sq_batch_times: NULL batch");
    if (build_ms) *build_ms = b->build_ms;
    if (trace_ms) *trace_ms = b->trace_ms;
    if (run_ms) *run_ms = b->run_ms;
    return SQ_OK;
}

int sq_batch_autotune(sq_ctx* c, sq_batch* b, sq_builder kind, uint32_t repeats, sq_tune_result* out) {
    if (!c || !b || !out || b->ctx != c) SQ_RET(SQ_ERR_INVALID, "sq_batch_autotune: bad argument");
    if (!b->cams_set) SQ_RET(SQ_ERR_INVALID, "sq_batch_autotune: set cameras first");
    if (repeats == 0) repeats = 3;
    *out = sq_tune_result{};
    sq_build_params bp{};
    auto time_build = [&](uint32_t digit, uint32_t thr, float* best) -> int {
        bp.radix_digit_bits = digit;
        bp.subtree_threshold = thr;
        bp.layout = b->layout;
        float m = 1e30f;
        for (uint32_t r = 0; r < repeats + 1; r++) {
            SQ_TRY(sq_batch_build(c, b, kind, &bp));
            if (r > 0) m = std::min(m, b->build_ms);  // first run warms up
        }
        *best = m;
        out->candidates_tried++;
        return SQ_OK;
    };
    float best_b = 1e30f;
    uint32_t bd = 8, bt = 2048, bn = 0;
    if (sah_kind(kind)) {
        // the bin count shapes the trees: each candidate is built and traced
        // with the batch's trace parameters and the fastest sum kept
        float best_sum = 1e30f;
        for (uint32_t nb : {8u, 16u, 32u}) {
            bp.num_bins = nb;
            bp.layout = b->layout;
            float bms = 1e30f, tms = 1e30f;
            for (uint32_t r = 0; r < repeats + 1; r++) {
                SQ_TRY(sq_batch_build(c, b, kind, &bp));
                SQ_TRY(sq_batch_trace(c, b));
                if (r > 0) {
                    bms = std::min(bms, b->build_ms);
                    tms = std::min(tms, b->trace_ms);
                }
            }
            out->candidates_tried++;
            if (nb == 8) out->default_build_ms = bms;
            if (bms + tms < best_sum) { best_sum = bms + tms; best_b = bms; bn = nb; }
        }
        bp.num_bins = bn;
        bd = 0;
        bt = 0;
    } else {
        const uint32_t digits[] = {8, 7, 6, 5, 4};
        const uint32_t thresholds[] = {2048, 1024, 512};
        for (uint32_t d : digits)
            for (uint32_t t : thresholds) {
                float ms = 0;
                SQ_TRY(time_build(d, t, &ms));
                if (d == 8 && t == 2048) out->default_build_ms = ms;
                if (ms < best_b) { best_b = ms; bd = d; bt = t; }
            }
        bp.radix_digit_bits = bd;
        bp.subtree_threshold = bt;
    }
    SQ_TRY(sq_batch_build(c, b, kind, &bp));
    const sq_trace_params base = b->tparams;
    const std::vector<sq_camera> cams = b->cams;
    const uint32_t outputs = b->outputs;
    const uint32_t slots[] = {16, 8, 32};
    // (workgroup size, workgroups per CU; 0 = the default, 16 waves per CU)
    const uint32_t shapes[][2] = {{64, 0}, {64, 8}, {64, 12}, {64, 24}, {128, 0}};
    float best_t = 1e30f;
    uint32_t bs = 16, bb = 64, bw = 0, bk = SQ_KERNEL_RAYS, bl4 = 0;
    auto time_trace = [&](uint32_t kernel, uint32_t sl, uint32_t bl, uint32_t bpc, uint32_t lists,
                          float* best) -> int {
        sq_trace_params tp = base;
        tp.kernel = kernel;
        tp.stack_slots = sl;
        tp.block_size = bl;
        tp.blocks_per_cu = bpc;
        tp.lists_per_step = lists;
        SQ_TRY(sq_batch_set_cameras(c, b, cams.data(), &tp, outputs));
        float m = 1e30f;
        for (uint32_t r = 0; r < repeats + 1; r++) {
            SQ_TRY(sq_batch_trace(c, b));
            if (r > 0) m = std::min(m, b->trace_ms);
        }
        out->candidates_tried++;
        *best = m;
        return SQ_OK;
    };
    for (uint32_t sl : slots)
        for (const auto& sh : shapes) {
            float m = 0;
            SQ_TRY(time_trace(SQ_KERNEL_RAYS, sl, sh[0], sh[1], 2, &m));
            if (sl == 16 && sh[0] == 64 && sh[1] == 0) out->default_trace_ms = m;
            if (m < best_t) { best_t = m; bs = sl; bb = sh[0]; bw = sh[1]; bk = SQ_KERNEL_RAYS; bl4 = 2; }
        }
    if (base.semantics == SQ_TRACE_EXHAUSTIVE) {
        // four lists per step (two waves per SIMD at most)
        for (uint32_t sl : {32u, 16u}) {
            float m = 0;
            SQ_TRY(time_trace(SQ_KERNEL_RAYS, sl, 64, 8, 4, &m));
            if (m < best_t) { best_t = m; bs = sl; bb = 64; bw = 8; bk = SQ_KERNEL_RAYS; bl4 = 4; }
        }
        // wave packets: exhaustive semantics only
        float m = 0;
        SQ_TRY(time_trace(SQ_KERNEL_PACKETS, 0, 0, 0, 0, &m));
        if (m < best_t) { best_t = m; bk = SQ_KERNEL_PACKETS; bl4 = 0; }
    }
    sq_trace_params tp = base;
    tp.kernel = bk;
    tp.stack_slots = bs;
    tp.block_size = bb;
    tp.blocks_per_cu = bw;
    tp.lists_per_step = bl4;
    SQ_TRY(sq_batch_set_cameras(c, b, cams.data(), &tp, outputs));
    out->kernel = bk;
    out->lists_per_step = bl4;
    out->radix_digit_bits = bd;
    out->subtree_threshold = bt;
    out->num_bins = bn;
    out->stack_slots = bs;
    out->block_size = bb;
    out->blocks_per_cu = bw;
    out->build_ms = best_b;
    out->trace_ms = best_t;
    return SQ_OK;
}

}  // extern "C"
