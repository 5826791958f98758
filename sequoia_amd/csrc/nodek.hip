// nodek.hip — wide nodes (Node<K>, aabb.h:26-32) for K = 4 and 8:
// conversion of a device Node1 tree, and the tree's entry points.
//
// Node<K> keeps the reference's structure-of-arrays layout (vec3f<K> min,
// simdi<K> flags, vec3f<K> max, simdi<K> child): 8 K 32-bit words, min x[K],
// min y[K], min z[K], flags[K], max x[K], max y[K], max z[K], child[K]. The
// reference never builds or reads it, so its meaning is defined here:
//   * slot j is one child: a Node1 node (its box as stored there) or a run of
//     a Node1 list (the union of the run's boxes); flags[j] follows Node1's
//     count | flags << 16 word;
//   * a leaf slot copies a Node1 leaf (bit 16 set, count triangles from
//     child[j], the Node1 tree's triangle ids);
//   * an interior slot names wide node child[j] and holds how many of its
//     slots are used (1..K, used slots first); a Node1 node with an empty
//     child list keeps a slot with word 0 and child 0xFFFFFFFF;
//   * unused slots: min +inf, max -inf, word 0, child 0xFFFFFFFF.
// Conversion: a wide node holds a list (f, c) of Node1 nodes (the child list
// of an interior node, or a run). Its slots start as the list's nodes
// (c <= K) or as K - m single nodes and m runs over the rest
// (m = min(K, ceil((c - K) / (K - 1))), the first runs one node longer where
// the rest does not divide evenly). Then, unless the caller asks for no
// collapse, while some slot can be replaced in place by what it stands for
// (an interior node's child list, a run's nodes) without exceeding K slots,
// the one with the largest half surface area dx*dy + dy*dz + dz*dx (plain
// float operations in that order; the first slot on ties) is. Nodes are
// numbered breadth first: level by level, a level's nodes in the order of the
// slots that name them; a slot's used count is written by the level that
// builds the node it names. One level is three launches: collapse the
// lists, scan the interior-slot counts, emit.
#include "internal.h"

#include <cstring>

namespace sq {

int grow_copy(hipStream_t stream, void** p, size_t* cap, size_t want, size_t keep, size_t esz);

namespace {

constexpr uint32_t kNkBlock = 256;
constexpr uint32_t kNkScanBlock = 1024;
constexpr uint32_t kUnused = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x10000u;
constexpr uint32_t kMaxLevels = 4096;
constexpr uint32_t kNoParent = 0xFFFFFFFFu;

struct NkItem {
    uint32_t f, c;       // the Node1 list the wide node holds
    uint32_t parent;     // wide node whose slot names it (kNoParent: the root record)
    uint32_t slot;
};

struct NkSlot {
    uint32_t f, c;  // the Node1 run it stands for (c == 1: that node)
    float mn[3], mx[3];
};

struct NkArgs {
    const sq_node* nodes;
    uint32_t node_count;
    uint32_t K;
    uint32_t expand;
    const NkItem* items;
    uint32_t nitems;
    uint32_t base;       // wide index of items[0]
    NkSlot* slots;       // K per item
    uint32_t* used;      // per item
    uint32_t* cnt;       // per item: interior slots naming a node
    uint32_t* off;       // exclusive scan of cnt within its scan block
    uint32_t* block_tot; // per scan block, then the exclusive scan of those
    NkItem* next;
    uint32_t next_cap;
    uint32_t* out;       // 8 K words per node
    uint32_t node_cap;
    uint32_t* counters;  // [0] next items, [2] error code
    sq_nodek_root* root;
};

__device__ __forceinline__ float nk_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float nk_max(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ bool run_ok(uint32_t f, uint32_t c, uint32_t n) { return c != 0 && f < n && c <= n - f; }

__device__ __forceinline__ NkSlot make_slot(const sq_node* nodes, uint32_t f, uint32_t c) {
    NkSlot s;
    s.f = f;
    s.c = c;
    const sq_node a = nodes[f];
    s.mn[0] = a.min.x; s.mn[1] = a.min.y; s.mn[2] = a.min.z;
    s.mx[0] = a.max.x; s.mx[1] = a.max.y; s.mx[2] = a.max.z;
    for (uint32_t i = 1; i < c; i++) {
        const sq_node b = nodes[f + i];
        const float bn[3] = {b.min.x, b.min.y, b.min.z}, bx[3] = {b.max.x, b.max.y, b.max.z};
        for (int k = 0; k < 3; k++) {
            s.mn[k] = nk_min(s.mn[k], bn[k]);
            s.mx[k] = nk_max(s.mx[k], bx[k]);
        }
    }
    return s;
}

__device__ __forceinline__ float half_area(const NkSlot& s) {
    const float dx = s.mx[0] - s.mn[0], dy = s.mx[1] - s.mn[1], dz = s.mx[2] - s.mn[2];
    const float a = dx * dy, b = dy * dz, c = dz * dx;
    return (a + b) + c;
}

// 0: leaf, 1: interior naming a node, 2: interior with an empty list.
__device__ __forceinline__ uint32_t slot_kind(const NkArgs& a, const NkSlot& s, uint32_t& lf, uint32_t& lc) {
    if (s.c != 1) {
        lf = s.f;
        lc = s.c;
        return 1;
    }
    const sq_node y = a.nodes[s.f];
    if (y.flags & 1) return 0;
    lf = y.child;
    lc = y.count;
    return lc == 0 ? 2 : 1;
}

__device__ __forceinline__ void nk_fail(const NkArgs& a, uint32_t code) { atomicMax(&a.counters[2], code); }

__global__ void k_nk_root(NkArgs a, uint32_t root, NkItem* items) {
    if (threadIdx.x != 0) return;
    a.counters[0] = 0;
    a.counters[2] = 0;
    if (root >= a.node_count) { a.counters[2] = 1; return; }
    const sq_node R = a.nodes[root];
    sq_nodek_root r{};
    r.min[0] = R.min.x; r.min[1] = R.min.y; r.min[2] = R.min.z;
    r.max[0] = R.max.x; r.max[1] = R.max.y; r.max[2] = R.max.z;
    if (R.flags & 1) {
        r.kind = 1;
        r.child = R.child;
        r.count = R.count;
    } else if (R.count != 0) {
        if (!run_ok(R.child, R.count, a.node_count)) { a.counters[2] = 1; return; }
        r.kind = 2;
        r.child = 0;
        items[0] = NkItem{R.child, R.count, kNoParent, 0};
        a.counters[0] = 1;
    }
    *a.root = r;
}

// Collapses each item's list into its slots; counts the slots that name a node.
__global__ __launch_bounds__(kNkBlock) void k_nk_collapse(NkArgs a) {
    const uint32_t i = blockIdx.x * kNkBlock + threadIdx.x;
    if (i >= a.nitems) return;
    const NkItem it = a.items[i];
    const uint32_t K = a.K;
    NkSlot* sl = a.slots + (size_t)i * K;
    uint32_t n = 0;
    bool bad = !run_ok(it.f, it.c, a.node_count);
    if (!bad) {
        if (it.c <= K) {
            for (uint32_t j = 0; j < it.c; j++) sl[j] = make_slot(a.nodes, it.f + j, 1);
            n = it.c;
        } else {
            uint32_t m = (it.c - K + K - 2) / (K - 1);
            if (m > K) m = K;
            const uint32_t s = K - m, q = (it.c - s) / m, r = (it.c - s) % m;
            for (uint32_t j = 0; j < s; j++) sl[j] = make_slot(a.nodes, it.f + j, 1);
            uint32_t at = it.f + s;
            for (uint32_t j = 0; j < m; j++) {
                const uint32_t len = q + (j < r ? 1u : 0u);
                sl[s + j] = make_slot(a.nodes, at, len);
                at += len;
            }
            n = K;
        }
    }
    while (!bad && a.expand) {
        int best = -1;
        float best_a = 0.0f;
        uint32_t bf = 0, bc = 0;
        for (uint32_t j = 0; j < n; j++) {
            uint32_t lf = 0, lc = 0;
            const NkSlot s = sl[j];
            if (slot_kind(a, s, lf, lc) != 1) continue;
            if (!run_ok(lf, lc, a.node_count)) { bad = true; break; }
            if (n - 1 + lc > K) continue;
            const float ar = half_area(s);
            if (best < 0 || ar > best_a) { best = (int)j; best_a = ar; bf = lf; bc = lc; }
        }
        if (bad || best < 0) break;
        for (uint32_t j = n; j-- > (uint32_t)best + 1;) sl[j + bc - 1] = sl[j];
        for (uint32_t j = 0; j < bc; j++) sl[best + j] = make_slot(a.nodes, bf + j, 1);
        n += bc - 1;
    }
    uint32_t named = 0;
    if (bad) {
        nk_fail(a, 1);
        n = 0;
    }
    for (uint32_t j = 0; j < n; j++) {
        uint32_t lf, lc;
        if (slot_kind(a, sl[j], lf, lc) == 1) named++;
    }
    a.used[i] = n;
    a.cnt[i] = named;
}

// Exclusive scan of cnt per block of kNkScanBlock items (block totals kept).
__global__ __launch_bounds__(kNkScanBlock) void k_nk_scan_block(NkArgs a) {
    __shared__ uint32_t s_w[kNkScanBlock / 64];
    const uint32_t i = blockIdx.x * kNkScanBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t v = i < a.nitems ? a.cnt[i] : 0u;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    uint32_t b = 0, t = 0;
    for (uint32_t w = 0; w < kNkScanBlock / 64; w++) {
        if (w < wave) b += s_w[w];
        t += s_w[w];
    }
    if (i < a.nitems) a.off[i] = b + x - v;
    if (threadIdx.x == 0) a.block_tot[blockIdx.x] = t;
}

// Exclusive scan of the block totals (one workgroup, serial over blocks);
// counters[0] = the next level's items.
__global__ void k_nk_scan_totals(NkArgs a, uint32_t nblocks) {
    if (threadIdx.x != 0) return;
    uint32_t run = 0;
    for (uint32_t b = 0; b < nblocks; b++) {
        const uint32_t t = a.block_tot[b];
        a.block_tot[b] = run;
        run += t;
    }
    a.counters[0] = run;
}

__global__ __launch_bounds__(kNkBlock) void k_nk_emit(NkArgs a) {
    const uint32_t i = blockIdx.x * kNkBlock + threadIdx.x;
    if (i >= a.nitems) return;
    const uint32_t K = a.K, W = 8 * K;
    const NkItem it = a.items[i];
    const uint32_t n = a.used[i];
    const uint32_t w = a.base + i;
    if (w >= a.node_cap) { nk_fail(a, 2); return; }
    if (it.parent == kNoParent) a.root->count = n;
    else if (it.parent < a.node_cap && it.slot < K) a.out[(size_t)it.parent * W + 3 * K + it.slot] = n;
    else nk_fail(a, 2);
    uint32_t* o = a.out + (size_t)w * W;
    const NkSlot* sl = a.slots + (size_t)i * K;
    uint32_t rank = a.off[i] + a.block_tot[i / kNkScanBlock];
    for (uint32_t j = 0; j < K; j++) {
        if (j >= n) {
            for (int k = 0; k < 3; k++) {
                o[k * K + j] = 0x7f800000u;
                o[(4 + k) * K + j] = 0xff800000u;
            }
            o[3 * K + j] = 0;
            o[7 * K + j] = kUnused;
            continue;
        }
        const NkSlot s = sl[j];
        for (int k = 0; k < 3; k++) {
            o[k * K + j] = __float_as_uint(s.mn[k]);
            o[(4 + k) * K + j] = __float_as_uint(s.mx[k]);
        }
        uint32_t lf = 0, lc = 0;
        const uint32_t kind = slot_kind(a, s, lf, lc);
        if (kind == 0) {
            const sq_node y = a.nodes[s.f];
            o[3 * K + j] = (uint32_t)y.count | ((uint32_t)y.flags << 16);
            o[7 * K + j] = y.child;
        } else if (kind == 2) {
            o[3 * K + j] = 0;
            o[7 * K + j] = kUnused;
        } else {
            o[3 * K + j] = 0;  // written by the next level
            o[7 * K + j] = a.base + a.nitems + rank;
            if (rank < a.next_cap) a.next[rank] = NkItem{lf, lc, w, j};
            else nk_fail(a, 2);
            rank++;
        }
    }
}

int nk_error(uint32_t code) {
    if (code == 1) SQ_RET(SQ_ERR_INVALID, "wide nodes: the Node1 tree has an index out of range");
    SQ_RET(SQ_ERR_HIP, "wide nodes: conversion capacity exceeded (internal error)");
}

}  // namespace

// Node1 tree -> Node<K> tree on the device, level by level; the node array
// grows as levels arrive (a reference-layout tree can reach a subtree through
// two paths, which then gets a wide node per path).
int nodek_convert(sq_ctx* c, const sq_bvh* b, uint32_t K, bool expand, sq_bvhk* out) {
    size_t node_cap = b->capacity / 2 + 16;
    void* pn = nullptr;
    SQ_CHECK_HIP(hipMalloc(&pn, node_cap * 32 * K));
    out->d_nodes = (uint32_t*)pn;
    SQ_TRY(c->counters.ensure(64));
    SQ_TRY(c->aux[0].ensure(sizeof(NkItem) * 4));
    SQ_TRY(pinned_ensure(c, 64 + sizeof(sq_nodek_root)));
    SQ_TRY(c->spill.ensure(sizeof(sq_nodek_root) + 256));
    NkArgs a{};
    a.nodes = b->d_nodes;
    a.node_count = b->capacity;
    a.K = K;
    a.expand = expand ? 1u : 0u;
    a.counters = c->counters.as<uint32_t>();
    a.root = (sq_nodek_root*)c->spill.p;
    sq::DevBuf* items[2] = {&c->aux[0], &c->aux[1]};
    uint32_t* h = (uint32_t*)c->pinned;
    hipLaunchKernelGGL(k_nk_root, dim3(1), dim3(64), 0, c->stream, a, b->root, items[0]->as<NkItem>());
    SQ_CHECK_HIP(hipGetLastError());
    SQ_CHECK_HIP(hipMemcpyAsync(h, a.counters, 12, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (h[2]) return nk_error(h[2]);
    uint32_t nitems = h[0], base = 0, levels = 0;
    int cur = 0;
    while (nitems) {
        if (++levels > kMaxLevels) SQ_RET(SQ_ERR_INVALID, "wide nodes: the Node1 tree is deeper than %u", kMaxLevels);
        const uint32_t g = (nitems + kNkBlock - 1) / kNkBlock, gs = (nitems + kNkScanBlock - 1) / kNkScanBlock;
        SQ_TRY(c->aux[2].ensure((size_t)nitems * K * sizeof(NkSlot)));
        SQ_TRY(c->keys[0].ensure((size_t)nitems * 8));
        SQ_TRY(c->keys[1].ensure((size_t)nitems * 4));
        SQ_TRY(c->hist.ensure((size_t)gs * 4 + 64));
        SQ_TRY(grow_copy(c->stream, &pn, &node_cap, (size_t)base + nitems, base, (size_t)32 * K));
        out->d_nodes = (uint32_t*)pn;
        a.out = out->d_nodes;
        a.node_cap = (uint32_t)node_cap;
        a.items = items[cur]->as<NkItem>();
        a.slots = c->aux[2].as<NkSlot>();
        a.used = c->keys[0].as<uint32_t>();
        a.cnt = a.used + nitems;
        a.off = c->keys[1].as<uint32_t>();
        a.block_tot = c->hist.as<uint32_t>();
        a.nitems = nitems;
        a.base = base;
        hipLaunchKernelGGL(k_nk_collapse, dim3(g), dim3(kNkBlock), 0, c->stream, a);
        hipLaunchKernelGGL(k_nk_scan_block, dim3(gs), dim3(kNkScanBlock), 0, c->stream, a);
        hipLaunchKernelGGL(k_nk_scan_totals, dim3(1), dim3(64), 0, c->stream, a, gs);
        SQ_CHECK_HIP(hipGetLastError());
        SQ_CHECK_HIP(hipMemcpyAsync(h, a.counters, 12, hipMemcpyDeviceToHost, c->stream));
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
        if (h[2]) return nk_error(h[2]);
        const uint32_t next_items = h[0];
        // stack entries hold node index << 3 | (used - 1)
        if ((uint64_t)base + nitems + next_items >= (1ull << 29))
            SQ_RET(SQ_ERR_UNSUPPORTED, "wide nodes: more than 2^29 nodes");
        SQ_TRY(items[cur ^ 1]->ensure((size_t)(next_items ? next_items : 1) * sizeof(NkItem)));
        a.next = items[cur ^ 1]->as<NkItem>();
        a.next_cap = next_items;
        hipLaunchKernelGGL(k_nk_emit, dim3(g), dim3(kNkBlock), 0, c->stream, a);
        SQ_CHECK_HIP(hipGetLastError());
        base += nitems;
        nitems = next_items;
        cur ^= 1;
    }
    SQ_CHECK_HIP(hipMemcpyAsync(h, a.counters, 12, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipMemcpyAsync(h + 16, a.root, sizeof(sq_nodek_root), hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (h[2]) return nk_error(h[2]);
    std::memcpy(&out->root, h + 16, sizeof(sq_nodek_root));
    out->nodes = base;
    out->levels = levels;
    return SQ_OK;
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_bvhk_create(sq_ctx* c, const sq_bvh* b, uint32_t k, uint32_t flags, sq_bvhk** out) {
    if (!c || !b || !out) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_create: NULL argument");
    if (!b->scene) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_create: the tree has no scene");
    if (k != 4 && k != 8) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_create: k = %u (4 or 8)", k);
    if (flags & ~(uint32_t)SQ_BVHK_NO_COLLAPSE) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_create: flags 0x%x", flags);
    *out = nullptr;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    sq_bvhk* t = new sq_bvhk();
    t->ctx = c;
    t->scene = b->scene;
    t->k = k;
    int rc = nodek_convert(c, b, k, !(flags & SQ_BVHK_NO_COLLAPSE), t);
    if (rc != SQ_OK) {
        sq_bvhk_free(t);
        return rc;
    }
    *out = t;
    return SQ_OK;
}

int sq_bvhk_info(const sq_bvhk* t, uint32_t* k, uint32_t* nodes, uint32_t* levels, sq_nodek_root* root) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_info: NULL tree");
    if (k) *k = t->k;
    if (nodes) *nodes = t->nodes;
    if (levels) *levels = t->levels;
    if (root) *root = t->root;
    return SQ_OK;
}

int sq_bvhk_download(const sq_bvhk* t, void* nodes, uint32_t node_cap) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_download: NULL tree");
    if (!nodes || !t->nodes) return SQ_OK;
    if (node_cap < t->nodes) SQ_RET(SQ_ERR_INVALID, "sq_bvhk_download: node capacity %u < %u", node_cap, t->nodes);
    sq_ctx* c = t->ctx;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    SQ_CHECK_HIP(hipMemcpyAsync(nodes, t->d_nodes, (size_t)t->nodes * 32 * t->k, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

void sq_bvhk_free(sq_bvhk* t) {
    if (!t) return;
    if (t->ctx) (void)hipSetDevice(t->ctx->device);
    if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
    if (t->d_nodes) (void)hipFree(t->d_nodes);
    delete t;
}

}  // extern "C"
