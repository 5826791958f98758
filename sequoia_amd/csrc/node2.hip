// node2.hip — quantised two-child nodes (Node2, aabb.h:44-74): conversion of a
// device Node1 tree, and the traversal entry points over it.
//
// Node2 keeps the reference's 32-byte field layout: a float centre, child a
// (24 bits) with the 8-bit size, child b (24 bits) with 3-bit counts and leaf
// bits, and twelve 8-bit bounds (a/b min/max per axis). The reference never
// builds or reads it, so its meaning is defined here:
//   * each Node2 describes the two children of one node; a child is a leaf
//     group (1..7 triangles, contiguous in a reordered triangle array; child =
//     first of them) or another Node2 (child = its index);
//   * size = e, the biased exponent of step = 2^(e-127): the smallest power of
//     two with 120 * step >= the largest distance from the centre to a face of
//     the node's box (t = h * (1/120), e = exponent of t, plus one unless t is
//     a power of two; e >= 1);
//   * bound q decodes to centre + (float)(q - 128) * step; a child's min is the
//     largest q whose decoded value is <= the Node1 min, its max the smallest q
//     whose decoded value is >= the Node1 max, so decoded boxes contain the
//     Node1 boxes (no hit can be lost) while a child pair costs one 32-byte
//     fetch instead of two.
// Conversion (breadth-first, deterministic): a list holding one interior
// Node1 node stands for that node's child list (keeping the outer box); a
// single leaf, or a list of leaves with at most max_leaf (1..7) triangles in
// all, becomes a leaf group; any other
// list becomes a Node2 whose children are its first c/2 nodes and the rest.
// Node2 indices follow levels, items in order, child a before b; leaf groups
// take triangle ranges in the same order, the root's group first. One level
// is three launches: classify the halves, scan the per-item counts, emit.
#include "trace_common.h"

#include <cstring>

namespace sq {

constexpr uint32_t kN2Block = 256;
constexpr uint32_t kN2ScanBlock = 1024;

struct N2Item {
    uint32_t f, c;
    float box[6];
};

struct N2Half {
    uint32_t kind, f, c, tris;  // kind 1: leaf group, 2: Node2
    float box[6];
};

struct N2Args {
    const sq_node* nodes;
    uint32_t node_count;
    const sq_triangle* tris;
    uint32_t ntris;
    const N2Item* items;
    uint32_t nitems;
    uint32_t base;       // Node2 index of items[0]
    N2Half* halves;      // 2 per item
    uint32_t* cnt;       // per item: [2i] Node2 children, [2i+1] leaf triangles
    uint32_t* off;       // exclusive scan of cnt (same layout)
    uint32_t* block_tot; // per scan block: 2 words
    N2Item* next;
    uint32_t next_base;
    uint32_t tri_base;
    sq_node2* out;
    sq_triangle* tris2;
    uint32_t* perm;
    uint32_t* counters;  // [0] next items, [1] triangles of the level, [2] error code
    uint32_t node_cap, item_cap, tri_cap;
    uint32_t max_leaf;   // largest leaf group made of several Node1 leaves (1..7)
    sq_node2_root* root;
};

__device__ __forceinline__ float n2_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float n2_max(float a, float b) { return (a < b) ? b : a; }

__device__ __forceinline__ void node_box6(const sq_node& n, float b[6]) {
    const float* mn = &n.min.x;
    const float* mx = &n.max.x;
    for (int k = 0; k < 3; k++) { b[k] = mn[k]; b[3 + k] = mx[k]; }
}

// The list (f, c) as a child (resolve() of the CPU restatement). Returns 0 or
// an error code (1: index out of range, 2: a leaf of more than 7 triangles).
__device__ uint32_t n2_resolve(const N2Args& a, uint32_t f, uint32_t c, const float box[6], N2Half& out) {
    for (uint32_t guard = 0; guard < 64 && c == 1; guard++) {
        if (f >= a.node_count) return 1;
        const sq_node x = a.nodes[f];
        if (x.flags & 1) break;
        f = x.child;
        c = x.count;
        if (c == 0) return 1;
    }
    if (c == 0 || f >= a.node_count || c > a.node_count - f) return 1;
    bool all_leaves = true;
    uint32_t tris = 0;
    for (uint32_t i = 0; i < c && all_leaves; i++) {
        const sq_node x = a.nodes[f + i];
        if (!(x.flags & 1)) all_leaves = false;
        else tris += x.count;
    }
    out.f = f;
    out.c = c;
    for (int k = 0; k < 6; k++) out.box[k] = box[k];
    if (all_leaves && tris <= (c == 1 ? 7u : a.max_leaf)) {
        out.kind = 1;
        out.tris = tris;
        return 0;
    }
    if (c == 1) return 2;
    out.kind = 2;
    out.tris = 0;
    return 0;
}

__device__ __forceinline__ void n2_fail(const N2Args& a, uint32_t code) { atomicMax(&a.counters[2], code); }

// Copies a leaf group's triangles to tris2[first ..] (original ids to perm).
__device__ uint32_t n2_copy_leaf(const N2Args& a, const N2Half& h, uint32_t first) {
    uint32_t k = first;
    for (uint32_t i = 0; i < h.c; i++) {
        const sq_node x = a.nodes[h.f + i];
        for (uint32_t j = 0; j < x.count; j++) {
            const uint32_t tid = x.child + j;
            if (tid >= a.ntris || k >= a.tri_cap) return 1;
            a.tris2[k] = a.tris[tid];
            a.perm[k] = tid;
            k++;
        }
    }
    return 0;
}

__global__ void k_n2_root(N2Args a, uint32_t root, N2Item* items) {
    if (threadIdx.x != 0) return;
    a.counters[0] = 0;
    a.counters[1] = 0;
    a.counters[2] = 0;
    if (root >= a.node_count) { a.counters[2] = 1; return; }
    float box[6];
    node_box6(a.nodes[root], box);
    N2Half h;
    const uint32_t e = n2_resolve(a, root, 1, box, h);
    if (e) { a.counters[2] = e; return; }
    sq_node2_root r{};
    for (int k = 0; k < 3; k++) { r.min[k] = box[k]; r.max[k] = box[3 + k]; }
    if (h.kind == 1) {
        r.kind = 1;
        r.child = 0;
        r.count = h.tris;
        if (n2_copy_leaf(a, h, 0)) a.counters[2] = 1;
        a.counters[1] = h.tris;
    } else {
        r.kind = 2;
        r.child = 0;
        N2Item it;
        it.f = h.f;
        it.c = h.c;
        for (int k = 0; k < 6; k++) it.box[k] = box[k];
        items[0] = it;
        a.counters[0] = 1;
    }
    *a.root = r;
}

__global__ __launch_bounds__(kN2Block) void k_n2_classify(N2Args a) {
    const uint32_t i = blockIdx.x * kN2Block + threadIdx.x;
    if (i >= a.nitems) return;
    const N2Item it = a.items[i];
    const uint32_t h = it.c / 2;
    const uint32_t hf[2] = {it.f, it.f + h}, hc[2] = {h, it.c - h};
    uint32_t nodes = 0, tris = 0;
    for (int s = 0; s < 2; s++) {
        float b[6];
        node_box6(a.nodes[hf[s]], b);
        for (uint32_t j = 1; j < hc[s]; j++) {
            float xb[6];
            node_box6(a.nodes[hf[s] + j], xb);
            for (int k = 0; k < 3; k++) {
                b[k] = n2_min(b[k], xb[k]);
                b[3 + k] = n2_max(b[3 + k], xb[3 + k]);
            }
        }
        N2Half hv;
        const uint32_t e = n2_resolve(a, hf[s], hc[s], b, hv);
        if (e) {
            n2_fail(a, e);
            hv.kind = 1; hv.f = hf[s]; hv.c = 0; hv.tris = 0;
            for (int k = 0; k < 6; k++) hv.box[k] = b[k];
        }
        a.halves[2 * i + s] = hv;
        if (hv.kind == 2) nodes++;
        else tris += hv.tris;
    }
    a.cnt[2 * i] = nodes;
    a.cnt[2 * i + 1] = tris;
}

// Exclusive scan of (nodes, tris) pairs: per block of kN2ScanBlock items,
// then the block totals (one workgroup, serial over blocks), then the add.
__global__ __launch_bounds__(kN2ScanBlock) void k_n2_scan_block(N2Args a) {
    __shared__ uint32_t s0[kN2ScanBlock / 64], s1[kN2ScanBlock / 64];
    const uint32_t i = blockIdx.x * kN2ScanBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t v0 = i < a.nitems ? a.cnt[2 * i] : 0u, v1 = i < a.nitems ? a.cnt[2 * i + 1] : 0u;
    uint32_t x0 = v0, x1 = v1;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y0 = __shfl_up(x0, o, 64), y1 = __shfl_up(x1, o, 64);
        if (lane >= (uint32_t)o) { x0 += y0; x1 += y1; }
    }
    if (lane == 63) { s0[wave] = x0; s1[wave] = x1; }
    __syncthreads();
    uint32_t b0 = 0, b1 = 0, t0 = 0, t1 = 0;
    for (uint32_t w = 0; w < kN2ScanBlock / 64; w++) {
        if (w < wave) { b0 += s0[w]; b1 += s1[w]; }
        t0 += s0[w];
        t1 += s1[w];
    }
    if (i < a.nitems) {
        a.off[2 * i] = b0 + x0 - v0;
        a.off[2 * i + 1] = b1 + x1 - v1;
    }
    if (threadIdx.x == 0) {
        a.block_tot[2 * blockIdx.x] = t0;
        a.block_tot[2 * blockIdx.x + 1] = t1;
    }
}

__global__ void k_n2_scan_totals(N2Args a, uint32_t nblocks) {
    if (threadIdx.x != 0) return;
    uint32_t r0 = 0, r1 = 0;
    for (uint32_t b = 0; b < nblocks; b++) {
        const uint32_t t0 = a.block_tot[2 * b], t1 = a.block_tot[2 * b + 1];
        a.block_tot[2 * b] = r0;
        a.block_tot[2 * b + 1] = r1;
        r0 += t0;
        r1 += t1;
    }
    a.counters[0] = r0;
    a.counters[1] = r1;
}

__device__ __forceinline__ float n2_dec(float c, uint32_t q, float step) { return c + (float)((int)q - 128) * step; }

__device__ __forceinline__ void n2_quantise(float c, float step, float inv, float lo, float hi, uint8_t& qlo_out,
                                            uint8_t& qhi_out, bool& bad) {
    const float gl = floorf((lo - c) * inv) + 128.0f, gh = ceilf((hi - c) * inv) + 128.0f;
    int qlo = gl < 0.0f ? 0 : (gl > 255.0f ? 255 : (int)gl);
    int qhi = gh < 0.0f ? 0 : (gh > 255.0f ? 255 : (int)gh);
    while (qlo > 0 && n2_dec(c, (uint32_t)qlo, step) > lo) qlo--;
    while (qlo < 255 && n2_dec(c, (uint32_t)qlo + 1, step) <= lo) qlo++;
    while (qhi < 255 && n2_dec(c, (uint32_t)qhi, step) < hi) qhi++;
    while (qhi > 0 && n2_dec(c, (uint32_t)qhi - 1, step) >= hi) qhi--;
    if (n2_dec(c, (uint32_t)qlo, step) > lo || n2_dec(c, (uint32_t)qhi, step) < hi) bad = true;
    qlo_out = (uint8_t)qlo;
    qhi_out = (uint8_t)qhi;
}

__global__ __launch_bounds__(kN2Block) void k_n2_emit(N2Args a) {
    const uint32_t i = blockIdx.x * kN2Block + threadIdx.x;
    if (i >= a.nitems) return;
    const uint32_t blk = i / kN2ScanBlock;
    const uint32_t node_off = a.off[2 * i] + a.block_tot[2 * blk];
    const uint32_t tri_off = a.off[2 * i + 1] + a.block_tot[2 * blk + 1];
    const N2Item it = a.items[i];
    float c[3];
    bool bad = false;
    for (int k = 0; k < 3; k++) {
        if (!isfinite(it.box[k]) || !isfinite(it.box[3 + k])) bad = true;
        c[k] = (it.box[k] + it.box[3 + k]) * 0.5f;
    }
    float h = 0.0f;
    for (int k = 0; k < 3; k++) {
        h = n2_max(h, c[k] - it.box[k]);
        h = n2_max(h, it.box[3 + k] - c[k]);
    }
    const float t = h * (1.0f / 120.0f);
    const uint32_t bits = __float_as_uint(t), E = (bits >> 23) & 0xFFu, M = bits & 0x7FFFFFu;
    uint32_t e = E == 0 ? 1u : E + (M != 0 ? 1u : 0u);
    if (e > 254) { bad = true; e = 254; }
    const float step = __uint_as_float(e << 23), inv = __uint_as_float((254u - e) << 23);
    sq_node2 o{};
    for (int k = 0; k < 3; k++) o.c[k] = c[k];
    uint32_t child[2], cnt[2], leaf[2], nodes = 0, tris = 0;
    for (int s = 0; s < 2; s++) {
        const N2Half hv = a.halves[2 * i + s];
        for (int k = 0; k < 3; k++)
            n2_quantise(c[k], step, inv, hv.box[k], hv.box[3 + k], o.q[4 * k + 2 * s], o.q[4 * k + 2 * s + 1], bad);
        if (hv.kind == 2) {
            leaf[s] = 0;
            cnt[s] = 0;
            child[s] = a.next_base + node_off + nodes;
            N2Item nx;
            nx.f = hv.f;
            nx.c = hv.c;
            for (int k = 0; k < 6; k++) nx.box[k] = hv.box[k];
            const uint32_t slot = node_off + nodes;
            if (slot < a.item_cap) a.next[slot] = nx;
            else bad = true;
            nodes++;
        } else {
            leaf[s] = 1;
            cnt[s] = hv.tris;
            child[s] = a.tri_base + tri_off + tris;
            if (n2_copy_leaf(a, hv, child[s])) bad = true;
            tris += hv.tris;
        }
        if (child[s] >= (1u << 24)) bad = true;
    }
    o.a_child_size = (child[0] & 0xFFFFFFu) | (e << 24);
    o.b_child_bits = (child[1] & 0xFFFFFFu) | (cnt[0] << 24) | (leaf[0] << 27) | (cnt[1] << 28) | (leaf[1] << 31);
    if (a.base + i < a.node_cap) a.out[a.base + i] = o;
    else bad = true;
    if (bad) n2_fail(a, 3);
}

// Grows a device array to `want` elements of `esz` bytes, keeping the first
// `keep` elements (also used by the wide-node conversion).
int grow_copy(hipStream_t stream, void** p, size_t* cap, size_t want, size_t keep, size_t esz) {
    if (want <= *cap) return SQ_OK;
    size_t n = *cap + *cap / 2;
    if (n < want) n = want;
    void* q = nullptr;
    SQ_CHECK_HIP(hipMalloc(&q, n * esz));
    if (*p && keep) SQ_CHECK_HIP(hipMemcpyAsync(q, *p, keep * esz, hipMemcpyDeviceToDevice, stream));
    SQ_CHECK_HIP(hipStreamSynchronize(stream));
    if (*p) SQ_CHECK_HIP(hipFree(*p));
    *p = q;
    *cap = n;
    return SQ_OK;
}

static int n2_error(uint32_t code) {
    if (code == 2) SQ_RET(SQ_ERR_UNSUPPORTED, "node2: a Node1 leaf holds more than 7 triangles");
    if (code == 3) SQ_RET(SQ_ERR_UNSUPPORTED, "node2: a box cannot be quantised or an index exceeds 24 bits");
    SQ_RET(SQ_ERR_INVALID, "node2: the Node1 tree has an index out of range");
}

// Node1 tree -> Node2 tree on the device. A reference-layout tree can reach a
// subtree through two paths (right-singleton lists); the conversion then
// holds it twice, so the output arrays grow level by level as needed.
int node2_convert(sq_ctx* c, const sq_bvh* b, uint32_t max_leaf, sq_bvh2* out) {
    const uint32_t ntris = b->scene->n;
    size_t node_cap = b->capacity ? b->capacity : 1, tri_cap = ntris > 8 ? ntris : 8;
    SQ_CHECK_HIP(hipMalloc(&out->d_nodes, node_cap * sizeof(sq_node2)));
    SQ_CHECK_HIP(hipMalloc(&out->d_tris, tri_cap * sizeof(sq_triangle)));
    SQ_CHECK_HIP(hipMalloc(&out->d_perm, tri_cap * 4));
    SQ_TRY(c->counters.ensure(64));
    SQ_TRY(c->aux[0].ensure(sizeof(N2Item) * 4));
    SQ_TRY(pinned_ensure(c, 64 + sizeof(sq_node2_root)));
    SQ_TRY(c->spill.ensure(sizeof(sq_node2_root) + 256));
    N2Args a{};
    a.nodes = b->d_nodes;
    a.node_count = b->capacity;
    a.tris = b->scene->d_tris;
    a.ntris = ntris;
    a.counters = c->counters.as<uint32_t>();
    a.root = (sq_node2_root*)c->spill.p;
    a.out = out->d_nodes;
    a.tris2 = out->d_tris;
    a.perm = out->d_perm;
    a.node_cap = (uint32_t)node_cap;
    a.tri_cap = (uint32_t)tri_cap;
    // default 1: larger groups trade box tests for triangle tests, which cost
    // more (1M soup, closest hits: 2.11 ms at 1, 2.35 at 2, 3.33 at 7)
    a.max_leaf = max_leaf == 0 ? 1u : (max_leaf > 7 ? 7u : max_leaf);
    sq::DevBuf* items[2] = {&c->aux[0], &c->aux[1]};
    uint32_t* h = (uint32_t*)c->pinned;
    hipLaunchKernelGGL(k_n2_root, dim3(1), dim3(64), 0, c->stream, a, b->root, items[0]->as<N2Item>());
    SQ_CHECK_HIP(hipGetLastError());
    SQ_CHECK_HIP(hipMemcpyAsync(h, a.counters, 12, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipMemcpyAsync(h + 16, a.root, sizeof(sq_node2_root), hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (h[2]) return n2_error(h[2]);
    std::memcpy(&out->root, h + 16, sizeof(sq_node2_root));
    uint32_t nitems = h[0], base = 0, tri_base = h[1], levels = 1;
    int cur = 0;
    while (nitems) {
        const uint32_t g = (nitems + kN2Block - 1) / kN2Block, gs = (nitems + kN2ScanBlock - 1) / kN2ScanBlock;
        SQ_TRY(c->aux[2].ensure((size_t)nitems * 2 * sizeof(N2Half)));
        SQ_TRY(c->keys[0].ensure((size_t)nitems * 8));
        SQ_TRY(c->keys[1].ensure((size_t)nitems * 8));
        SQ_TRY(c->hist.ensure((size_t)gs * 8 + 64));
        void* pn = out->d_nodes;
        SQ_TRY(grow_copy(c->stream, &pn, &node_cap, (size_t)base + nitems, base, sizeof(sq_node2)));
        out->d_nodes = (sq_node2*)pn;
        a.out = out->d_nodes;
        a.node_cap = (uint32_t)node_cap;
        a.items = items[cur]->as<N2Item>();
        a.halves = c->aux[2].as<N2Half>();
        a.cnt = c->keys[0].as<uint32_t>();
        a.off = c->keys[1].as<uint32_t>();
        a.block_tot = c->hist.as<uint32_t>();
        a.nitems = nitems;
        a.base = base;
        a.next_base = base + nitems;
        a.tri_base = tri_base;
        hipLaunchKernelGGL(k_n2_classify, dim3(g), dim3(kN2Block), 0, c->stream, a);
        hipLaunchKernelGGL(k_n2_scan_block, dim3(gs), dim3(kN2ScanBlock), 0, c->stream, a);
        hipLaunchKernelGGL(k_n2_scan_totals, dim3(1), dim3(64), 0, c->stream, a, gs);
        SQ_CHECK_HIP(hipGetLastError());
        SQ_CHECK_HIP(hipMemcpyAsync(h, a.counters, 12, hipMemcpyDeviceToHost, c->stream));
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
        if (h[2]) return n2_error(h[2]);
        const uint32_t next_items = h[0], level_tris = h[1];
        if ((uint64_t)base + nitems + next_items >= (1ull << 24) || (uint64_t)tri_base + level_tris >= (1ull << 24))
            SQ_RET(SQ_ERR_UNSUPPORTED, "node2: more than 2^24 nodes or triangles");
        SQ_TRY(items[cur ^ 1]->ensure((size_t)(next_items ? next_items : 1) * sizeof(N2Item)));
        void* pt = out->d_tris;
        void* pp = out->d_perm;
        size_t tcap2 = tri_cap;
        SQ_TRY(grow_copy(c->stream, &pt, &tri_cap, (size_t)tri_base + level_tris, tri_base, sizeof(sq_triangle)));
        SQ_TRY(grow_copy(c->stream, &pp, &tcap2, (size_t)tri_base + level_tris, tri_base, 4));
        out->d_tris = (sq_triangle*)pt;
        out->d_perm = (uint32_t*)pp;
        a.tris2 = out->d_tris;
        a.perm = out->d_perm;
        a.tri_cap = (uint32_t)tri_cap;
        a.next = items[cur ^ 1]->as<N2Item>();
        a.item_cap = next_items;
        hipLaunchKernelGGL(k_n2_emit, dim3(g), dim3(kN2Block), 0, c->stream, a);
        SQ_CHECK_HIP(hipGetLastError());
        SQ_CHECK_HIP(hipMemcpyAsync(h, a.counters, 12, hipMemcpyDeviceToHost, c->stream));
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
        if (h[2]) return n2_error(h[2]);
        base += nitems;
        tri_base += level_tris;
        nitems = next_items;
        cur ^= 1;
        levels++;
    }
    out->nodes = base;
    out->tris = tri_base;
    out->levels = levels;
    return SQ_OK;
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_bvh2_create(sq_ctx* c, const sq_bvh* b, uint32_t max_leaf, sq_bvh2** out) {
    if (!c || !b || !out) SQ_RET(SQ_ERR_INVALID, "sq_bvh2_create: NULL argument");
    if (!b->scene) SQ_RET(SQ_ERR_INVALID, "sq_bvh2_create: the tree has no scene");
    *out = nullptr;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    sq_bvh2* t = new sq_bvh2();
    t->ctx = c;
    int rc = node2_convert(c, b, max_leaf, t);
    if (rc != SQ_OK) {
        sq_bvh2_free(t);
        return rc;
    }
    *out = t;
    return SQ_OK;
}

int sq_bvh2_info(const sq_bvh2* t, uint32_t* nodes, uint32_t* tris, uint32_t* levels, sq_node2_root* root) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_bvh2_info: NULL tree");
    if (nodes) *nodes = t->nodes;
    if (tris) *tris = t->tris;
    if (levels) *levels = t->levels;
    if (root) *root = t->root;
    return SQ_OK;
}

int sq_bvh2_download(const sq_bvh2* t, sq_node2* nodes, uint32_t node_cap, sq_triangle* tris, uint32_t* perm,
                     uint32_t tri_cap) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_bvh2_download: NULL tree");
    sq_ctx* c = t->ctx;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    if (nodes && t->nodes) {
        if (node_cap < t->nodes) SQ_RET(SQ_ERR_INVALID, "sq_bvh2_download: node capacity %u < %u", node_cap, t->nodes);
        SQ_CHECK_HIP(hipMemcpyAsync(nodes, t->d_nodes, (size_t)t->nodes * sizeof(sq_node2), hipMemcpyDeviceToHost,
                                    c->stream));
    }
    if ((tris || perm) && t->tris) {
        if (tri_cap < t->tris) SQ_RET(SQ_ERR_INVALID, "sq_bvh2_download: triangle capacity %u < %u", tri_cap, t->tris);
        if (tris)
            SQ_CHECK_HIP(hipMemcpyAsync(tris, t->d_tris, (size_t)t->tris * sizeof(sq_triangle), hipMemcpyDeviceToHost,
                                        c->stream));
        if (perm)
            SQ_CHECK_HIP(hipMemcpyAsync(perm, t->d_perm, (size_t)t->tris * 4, hipMemcpyDeviceToHost, c->stream));
    }
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

void sq_bvh2_free(sq_bvh2* t) {
    if (!t) return;
    if (t->ctx) (void)hipSetDevice(t->ctx->device);
    if (t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
    if (t->d_nodes) (void)hipFree(t->d_nodes);
    if (t->d_tris) (void)hipFree(t->d_tris);
    if (t->d_perm) (void)hipFree(t->d_perm);
    delete t;
}

}  // extern "C"
