// oracle.cpp — CPU restatement of Sequoia's BVH builders and tracers.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path. Built with
// -ffp-contract=off; every fused multiply-add below is an explicit fmaf that
// reproduces the contraction g++ 11.4 -O3 chose for the reference source
// expression at that site (sites listed where they occur).
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Scalar min/max exactly as the reference spells them.
//   std::min(a,b) == (b < a) ? b : a          (vec3.h:179 via std::min)
//   std::max(a,b) == (a < b) ? b : a          (vec3.h:180 via std::max)
//   simd max      == a > b ? a : b            (simd_generic.h:170-174)
//   simd min      == std::min                  (simd_generic.h:164-168)
// They differ from IEEE fminf/fmaxf for NaN and for +-0 ties.
// ---------------------------------------------------------------------------
inline float std_min(float a, float b) { return (b < a) ? b : a; }
inline float std_max(float a, float b) { return (a < b) ? b : a; }
inline float simd_max(float a, float b) { return (a > b) ? a : b; }
inline float simd_min(float a, float b) { return (b < a) ? b : a; }

struct Box {
    float mn[3], mx[3];
    void reset() { for (int k = 0; k < 3; k++) { mn[k] = FLT_MAX; mx[k] = -FLT_MAX; } }
    // AABB1::grow / combine (aabb.h:113-117)
    void grow(const Box& a) {
        for (int k = 0; k < 3; k++) { mn[k] = std_min(mn[k], a.mn[k]); mx[k] = std_max(mx[k], a.mx[k]); }
    }
    void grow_pt(const float p[3]) {
        for (int k = 0; k < 3; k++) { mn[k] = std_min(mn[k], p[k]); mx[k] = std_max(mx[k], p[k]); }
    }
};

inline const float* vp(const sqo_vec3& v) { return &v.x; }

// AABB1(const Triangle&) / setup(): min(min(v0,v1),v2), max(max(v0,v1),v2)  (aabb.h:107-110)
inline Box tri_box(const sqo_tri& t) {
    Box b;
    const float *a = vp(t.v0), *c = vp(t.v1), *d = vp(t.v2);
    for (int k = 0; k < 3; k++) {
        b.mn[k] = std_min(std_min(a[k], c[k]), d[k]);
        b.mx[k] = std_max(std_max(a[k], c[k]), d[k]);
    }
    return b;
}

// AABB1::sa / AABBh::sa / AABBm::sa (aabb.h:119-123,141-144,166-170).
// g++ emits fma(ly,lz, fma(lx,ly, lx*lz)) in choose_plane, choose_plane_sse
// and choose_plane_m128 (builder.cpp:507,558,589); AABB1 adds the 2x.
inline float box_sa(const Box& b, uint32_t sa_kind) {
    float lx = b.mx[0] - b.mn[0], ly = b.mx[1] - b.mn[1], lz = b.mx[2] - b.mn[2];
    float s = fmaf(ly, lz, fmaf(lx, ly, lx * lz));
    return sa_kind == SQO_SA_SCALAR ? s + s : s;
}

inline void write_leaf(sqo_node& n, const Box& b, uint32_t child) {
    for (int k = 0; k < 3; k++) { n.min[k] = b.mn[k]; n.max[k] = b.mx[k]; }
    n.child = child; n.count = 1; n.flags = 1;
}
inline void write_inner(sqo_node& n, const Box& b, uint32_t child, uint32_t count) {
    for (int k = 0; k < 3; k++) { n.min[k] = b.mn[k]; n.max[k] = b.mx[k]; }
    n.child = child; n.count = (uint16_t)count; n.flags = 0;
}

// pdep-style interleave: coordinate bit i -> code bit 3i + axis, bits that
// fall past the word are dropped as pdep drops them (algorithms.cpp:8-22).
template <typename T>
T interleave(T x, T y, T z) {
    const int W = (int)sizeof(T) * 8;
    const T v[3] = {x, y, z};
    T r = 0;
    for (int a = 0; a < 3; a++)
        for (int i = 0; 3 * i + a < W; i++) r |= ((v[a] >> i) & (T)1) << (3 * i + a);
    return r;
}

// float -> unsigned conversion as the x86-64 build performs it for values the
// Morton quantiser can produce: truncation; NaN (zero-extent axis) -> 0.
inline uint64_t quantise(float v) {
    if (!(v == v)) return 0;
    if (v <= 0.0f) return 0;
    return (uint64_t)v;
}

}  // namespace

extern "C" {

uint32_t sqo_interleave32(uint32_t x, uint32_t y, uint32_t z) { return interleave<uint32_t>(x, y, z); }
uint64_t sqo_interleave64(uint64_t x, uint64_t y, uint64_t z) { return interleave<uint64_t>(x, y, z); }

// Serial scene AABB over per-triangle boxes (builder.cpp:2208-2215).
void sqo_scene_bounds(const sqo_tri* tris, uint32_t n, float out_min[3], float out_max[3]) {
    Box s; s.reset();
    for (uint32_t i = 0; i < n; i++) s.grow(tri_box(tris[i]));
    for (int k = 0; k < 3; k++) { out_min[k] = s.mn[k]; out_max[k] = s.mx[k]; }
}

// setup_morton / setup_morton2 (builder.cpp:2204-2237, 2247-2279); 21 bits per
// axis is the 63-bit extension with scale 0x1FFFFF.
void sqo_morton_codes(const sqo_tri* tris, uint32_t n, uint32_t bits, uint64_t* codes) {
    float smn[3], smx[3];
    sqo_scene_bounds(tris, n, smn, smx);
    const float scale = (float)((1u << bits) - 1u);
    float ext[3];
    for (int k = 0; k < 3; k++) ext[k] = smx[k] - smn[k];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Box b = tri_box(tris[i]);
        uint64_t q[3];
        for (int k = 0; k < 3; k++) {
            float c = (b.mn[k] + b.mx[k]) * 0.5f;
            c = (c - smn[k]) / ext[k];
            q[k] = quantise(c * scale);
        }
        if (bits <= 10) codes[i] = sqo_interleave32((uint32_t)q[0], (uint32_t)q[1], (uint32_t)q[2]);
        else codes[i] = sqo_interleave64(q[0], q[1], q[2]);
    }
}

// LSD radix sort of (key, value) pairs, stable per pass (algorithms.h:11-113).
void sqo_radix_sort32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t b) {
    std::vector<uint32_t> k2(n), v2(n);
    uint32_t* ks = keys; uint32_t* vs = vals; uint32_t* kd = k2.data(); uint32_t* vd = v2.data();
    const uint32_t radix = 1u << b, mask = radix - 1;
    std::vector<uint32_t> cnt(radix);
    uint32_t passes = (32 + b - 1) / b;
    for (uint32_t p = 0; p < passes; p++) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (uint32_t i = 0; i < n; i++) cnt[(ks[i] >> (p * b)) & mask]++;
        uint32_t s = 0;
        for (uint32_t d = 0; d < radix; d++) { uint32_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint32_t i = 0; i < n; i++) {
            uint32_t pos = cnt[(ks[i] >> (p * b)) & mask]++;
            kd[pos] = ks[i]; vd[pos] = vs[i];
        }
        std::swap(ks, kd); std::swap(vs, vd);
    }
    if (ks != keys) { std::memcpy(keys, ks, n * 4); std::memcpy(vals, vs, n * 4); }
}

void sqo_radix_sort64(uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t b, uint32_t key_bits) {
    std::vector<uint64_t> k2(n); std::vector<uint32_t> v2(n);
    uint64_t* ks = keys; uint32_t* vs = vals; uint64_t* kd = k2.data(); uint32_t* vd = v2.data();
    const uint32_t radix = 1u << b; const uint64_t mask = radix - 1;
    std::vector<uint32_t> cnt(radix);
    uint32_t passes = (key_bits + b - 1) / b;
    for (uint32_t p = 0; p < passes; p++) {
        std::fill(cnt.begin(), cnt.end(), 0);
        for (uint32_t i = 0; i < n; i++) cnt[(ks[i] >> (p * b)) & mask]++;
        uint32_t s = 0;
        for (uint32_t d = 0; d < radix; d++) { uint32_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint32_t i = 0; i < n; i++) {
            uint32_t pos = cnt[(ks[i] >> (p * b)) & mask]++;
            kd[pos] = ks[i]; vd[pos] = vs[i];
        }
        std::swap(ks, kd); std::swap(vs, vd);
    }
    if (ks != keys) { std::memcpy(keys, ks, (size_t)n * 8); std::memcpy(vals, vs, (size_t)n * 4); }
}

static inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

// find_split_morton{,2} (builder.cpp:2281-2349): binary search for the last
// index sharing first's prefix beyond the common prefix; all-equal -> n>>1.
// lzcnt over 64-bit keys is used for every width: the split is invariant to
// the constant offset a narrower word would add to both sides.
uint32_t sqo_find_split(const uint64_t* codes, uint32_t count) {
    uint64_t first = codes[0], last_code = codes[count - 1];
    if (first == last_code) return count >> 1;
    int common = clz64(first ^ last_code);
    int split = 0, last = (int)count - 1, step = last;
    do {
        step = (step + 1) >> 1;
        int ns = split + step;
        if (ns < last) {
            if (clz64(first ^ codes[ns]) > common) split = ns;
        }
    } while (step > 1);
    return (uint32_t)split;
}

static uint32_t emit_rec(const uint64_t* codes, const uint32_t* idx, uint32_t begin, uint32_t end,
                         sqo_node* nodes, uint32_t free_index, uint32_t layout) {
    // recurse_morton (builder.cpp:2351-2409)
    uint32_t count = end - begin;
    if (count < 8) {
        for (uint32_t i = 0; i < count; i++) {
            sqo_node& lf = nodes[free_index + i];
            lf.child = idx[begin + i]; lf.count = 1; lf.flags = 1;
        }
        return count;
    }
    uint32_t split = sqo_find_split(codes + begin, count);
    uint32_t L = split + 1, R = count - L;
    uint32_t new_nodes = (L != 1) + (R != 1);
    uint32_t left_free = free_index + new_nodes;
    uint32_t right_free = free_index + new_nodes + (L - 1) * 2 + (L == 1);
    if (layout == SQO_LAYOUT_FIXED && R == 1 && L != 1) {
        // Right singleton gets the second pair slot; left children follow it.
        left_free = free_index + 2;
        right_free = free_index + 1;
    }
    uint32_t lc = emit_rec(codes, idx, begin, begin + L, nodes, left_free, layout);
    uint32_t rc = emit_rec(codes, idx, begin + L, end, nodes, right_free, layout);
    if (L != 1) { sqo_node& a = nodes[free_index]; a.child = left_free; a.count = (uint16_t)lc; a.flags = 0; }
    if (R != 1) { sqo_node& b = nodes[free_index + (L != 1)]; b.child = right_free; b.count = (uint16_t)rc; b.flags = 0; }
    return 2;
}

// build_hierarchy_morton (builder.cpp:2474-2508): root at 0 with child 1.
uint32_t sqo_emit_morton(const uint64_t* codes, const uint32_t* idx, uint32_t n, uint32_t layout,
                         sqo_node* nodes) {
    std::memset(nodes, 0, sizeof(sqo_node) * (size_t)2 * n);
    uint32_t c = emit_rec(codes, idx, 0, n, nodes, 1, layout);
    nodes[0].child = 1; nodes[0].count = (uint16_t)c; nodes[0].flags = 0;
    return 0;
}

static Box refit_rec(sqo_node* nodes, uint32_t i, const sqo_tri* tris) {
    sqo_node& nd = nodes[i];
    Box b;
    if (nd.flags & 1) {
        b = tri_box(tris[nd.child]);
        for (uint32_t j = 1; j < nd.count; j++) b.grow(tri_box(tris[nd.child + j]));
    } else {
        b.reset();
        for (uint32_t j = 0; j < nd.count; j++) {
            Box c = refit_rec(nodes, nd.child + j, tris);
            if (j == 0) b = c; else b.grow(c);
        }
    }
    for (int k = 0; k < 3; k++) { nd.min[k] = b.mn[k]; nd.max[k] = b.mx[k]; }
    return b;
}

// New (no reference counterpart): leaf box = triangle bounds as recurse()
// writes them (builder.cpp:1080-1085); interior = child-list union in order.
void sqo_refit(sqo_node* nodes, uint32_t node_count, uint32_t root, const sqo_tri* tris) {
    for (uint32_t i = 0; i < node_count; i++) {
        sqo_node& nd = nodes[i];
        if ((nd.flags & 1) && nd.count == 1) {
            Box b = tri_box(tris[nd.child]);
            for (int k = 0; k < 3; k++) { nd.min[k] = b.mn[k]; nd.max[k] = b.mx[k]; }
        }
    }
    refit_rec(nodes, root, tris);
}

uint32_t sqo_build_morton(const sqo_tri* tris, uint32_t n, uint32_t bits, uint32_t layout, uint32_t refit,
                          sqo_node* nodes, uint64_t* out_codes, uint32_t* out_idx) {
    std::vector<uint64_t> codes(n); std::vector<uint32_t> idx(n);
    sqo_morton_codes(tris, n, bits, codes.data());
    for (uint32_t i = 0; i < n; i++) idx[i] = i;
    sqo_radix_sort64(codes.data(), idx.data(), n, 8, 3 * bits);
    uint32_t root = sqo_emit_morton(codes.data(), idx.data(), n, layout, nodes);
    if (refit) sqo_refit(nodes, 2 * n, root, tris);
    if (out_codes) std::memcpy(out_codes, codes.data(), (size_t)n * 8);
    if (out_idx) std::memcpy(out_idx, idx.data(), (size_t)n * 4);
    return root;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Binned SAH
// ---------------------------------------------------------------------------
namespace {

struct Bin {
    Box p; int count; Box c;
    void reset() { p.reset(); c.reset(); count = 0; }
    Bin combine(const Bin& o) const { Bin r = *this; r.p.grow(o.p); r.c.grow(o.c); r.count += o.count; return r; }
};

struct SahCtx {
    std::vector<Box> aabbs;
    std::vector<float> centres;  // stride 4
    uint32_t nb, part, sa_kind, layout;
    sqo_node* nodes;
    std::vector<Bin> bins;
    std::vector<uint32_t> scratch_l, scratch_r;
};

// setup / setup_reduction (builder.cpp:54-147).
void sah_setup(SahCtx& c, const sqo_tri* tris, uint32_t n, Box& p_aabb, Box& c_aabb) {
    c.aabbs.resize(n); c.centres.assign((size_t)n * 4, 0.0f);
    p_aabb.reset(); c_aabb.reset();
    for (uint32_t i = 0; i < n; i++) {
        Box a = tri_box(tris[i]);
        c.aabbs[i] = a;
        float ct[3];
        for (int k = 0; k < 3; k++) ct[k] = (a.mn[k] + a.mx[k]) * 0.5f;
        p_aabb.grow(a); c_aabb.grow_pt(ct);
        for (int k = 0; k < 3; k++) c.centres[(size_t)i * 4 + k] = ct[k];
    }
}

// choose_axis (builder.cpp:334-345)
int choose_axis(const Box& b) {
    float lx = b.mx[0] - b.mn[0], ly = b.mx[1] - b.mn[1], lz = b.mx[2] - b.mn[2];
    int r = 0;
    r += 2 * (lz > lx && lz > ly);
    r += 1 * (ly > lx && ly >= lz);
    return r;
}

inline float bin_k1(uint32_t nb, const Box& cb, int axis) {
    const float epsilon = 1.1920929e-7f;
    return (float)nb * (1 - epsilon) / (cb.mx[axis] - cb.mn[axis]);
}
inline int bin_of(float k1, float c, float cmin) { return (int)(k1 * (c - cmin)); }

// bin_centroids (builder.cpp:372-398). An id of nb (possible after rounding,
// unchecked in the reference) is clamped to nb-1.
void bin_centroids(SahCtx& c, const uint32_t* ids, uint32_t count, const Box& cb, int axis) {
    float k1 = bin_k1(c.nb, cb, axis);
    for (auto& b : c.bins) b.reset();
    for (uint32_t i = 0; i < count; i++) {
        uint32_t id = ids[i];
        int bi = bin_of(k1, c.centres[(size_t)id * 4 + axis], cb.mn[axis]);
        if (bi < 0) bi = 0;
        if (bi >= (int)c.nb) bi = (int)c.nb - 1;
        c.bins[bi].p.grow(c.aabbs[id]);
        c.bins[bi].c.grow_pt(&c.centres[(size_t)id * 4]);
        c.bins[bi].count += 1;
    }
}

// choose_plane (builder.cpp:488-523): score = fma(sa_l, n_l, sa_r * n_r).
int choose_plane(SahCtx& c, Box& lp, Box& rp, Box& lc, Box& rc) {
    const uint32_t nb = c.nb;
    std::vector<Bin> l2r(nb - 1);
    l2r[0] = c.bins[0];
    for (uint32_t i = 1; i < nb - 1; i++) l2r[i] = l2r[i - 1].combine(c.bins[i]);
    Bin r2l = c.bins[nb - 1];
    int result = 0; float best = FLT_MAX; bool set = false;
    for (int i = (int)nb - 2; i >= 0; i--) {
        float score = fmaf(box_sa(l2r[i].p, c.sa_kind), (float)l2r[i].count,
                           box_sa(r2l.p, c.sa_kind) * (float)r2l.count);
        if (score < best) {
            best = score; result = i; set = true;
            lp = l2r[i].p; rp = r2l.p; lc = l2r[i].c; rc = r2l.c;
        }
        r2l = r2l.combine(c.bins[i]);
    }
    if (!set) {  // reference leaves the boxes uninitialised here
        Bin rest = c.bins[nb - 1];
        for (uint32_t i = nb - 2; i >= 1; i--) rest = rest.combine(c.bins[i]);
        lp = l2r[0].p; lc = l2r[0].c; rp = rest.p; rc = rest.c;
    }
    return result;
}

// partition (builder.cpp:599-628): in-place Hoare swap.
uint32_t partition_hoare(SahCtx& c, int plane, int axis, const Box& cb, uint32_t* ids, uint32_t count) {
    float k1 = bin_k1(c.nb, cb, axis);
    int64_t a = 0, b = (int64_t)count - 1;
    auto side = [&](uint32_t id) { return bin_of(k1, c.centres[(size_t)id * 4 + axis], cb.mn[axis]); };
    while (a < b) {
        if (side(ids[a]) <= plane && a <= b) a++;
        else if (side(ids[b]) > plane && a <= b) b--;
        else if (a < b) std::swap(ids[a], ids[b]);
    }
    return (uint32_t)a;
}

// partition_sse / partition_m128 (builder.cpp:683-903): stable two-way split.
uint32_t partition_stable(SahCtx& c, int plane, int axis, const Box& cb, uint32_t* ids, uint32_t count) {
    float k1 = bin_k1(c.nb, cb, axis);
    c.scratch_l.clear(); c.scratch_r.clear();
    for (uint32_t i = 0; i < count; i++) {
        uint32_t id = ids[i];
        if (bin_of(k1, c.centres[(size_t)id * 4 + axis], cb.mn[axis]) <= plane) c.scratch_l.push_back(id);
        else c.scratch_r.push_back(id);
    }
    std::copy(c.scratch_l.begin(), c.scratch_l.end(), ids);
    std::copy(c.scratch_r.begin(), c.scratch_r.end(), ids + c.scratch_l.size());
    return (uint32_t)c.scratch_l.size();
}

// recurse / recurse_omp / recurse_sse / recurse_m128 (builder.cpp:1053-1156,
// 1158-1264, 1266-1363, 1365-1477). Returns (first child slot, child count).
std::pair<uint32_t, uint32_t> sah_rec(SahCtx& c, uint32_t* ids, uint32_t count, const Box& cb,
                                      uint32_t free_index) {
    if (count <= 8 || box_sa(cb, c.sa_kind) <= 0.0f) {
        for (uint32_t i = 0; i < count; i++) write_leaf(c.nodes[free_index + i], c.aabbs[ids[i]], ids[i]);
        return {free_index, count};
    }
    int axis = choose_axis(cb);
    bin_centroids(c, ids, count, cb, axis);
    Box lp, rp, lc, rc;
    int plane = choose_plane(c, lp, rp, lc, rc);
    uint32_t mid = (c.part == SQO_PART_HOARE) ? partition_hoare(c, plane, axis, cb, ids, count)
                                              : partition_stable(c, plane, axis, cb, ids, count);
    uint32_t L = mid, R = count - mid;
    uint32_t pair_index = free_index;
    uint32_t left_index = free_index;
    free_index += (L != 1);
    uint32_t right_index = free_index;
    free_index += (R != 1);
    uint32_t left_free = free_index;
    uint32_t right_free = free_index + (L - 1) * 2 + (L == 1);
    if (c.layout == SQO_LAYOUT_FIXED && R == 1 && L != 1) {
        right_index = pair_index + 1;
        left_free = pair_index + 2;
        right_free = pair_index + 1;
    }
    auto lchild = sah_rec(c, ids, L, lc, left_free);
    auto rchild = sah_rec(c, ids + L, R, rc, right_free);
    if (L != 1) write_inner(c.nodes[left_index], lp, lchild.first, lchild.second);
    if (R != 1) write_inner(c.nodes[right_index], rp, rchild.first, rchild.second);
    return {pair_index, 2};
}

}  // namespace

extern "C" {

uint32_t sqo_sah_node_capacity(uint32_t n) { return 2 * n; }
uint32_t sqo_grid_node_capacity(uint32_t n, uint32_t dim) { return 2 * n + 2 * dim * dim * dim; }

// Builder::build_hierarchy / build_hierarchy_horizontal (builder.cpp:1576-1663):
// root at 2n-1 holding the scene box. The horizontal builder produces the same
// tree (its root-level reduction changes evaluation order only).
uint32_t sqo_build_sah(const sqo_tri* tris, uint32_t n, uint32_t nb, uint32_t part, uint32_t sa_kind,
                       uint32_t layout, sqo_node* nodes) {
    SahCtx c;
    c.nb = nb; c.part = part; c.sa_kind = sa_kind; c.layout = layout; c.nodes = nodes;
    c.bins.resize(nb);
    std::memset(nodes, 0, sizeof(sqo_node) * (size_t)2 * n);
    Box p_aabb, c_aabb;
    sah_setup(c, tris, n, p_aabb, c_aabb);
    std::vector<uint32_t> ids(n);
    for (uint32_t i = 0; i < n; i++) ids[i] = i;
    auto root = sah_rec(c, ids.data(), n, c_aabb, 0);
    uint32_t ri = 2 * n - 1;
    write_inner(nodes[ri], p_aabb, root.first, root.second);
    return ri;
}

static inline uint32_t compact_1_by_2(uint32_t x) {
    // builder.cpp:27-35
    x &= 0x09249249;
    x = (x ^ (x >> 2)) & 0x030c30c3;
    x = (x ^ (x >> 4)) & 0x0300f00f;
    x = (x ^ (x >> 8)) & 0xff0000ff;
    x = (x ^ (x >> 16)) & 0x000003ff;
    return x;
}

// build_hierarchy_grid{,_sse,_m128} (builder.cpp:1993-2194) with
// bin_centroids_to_grid{,_sse} (1665-1839): stable per-cell scatter in
// Morton cell order, per-cell SAH subtree at 2*cells + 2*begin, cell nodes at
// cells-1+i, heap top 0..cells-2, root 0.
uint32_t sqo_build_grid(const sqo_tri* tris, uint32_t n, uint32_t dim, uint32_t nb, uint32_t part,
                        uint32_t sa_kind, uint32_t cell_box, uint32_t layout, sqo_node* nodes) {
    const uint32_t cells = dim * dim * dim;
    SahCtx c;
    c.nb = nb; c.part = part; c.sa_kind = sa_kind; c.layout = layout; c.nodes = nodes;
    c.bins.resize(nb);
    std::memset(nodes, 0, sizeof(sqo_node) * (size_t)sqo_grid_node_capacity(n, dim));
    Box p_aabb, c_aabb;
    sah_setup(c, tris, n, p_aabb, c_aabb);
    const float epsilon = 1.1920929e-7f;
    float k1[3];
    for (int a = 0; a < 3; a++) k1[a] = (float)dim * (1 - epsilon) / (c_aabb.mx[a] - c_aabb.mn[a]);
    std::vector<uint32_t> cell_of(n);
    std::vector<Bin> gb(cells);
    for (auto& b : gb) b.reset();
    auto lin = [&](uint32_t kx, uint32_t ky, uint32_t kz) { return (kz * dim + ky) * dim + kx; };
    for (uint32_t i = 0; i < n; i++) {
        int bid[3];
        for (int a = 0; a < 3; a++) {
            bid[a] = bin_of(k1[a], c.centres[(size_t)i * 4 + a], c_aabb.mn[a]);
            if (bid[a] < 0) bid[a] = 0;
            if (bid[a] >= (int)dim) bid[a] = (int)dim - 1;
        }
        uint32_t li = lin(bid[0], bid[1], bid[2]);
        cell_of[i] = li;
        gb[li].p.grow(c.aabbs[i]); gb[li].c.grow_pt(&c.centres[(size_t)i * 4]); gb[li].count += 1;
    }
    std::vector<uint32_t> begin(cells + 1), order_of_lin(cells * 1);
    uint32_t prefix = 0;
    for (uint32_t i = 0; i < cells; i++) {
        uint32_t kx = compact_1_by_2(i >> 0), ky = compact_1_by_2(i >> 1), kz = compact_1_by_2(i >> 2);
        uint32_t li = lin(kx, ky, kz);
        order_of_lin[li] = i;
        begin[i] = prefix;
        prefix += gb[li].count;
    }
    begin[cells] = n;
    std::vector<uint32_t> ids(n), fill(cells, 0);
    for (uint32_t i = 0; i < n; i++) {
        uint32_t ci = order_of_lin[cell_of[i]];
        ids[begin[ci] + fill[ci]++] = i;
    }
    const uint32_t offset = cells * 2;
    for (uint32_t i = 0; i < cells; i++) {
        uint32_t kx = compact_1_by_2(i >> 0), ky = compact_1_by_2(i >> 1), kz = compact_1_by_2(i >> 2);
        const Bin& cb = gb[lin(kx, ky, kz)];
        auto child = sah_rec(c, ids.data() + begin[i], begin[i + 1] - begin[i], cb.c, offset + begin[i] * 2);
        write_inner(nodes[i + cells - 1], cell_box == SQO_CELLBOX_TIGHT ? cb.p : p_aabb, child.first,
                    child.second);
    }
    for (int i = (int)cells - 2; i >= 0; i--) {
        sqo_node& nd = nodes[i];
        const sqo_node &a = nodes[i * 2 + 1], &b = nodes[i * 2 + 2];
        for (int k = 0; k < 3; k++) { nd.min[k] = std_min(a.min[k], b.min[k]); nd.max[k] = std_max(a.max[k], b.max[k]); }
        nd.child = i * 2 + 1; nd.count = 2; nd.flags = 0;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Rays. trace()/trace_packet()/trace_bundle() (tracer.cpp:202-216,247-262,
// 316-334): ndc = fma((c+0.5)/dim, 2, -1); x scaled by w/h; p = (x, y, 1);
// |p|^2 as fma(pz,pz, fma(px,px, py*py)); d = p/|p|. The tan(half fov)
// factor generalises the fixed 90-degree camera (1.0 leaves it bit-exact).
// ---------------------------------------------------------------------------
static inline uint32_t hash_u32(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

void sqo_camera_ray(const sqo_camera* cam, uint32_t px, uint32_t py, uint32_t s, sqo_ray* r) {
    float jx = 0.5f, jy = 0.5f;
    if (cam->jitter) {
        uint32_t pix = py * cam->width + px;
        uint32_t h1 = hash_u32(cam->seed ^ hash_u32(pix * cam->spp + s));
        uint32_t h2 = hash_u32(h1 ^ 0x9E3779B9u);
        jx = (float)(h1 >> 8) * (1.0f / 16777216.0f);
        jy = (float)(h2 >> 8) * (1.0f / 16777216.0f);
    }
    float w = (float)cam->width, h = (float)cam->height;
    float nx = fmaf(((float)px + jx) / w, 2.0f, -1.0f);
    float ny = fmaf(((float)py + jy) / h, 2.0f, -1.0f);
    nx = nx * (w / h);
    float tx = nx * cam->tan_half_fov, ty = ny * cam->tan_half_fov, tz = 1.0f;
    float len = sqrtf(fmaf(tz, tz, fmaf(tx, tx, ty * ty)));
    r->ox = cam->origin[0]; r->oy = cam->origin[1]; r->oz = cam->origin[2];
    r->dx = tx / len; r->dy = ty / len; r->dz = tz / len;
    r->tmin = cam->tmin; r->tmax = cam->tmax;
}

void sqo_camera_rays(const sqo_camera* cam, sqo_ray* out) {
    const uint32_t W = cam->width, H = cam->height, S = cam->spp;
#pragma omp parallel for schedule(static)
    for (int64_t py = 0; py < (int64_t)H; py++)
        for (uint32_t px = 0; px < W; px++)
            for (uint32_t s = 0; s < S; s++) sqo_camera_ray(cam, px, (uint32_t)py, s, &out[((size_t)py * W + px) * S + s]);
}

// intersect_ray_aabb<1> (tracer.h:13-35)
int sqo_slab(const sqo_node* nd, const sqo_ray* r) {
    float ix = 1.0f / r->dx, iy = 1.0f / r->dy, iz = 1.0f / r->dz;
    float t1x = (nd->min[0] - r->ox) * ix, t2x = (nd->max[0] - r->ox) * ix;
    float t1y = (nd->min[1] - r->oy) * iy, t2y = (nd->max[1] - r->oy) * iy;
    float t1z = (nd->min[2] - r->oz) * iz, t2z = (nd->max[2] - r->oz) * iz;
    float tnx = simd_min(t1x, t2x), tny = simd_min(t1y, t2y), tnz = simd_min(t1z, t2z);
    float txx = simd_max(t1x, t2x), txy = simd_max(t1y, t2y), txz = simd_max(t1z, t2z);
    float front = simd_max(simd_max(tnx, tny), tnz);
    float back = simd_min(simd_min(txx, txy), txz);
    return (back >= front) & (back <= r->tmax);
}

// intersect_ray_triangle<1> (tracer.h:83-116) with g++'s contraction:
// cross a*b - c*d -> fma(-c, d, a*b); dot -> fma(z, z', fma(y, y', x*x')).
int sqo_tri_test(const sqo_tri* tri, const sqo_ray* r, float* t_out) {
    const float eps = 0.000000001f;
    float e1x = tri->v1.x - tri->v0.x, e1y = tri->v1.y - tri->v0.y, e1z = tri->v1.z - tri->v0.z;
    float e2x = tri->v2.x - tri->v0.x, e2y = tri->v2.y - tri->v0.y, e2z = tri->v2.z - tri->v0.z;
    float hx = fmaf(-e2y, r->dz, r->dy * e2z);
    float hy = fmaf(-e2z, r->dx, r->dz * e2x);
    float hz = fmaf(-e2x, r->dy, r->dx * e2y);
    float a = fmaf(e1z, hz, fmaf(e1y, hy, e1x * hx));
    bool miss = (a > -eps) & (a < eps);
    float f = 1.0f / a;
    float sx = r->ox - tri->v0.x, sy = r->oy - tri->v0.y, sz = r->oz - tri->v0.z;
    float u = f * fmaf(sz, hz, fmaf(sy, hy, sx * hx));
    miss |= (u < 0.0f) | (u > 1.0f);
    float qx = fmaf(-e1y, sz, sy * e1z);
    float qy = fmaf(-e1z, sx, sz * e1x);
    float qz = fmaf(-e1x, sy, sx * e1y);
    float v = f * fmaf(r->dz, qz, fmaf(r->dy, qy, r->dx * qx));
    miss |= (v < 0.0f) | ((u + v) > 1.0f);
    float t = f * fmaf(e2z, qz, fmaf(e2y, qy, e2x * qx));
    *t_out = t;
    return (!miss) & (t > eps);
}

}  // extern "C"

namespace {

struct HitRec { uint32_t count = 0; uint32_t id = 0xFFFFFFFFu; float t = INFINITY; uint32_t any = 0; };

inline void test_leaf(const sqo_node& nd, const sqo_tri* tris, const sqo_ray& r, HitRec& h) {
    for (uint32_t j = 0; j < nd.count; j++) {
        uint32_t tid = nd.child + j;
        float t;
        int m = sqo_tri_test(&tris[tid], &r, &t);
        h.any |= (uint32_t)m;
        if (m && t > r.tmin && t < r.tmax) {
            if (t < h.t || (t == h.t && tid < h.id)) { h.t = t; h.id = tid; }
        }
    }
}

// traverse (tracer.cpp:47-65): exhaustive; count = node tests.
void traverse(const sqo_node* nodes, uint32_t index, uint32_t count, const sqo_tri* tris, const sqo_ray& r,
              HitRec& h) {
    for (uint32_t i = 0; i < count; i++) {
        const sqo_node& nd = nodes[index + i];
        h.count++;
        if (sqo_slab(&nd, &r)) {
            if (nd.flags & 1) test_leaf(nd, tris, r, h);
            else traverse(nodes, nd.child, nd.count, tris, r, h);
        }
    }
}

}  // namespace

extern "C" {

void sqo_trace_rays(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_ray* rays,
                    uint32_t nrays, uint32_t* counts, uint32_t* hit_id, float* hit_t, uint32_t* any_hit) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        HitRec h;
        traverse(nodes, root, 1, tris, rays[i], h);
        if (counts) counts[i] = h.count;
        if (hit_id) hit_id[i] = h.id;
        if (hit_t) hit_t[i] = h.t;
        if (any_hit) any_hit[i] = h.any;
    }
}

// Closest-hit traversal: the box test additionally culls nodes whose entry
// distance exceeds the best hit by more than 2^-20 relative, so the result
// equals the exhaustive closest hit.
void sqo_trace_rays_closest(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_ray* rays,
                            uint32_t nrays, uint32_t* hit_id, float* hit_t) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        const sqo_ray& r = rays[i];
        HitRec h;
        std::vector<uint32_t> st;
        st.push_back(root);
        while (!st.empty()) {
            uint32_t ni = st.back(); st.pop_back();
            const sqo_node& nd = nodes[ni];
            if (!sqo_slab(&nd, &r)) continue;
            // entry distance for culling
            float ix = 1.0f / r.dx, iy = 1.0f / r.dy, iz = 1.0f / r.dz;
            float t1x = (nd.min[0] - r.ox) * ix, t2x = (nd.max[0] - r.ox) * ix;
            float t1y = (nd.min[1] - r.oy) * iy, t2y = (nd.max[1] - r.oy) * iy;
            float t1z = (nd.min[2] - r.oz) * iz, t2z = (nd.max[2] - r.oz) * iz;
            float front = simd_max(simd_max(simd_min(t1x, t2x), simd_min(t1y, t2y)), simd_min(t1z, t2z));
            if (front > h.t + h.t * (1.0f / 1048576.0f)) continue;
            if (nd.flags & 1) test_leaf(nd, tris, r, h);
            else for (uint32_t j = 0; j < nd.count; j++) st.push_back(nd.child + nd.count - 1 - j);
        }
        hit_id[i] = h.id; hit_t[i] = h.t;
    }
}

// tracer.cpp:222-225: uint8(min(1, c/200) * 255); samples averaged first.
uint8_t sqo_heat_byte(uint32_t count, uint32_t spp) {
    float c = (float)count;
    if (spp > 1) c = c / (float)spp;
    float v = c / 200.0f;
    v = (v < 1.0f) ? v : 1.0f;   // std::min(1.0f, v)
    return (uint8_t)(v * 255.0f);
}

void sqo_trace_camera(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_camera* cam,
                      uint32_t* pixel_counts, uint8_t* rgb, uint32_t* hit_id, float* hit_t) {
    const uint32_t W = cam->width, H = cam->height, S = cam->spp;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t py = 0; py < (int64_t)H; py++) {
        for (uint32_t px = 0; px < W; px++) {
            uint32_t sum = 0;
            for (uint32_t s = 0; s < S; s++) {
                sqo_ray r;
                sqo_camera_ray(cam, px, (uint32_t)py, s, &r);
                HitRec h;
                traverse(nodes, root, 1, tris, r, h);
                sum += h.count;
                size_t ri = ((size_t)py * W + px) * S + s;
                if (hit_id) hit_id[ri] = h.id;
                if (hit_t) hit_t[ri] = h.t;
            }
            size_t p = (size_t)py * W + px;
            if (pixel_counts) pixel_counts[p] = sum;
            if (rgb) { uint8_t b = sqo_heat_byte(sum, S); rgb[p * 3] = b; rgb[p * 3 + 1] = b; rgb[p * 3 + 2] = b; }
        }
    }
}

}  // extern "C"

namespace {

// traverse_packet_stack<K> (tracer.cpp:88-125): descend when any lane hits;
// one count per node test of the packet.
uint32_t packet_traverse(const sqo_node* nodes, uint32_t index, uint32_t count, const sqo_ray* rays, uint32_t K) {
    struct E { uint32_t index, count; };
    std::vector<E> st;
    st.push_back({index, count});
    uint32_t tests = 0;
    while (!st.empty()) {
        E e = st.back(); st.pop_back();
        for (uint32_t i = 0; i < e.count; i++) {
            const sqo_node& nd = nodes[e.index + i];
            int hit = 0;
            for (uint32_t k = 0; k < K; k++) hit |= sqo_slab(&nd, &rays[k]);
            tests++;
            if (hit && !(nd.flags & 1)) st.push_back({nd.child, nd.count});
        }
    }
    return tests;
}

const uint32_t kMortonLut[64] = {  // tracer.cpp:18-23: pixel order inside a bundle tile
    0x00, 0x01, 0x04, 0x05, 0x02, 0x03, 0x06, 0x07, 0x08, 0x09, 0x0C, 0x0D, 0x0A, 0x0B, 0x0E, 0x0F,
    0x10, 0x11, 0x14, 0x15, 0x12, 0x13, 0x16, 0x17, 0x18, 0x19, 0x1C, 0x1D, 0x1A, 0x1B, 0x1E, 0x1F,
    0x20, 0x21, 0x24, 0x25, 0x22, 0x23, 0x26, 0x27, 0x28, 0x29, 0x2C, 0x2D, 0x2A, 0x2B, 0x2E, 0x2F,
    0x30, 0x31, 0x34, 0x35, 0x32, 0x33, 0x36, 0x37, 0x38, 0x39, 0x3C, 0x3D, 0x3A, 0x3B, 0x3E, 0x3F};

struct Bundle { float dmin[3], dmax[3], pos[3]; };

// intersect_bundle_aabb (tracer.h:39-80)
void bundle_test(const Bundle& b, const sqo_node& nd, bool& any, bool& mid) {
    float t1a[3], t1b[3], t2a[3], t2b[3], t1[3], t2[3];
    for (int k = 0; k < 3; k++) {
        float imn = 1.0f / b.dmin[k], imx = 1.0f / b.dmax[k];
        float imd = 1.0f / ((b.dmin[k] + b.dmax[k]) * 0.5f);
        float lo = nd.min[k] - b.pos[k], hi = nd.max[k] - b.pos[k];
        t1a[k] = lo * imn; t1b[k] = lo * imx; t2a[k] = hi * imn; t2b[k] = hi * imx;
        t1[k] = lo * imd; t2[k] = hi * imd;
    }
    float tmin_mid[3], tmin_any[3], tmax_mid[3], tmax_any[3];
    for (int k = 0; k < 3; k++) {
        float t1min = simd_min(t1a[k], t1b[k]), t1max = simd_max(t1a[k], t1b[k]);
        float t2min = simd_min(t2a[k], t2b[k]), t2max = simd_max(t2a[k], t2b[k]);
        tmin_mid[k] = simd_min(t1[k], t2[k]);
        tmin_any[k] = simd_min(t1min, t2min);
        tmax_mid[k] = simd_max(t1[k], t2[k]);
        tmax_any[k] = simd_max(t1max, t2max);
    }
    float front_mid = std_max(std_max(tmin_mid[0], tmin_mid[1]), tmin_mid[2]);
    float front_any = std_max(std_max(tmin_any[0], tmin_any[1]), tmin_any[2]);
    float back_mid = std_min(std_min(tmax_mid[0], tmax_mid[1]), tmax_mid[2]);
    float back_any = std_min(std_min(tmax_any[0], tmax_any[1]), tmax_any[2]);
    mid = back_mid >= front_mid;
    any = back_any >= front_any;
}

// traverse_bundle (tracer.cpp:128-156)
void bundle_traverse(const sqo_node* nodes, uint32_t index, uint32_t count, const Bundle& b,
                     const sqo_ray* rays, uint32_t M, uint32_t& bundle_tests, uint32_t* packet_tests) {
    for (uint32_t i = 0; i < count; i++) {
        const sqo_node& nd = nodes[index + i];
        bool any, mid;
        bundle_test(b, nd, any, mid);
        bundle_tests++;
        if (mid) {
            if (!(nd.flags & 1)) bundle_traverse(nodes, nd.child, nd.count, b, rays, M, bundle_tests, packet_tests);
        } else if (any) {
            for (uint32_t k = 0; k < M; k++) packet_tests[k] += packet_traverse(nodes, index + i, 1, rays + 8 * k, 8);
        }
    }
}

}  // namespace

extern "C" {

// trace_packet<K> (tracer.cpp:232-281)
void sqo_trace_packet(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_camera* cam,
                      uint32_t K, uint32_t* packet_counts, uint8_t* rgb) {
    (void)tris;
    const uint32_t W = cam->width, H = cam->height;
    uint32_t kx = (uint32_t)std::floor(std::sqrt((double)K)), ky = K / kx;
    uint32_t pw = (W + kx - 1) / kx;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t jj = 0; jj < (int64_t)((H + ky - 1) / ky); jj++) {
        uint32_t j = (uint32_t)jj * ky;
        for (uint32_t i = 0; i < W; i += kx) {
            sqo_ray rays[8];
            uint32_t m = 0;
            for (uint32_t my = 0; my < ky; my++)
                for (uint32_t mx = 0; mx < kx; mx++) sqo_camera_ray(cam, i + mx, j + my, 0, &rays[m++]);
            uint32_t tests = packet_traverse(nodes, root, 1, rays, K);
            if (packet_counts) packet_counts[(size_t)jj * pw + i / kx] = tests;
            float v = ((float)tests / (float)K) / 200.0f;
            v = (v < 1.0f) ? v : 1.0f;
            uint8_t byte = (uint8_t)(v * 255.0f);
            for (uint32_t my = 0; my < ky; my++)
                for (uint32_t mx = 0; mx < kx; mx++) {
                    if (j + my >= H || i + mx >= W) continue;
                    size_t p = ((size_t)(j + my) * W + i + mx) * 3;
                    rgb[p] = byte; rgb[p + 1] = byte; rgb[p + 2] = byte;
                }
        }
    }
}

// trace_bundle<K> (tracer.cpp:292-362)
void sqo_trace_bundle(const sqo_node* nodes, uint32_t root, const sqo_tri* tris, const sqo_camera* cam,
                      uint32_t K, uint8_t* rgb) {
    (void)tris;
    const uint32_t W = cam->width, H = cam->height;
    const uint32_t N = 8, tile_w = (uint32_t)std::sqrt((double)K), tile_h = K / tile_w, M = K / N;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t jj = 0; jj < (int64_t)((H + tile_h - 1) / tile_h); jj++) {
        uint32_t j = (uint32_t)jj * tile_h;
        for (uint32_t i = 0; i < W; i += tile_w) {
            sqo_ray rays[64];
            Bundle b;
            for (int k = 0; k < 3; k++) { b.dmin[k] = FLT_MAX; b.dmax[k] = -FLT_MAX; b.pos[k] = cam->origin[k]; }
            for (uint32_t k = 0; k < K; k++) {
                uint32_t mx = kMortonLut[k] % tile_w, my = kMortonLut[k] / tile_w;
                sqo_camera_ray(cam, i + mx, j + my, 0, &rays[k]);
                const float d[3] = {rays[k].dx, rays[k].dy, rays[k].dz};
                for (int a = 0; a < 3; a++) { b.dmin[a] = simd_min(b.dmin[a], d[a]); b.dmax[a] = simd_max(b.dmax[a], d[a]); }
            }
            uint32_t btests = 0, ptests[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            bundle_traverse(nodes, root, 1, b, rays, M, btests, ptests);
            for (uint32_t k = 0; k < K; k++) {
                uint32_t mx = kMortonLut[k] % tile_w, my = kMortonLut[k] / tile_w, m = k / N;
                if (j + my >= H || i + mx >= W) continue;
                float tests = (float)btests / (float)K + (float)ptests[m] / (float)N;
                float v = tests / 200.0f;
                v = (v < 1.0f) ? v : 1.0f;
                uint8_t byte = (uint8_t)(v * 255.0f);
                size_t p = ((size_t)(j + my) * W + i + mx) * 3;
                rgb[p] = byte; rgb[p + 1] = byte; rgb[p + 2] = byte;
            }
        }
    }
}

}  // extern "C"
