// nodek.cpp — CPU restatement of the wide-node format (Node<K>, K = 2, 4, 8)
// and its traversal.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP conversion and traversal
// kernels (sequoia_amd/csrc/nodek.hip, the wide steps of trace.hip). Built
// with -ffp-contract=off.
//
// Layout: the reference's template <size_t K> struct Node (aabb.h:26-32):
// vec3f<K> min, simdi<K> flags, vec3f<K> max, simdi<K> child, i.e. 8 K
// 32-bit words per node: min x[K], min y[K], min z[K], flags[K], max x[K],
// max y[K], max z[K], child[K]. The reference never builds or reads a
// Node<K>, so its meaning is defined here (and mirrored by the GPU code):
//   * slot j of a node describes one child: a Node1 node of the source tree
//     (its box as stored there) or a run of several (the union of their
//     boxes), a word in Node1's count/flags order (count | flags << 16) and a
//     child index;
//   * a leaf slot copies a Node1 leaf: flags bit 0 set, count triangles from
//     child on (the same triangle array and ids as the Node1 tree);
//   * an interior slot names the wide node `child` that holds the Node1
//     node's children (or the run's nodes); its count is how many slots of
//     that node are used (1..K, used slots first);
//   * an interior Node1 node with an empty child list (a grid builder's
//     empty cell) keeps its slot, word 0 and child 0xFFFFFFFF: tested, and
//     leading nowhere;
//   * unused slots: min +inf, max -inf, word 0, child 0xFFFFFFFF (never
//     tested: a node's occupancy comes with the slot that names it).
// The root record holds the Node1 root's box (tested as stored) and either
// the root leaf (kind 1: child = first triangle, count), the root wide node
// (kind 2: child, count = its occupancy) or nothing (kind 0: an interior root
// with an empty list).
//
// Conversion. A wide node holds a list (f, c) of Node1 nodes: the child list
// of an interior node, or a run of one. Its slots start as the list's nodes
// (c <= K) or, for a longer list, as K - m single nodes followed by m runs
// over the rest (m = min(K, ceil((c - K) / (K - 1))), the fewest runs of at
// most K nodes that fit; the first runs one node longer where the rest does
// not divide evenly); a run is a slot whose box is the union of its nodes'
// boxes. While some slot can be replaced
// in place by what it stands for without exceeding K slots (an interior
// node: its child list; a longer run: its nodes), the one with the largest
// half surface area dx*dy + dy*dz + dz*dx (plain float operations in that
// order; the first slot on ties) is. `expand` = 0 skips this: each wide node
// is then one Node1 child list, and when no list is longer than K the
// traversal tests exactly the Node1 boxes the reference's traverse tests.
// Every interior slot then names a wide node for its list (an interior
// node's child list, or the run). Wide nodes are numbered breadth first:
// level by level, nodes of a level in the order of the slots that name them.
// A Node1 node reached through two paths (the reference layout's
// right-singleton lists) gets a wide node per path.
//
// Traversal: the root box costs one test; a visited wide node costs one test
// per used slot. Exhaustive mode visits every hit slot; closest mode visits
// nearer slots first and culls entries whose distance exceeds the best hit by
// more than 2^-20 relative. Hits: the closest valid triangle (t > tmin,
// t < tmax; lower id on equal t), with the Node1 tree's triangle ids.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline float kmin(float a, float b) { return (b < a) ? b : a; }
inline float kmax(float a, float b) { return (a < b) ? b : a; }

struct Slot {
    uint32_t f, c;  // the run of Node1 nodes it stands for
    float mn[3], mx[3];
};

// dx*dy + dy*dz + dz*dx (half the surface area).
inline float half_area(const Slot& s) {
    const float dx = s.mx[0] - s.mn[0], dy = s.mx[1] - s.mn[1], dz = s.mx[2] - s.mn[2];
    const float a = dx * dy, b = dy * dz, c = dz * dx;
    return (a + b) + c;
}

inline bool run_ok(uint32_t f, uint32_t c, uint32_t node_count) {
    return c != 0 && f < node_count && c <= node_count - f;
}

// The slot for run (f, c): its box is the first node's, widened by the rest.
inline Slot make_slot(const sqo_node* nodes, uint32_t f, uint32_t c) {
    Slot s;
    s.f = f;
    s.c = c;
    for (int k = 0; k < 3; k++) { s.mn[k] = nodes[f].min[k]; s.mx[k] = nodes[f].max[k]; }
    for (uint32_t i = 1; i < c; i++)
        for (int k = 0; k < 3; k++) {
            s.mn[k] = kmin(s.mn[k], nodes[f + i].min[k]);
            s.mx[k] = kmax(s.mx[k], nodes[f + i].max[k]);
        }
    return s;
}

inline bool slot_leaf(const sqo_node* nodes, const Slot& s) { return s.c == 1 && (nodes[s.f].flags & 1); }
// An interior Node1 node with an empty child list (a grid builder's empty
// cell): its slot is tested and leads nowhere.
inline bool slot_dead(const sqo_node* nodes, const Slot& s) {
    return s.c == 1 && !(nodes[s.f].flags & 1) && nodes[s.f].count == 0;
}

// The list a slot's wide node holds: an interior node's child list, or the run.
inline void slot_list(const sqo_node* nodes, const Slot& s, uint32_t* f, uint32_t* c) {
    if (s.c == 1) { *f = nodes[s.f].child; *c = nodes[s.f].count; }
    else { *f = s.f; *c = s.c; }
}

constexpr uint32_t kUnused = 0xFFFFFFFFu;
constexpr uint32_t kMaxLevels = 4096;

// Slots of the wide node for list (f, c). Returns 0 or -1 (malformed tree).
int collapse(const sqo_node* nodes, uint32_t node_count, uint32_t f, uint32_t c, uint32_t K, uint32_t expand,
             Slot* slots, uint32_t* used) {
    if (!run_ok(f, c, node_count)) return -1;
    uint32_t n = c;
    if (c <= K) {
        for (uint32_t j = 0; j < c; j++) slots[j] = make_slot(nodes, f + j, 1);
    } else {
        // s single nodes, then m runs over the rest (the first (c - s) % m
        // one longer): the fewest runs with which runs of at most K nodes
        // cover the list, at most K of them
        uint32_t m = (c - K + K - 2) / (K - 1);
        if (m > K) m = K;
        const uint32_t s = K - m, q = (c - s) / m, r = (c - s) % m;
        for (uint32_t j = 0; j < s; j++) slots[j] = make_slot(nodes, f + j, 1);
        for (uint32_t j = 0, at = f + s; j < m; j++) {
            const uint32_t len = q + (j < r ? 1u : 0u);
            slots[s + j] = make_slot(nodes, at, len);
            at += len;
        }
        n = K;
    }
    while (expand) {
        int best = -1;
        float best_a = 0.0f;
        for (uint32_t j = 0; j < n; j++) {
            if (slot_leaf(nodes, slots[j]) || slot_dead(nodes, slots[j])) continue;
            uint32_t lf, lc;
            slot_list(nodes, slots[j], &lf, &lc);
            if (!run_ok(lf, lc, node_count)) return -1;
            if (n - 1 + lc > K) continue;
            const float a = half_area(slots[j]);
            if (best < 0 || a > best_a) { best = (int)j; best_a = a; }
        }
        if (best < 0) break;
        uint32_t lf, lc;
        slot_list(nodes, slots[best], &lf, &lc);
        // shift the slots after `best` right by lc - 1, then insert the list
        for (uint32_t j = n; j-- > (uint32_t)best + 1;) slots[j + lc - 1] = slots[j];
        for (uint32_t j = 0; j < lc; j++) slots[best + j] = make_slot(nodes, lf + j, 1);
        n += lc - 1;
    }
    *used = n;
    return 0;
}

struct Item {
    uint32_t f, c;        // the Node1 list the wide node holds
    int64_t parent_word;  // word index of the naming slot's flags word, -1: the root record
};

}  // namespace

extern "C" {

int sqo_nodek_convert(const sqo_node* nodes, uint32_t node_count, uint32_t root, uint32_t K, uint32_t expand,
                      uint32_t* out, uint32_t node_cap, uint32_t* out_nodes, uint32_t* out_levels,
                      sqo_nodek_root* root_out) {
    if (K != 2 && K != 4 && K != 8) return -3;
    if (root >= node_count) return -1;
    const sqo_node& R = nodes[root];
    if (!(R.flags & 1) && R.count != 0 && !run_ok(R.child, R.count, node_count)) return -1;
    sqo_nodek_root r;
    std::memset(&r, 0, sizeof r);
    for (int k = 0; k < 3; k++) { r.min[k] = R.min[k]; r.max[k] = R.max[k]; }
    *out_nodes = 0;
    *out_levels = 0;
    if (R.flags & 1) {
        r.kind = 1;
        r.child = R.child;
        r.count = R.count;
        *root_out = r;
        return 0;
    }
    if (R.count == 0) {  // an empty tree: the root box is all there is
        *root_out = r;
        return 0;
    }
    r.kind = 2;
    r.child = 0;
    const uint32_t W = 8 * K;
    std::vector<Item> level{{R.child, R.count, -1}}, next;
    uint32_t base = 0, levels = 0;
    std::vector<Slot> slots(K);
    while (!level.empty()) {
        if (++levels > kMaxLevels) return -1;
        next.clear();
        const uint32_t next_base = base + (uint32_t)level.size();
        for (size_t i = 0; i < level.size(); i++) {
            const uint32_t w = base + (uint32_t)i;
            if (w >= node_cap) return -4;
            uint32_t n = 0;
            const int e = collapse(nodes, node_count, level[i].f, level[i].c, K, expand, slots.data(), &n);
            if (e) return e;
            if (level[i].parent_word < 0) r.count = n;
            else out[level[i].parent_word] = n;
            uint32_t* o = out + (size_t)w * W;
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
                const Slot& sl = slots[j];
                for (int k = 0; k < 3; k++) {
                    o[k * K + j] = f2u(sl.mn[k]);
                    o[(4 + k) * K + j] = f2u(sl.mx[k]);
                }
                if (slot_leaf(nodes, sl)) {
                    const sqo_node& Y = nodes[sl.f];
                    o[3 * K + j] = (uint32_t)Y.count | ((uint32_t)Y.flags << 16);
                    o[7 * K + j] = Y.child;
                } else if (slot_dead(nodes, sl)) {
                    o[3 * K + j] = 0;
                    o[7 * K + j] = kUnused;
                } else {
                    uint32_t lf, lc;
                    slot_list(nodes, sl, &lf, &lc);
                    o[3 * K + j] = 0;  // the named node's used slots, written with it
                    o[7 * K + j] = next_base + (uint32_t)next.size();
                    next.push_back({lf, lc, (int64_t)w * W + 3 * K + j});
                }
            }
        }
        base = next_base;
        level.swap(next);
    }
    *out_nodes = base;
    *out_levels = levels;
    *root_out = r;
    return 0;
}

}  // extern "C"

namespace {

struct HitK {
    uint32_t count = 0, id = 0xFFFFFFFFu;
    float t = INFINITY;
};

// The reference slab test on a box given by its six floats; front = the
// entry distance (largest near-plane t).
inline bool slab_box(const float mn[3], const float mx[3], const sqo_ray& r, float* front_out) {
    sqo_node nd;
    std::memset(&nd, 0, sizeof nd);
    for (int k = 0; k < 3; k++) { nd.min[k] = mn[k]; nd.max[k] = mx[k]; }
    const float ix = 1.0f / r.dx, iy = 1.0f / r.dy, iz = 1.0f / r.dz;
    const float t1x = (mn[0] - r.ox) * ix, t2x = (mx[0] - r.ox) * ix;
    const float t1y = (mn[1] - r.oy) * iy, t2y = (mx[1] - r.oy) * iy;
    const float t1z = (mn[2] - r.oz) * iz, t2z = (mx[2] - r.oz) * iz;
    const float tnx = (t2x < t1x) ? t2x : t1x, tny = (t2y < t1y) ? t2y : t1y, tnz = (t2z < t1z) ? t2z : t1z;
    const float a = tnx > tny ? tnx : tny;
    *front_out = a > tnz ? a : tnz;
    return sqo_slab(&nd, &r) != 0;
}

inline void leafk(const sqo_tri* tris, uint32_t first, uint32_t count, const sqo_ray& r, HitK& h) {
    for (uint32_t j = 0; j < count; j++) {
        const uint32_t tid = first + j;
        float t;
        if (sqo_tri_test(&tris[tid], &r, &t) && t > r.tmin && t < r.tmax)
            if (t < h.t || (t == h.t && tid < h.id)) { h.t = t; h.id = tid; }
    }
}

struct Entry {
    uint32_t node, used;
    float t;
};

}  // namespace

extern "C" {

void sqo_nodek_trace(const uint32_t* nodes, uint32_t K, const sqo_nodek_root* root, const sqo_tri* tris,
                     const sqo_ray* rays, uint32_t nrays, uint32_t closest, uint32_t* counts, uint32_t* hit_id,
                     float* hit_t) {
    const uint32_t W = 8 * K;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        const sqo_ray& r = rays[i];
        HitK h;
        h.count = 1;
        float front;
        if (root->kind != 0 && slab_box(root->min, root->max, r, &front)) {
            if (root->kind == 1) {
                leafk(tris, root->child, root->count, r, h);
            } else {
                std::vector<Entry> st;
                st.push_back({root->child, root->count, front});
                while (!st.empty()) {
                    const Entry e = st.back();
                    st.pop_back();
                    const float cull = h.t + h.t * (1.0f / 1048576.0f);
                    if (closest && e.t > cull) continue;
                    const uint32_t* o = nodes + (size_t)e.node * W;
                    bool hit[8];
                    float f[8];
                    for (uint32_t j = 0; j < e.used; j++) {
                        float mn[3], mx[3];
                        for (int k = 0; k < 3; k++) {
                            mn[k] = u2f(o[k * K + j]);
                            mx[k] = u2f(o[(4 + k) * K + j]);
                        }
                        h.count++;
                        hit[j] = slab_box(mn, mx, r, &f[j]);
                        if (closest) hit[j] = hit[j] && !(f[j] > cull);
                    }
                    // leaves at once, in slot order
                    for (uint32_t j = 0; j < e.used; j++)
                        if (hit[j] && (o[3 * K + j] & 0x10000u)) leafk(tris, o[7 * K + j], o[3 * K + j] & 0xFFFFu, r, h);
                    // interior slots: farthest pushed first, so the nearest is visited next
                    uint32_t order[8], m = 0;
                    for (uint32_t j = 0; j < e.used; j++)
                        if (hit[j] && !(o[3 * K + j] & 0x10000u) && (o[3 * K + j] & 0xFFFFu)) order[m++] = j;
                    if (closest)
                        std::stable_sort(order, order + m, [&](uint32_t a, uint32_t b) { return f[a] > f[b]; });
                    for (uint32_t q = 0; q < m; q++) {
                        const uint32_t j = order[q];
                        st.push_back({o[7 * K + j], o[3 * K + j] & 0xFFFFu, f[j]});
                    }
                }
            }
        }
        if (counts) counts[i] = h.count;
        if (hit_id) hit_id[i] = h.id;
        if (hit_t) hit_t[i] = h.t;
    }
}

}  // extern "C"
