// node2.cpp — CPU restatement of the quantised two-child node format (Node2)
// and its traversal.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP conversion and traversal
// kernels. Built with -ffp-contract=off.
//
// The reference declares Node2 (aabb.h:44-74: a float centre, two 24-bit
// child fields, an 8-bit size, 3-bit counts, leaf bits and twelve 8-bit
// bounds) but never builds or traverses it, so the encoding is defined here
// (and mirrored by the GPU code):
//   * a Node2 describes two children a and b of one node; each child is a
//     leaf group (1..7 triangles, contiguous in a reordered triangle array,
//     child = first of them) or another Node2 (child = its index);
//   * size = e, the biased exponent of step = 2^(e-127), the smallest power
//     of two with 120 * step >= the largest distance from the centre to the
//     node's box faces (t = h * (1/120) rounded, e = exponent of t, plus one
//     if t is not a power of two; e >= 1);
//   * a bound q decodes to centre + (float)(q - 128) * step; each child's
//     min is the largest q whose decoded value is <= the true min, its max
//     the smallest q whose decoded value is >= the true max, so the decoded
//     box always contains the child's Node1 box.
// Conversion from a Node1 tree (deterministic breadth-first order): a list
// of Node1 nodes with one interior node is replaced by that node's child list
// (keeping the outer box); a single leaf, or a list of leaves holding at most
// max_leaf (1..7) triangles in total, becomes a leaf group; any other list becomes a Node2 whose halves
// (first c/2 nodes, the rest) are its children. Node2 indices follow levels,
// items in order, child a before b; leaf groups take triangle ranges in the
// same order, the root's group first.
// Traversal: the root box (as stored, unquantised) costs one test; each
// visited Node2 costs one test per child. Exhaustive mode visits every hit
// child; closest mode visits nearer children first and culls children whose
// entry distance exceeds the best hit by more than 2^-20 relative.
#include "oracle.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

inline float n2_min(float a, float b) { return (b < a) ? b : a; }
inline float n2_max(float a, float b) { return (a < b) ? b : a; }

struct B6 {
    float mn[3], mx[3];
};

B6 node_box(const sqo_node& n) {
    B6 b;
    for (int k = 0; k < 3; k++) { b.mn[k] = n.min[k]; b.mx[k] = n.max[k]; }
    return b;
}

B6 list_box(const sqo_node* nodes, uint32_t f, uint32_t c) {
    B6 b = node_box(nodes[f]);
    for (uint32_t i = 1; i < c; i++)
        for (int k = 0; k < 3; k++) {
            b.mn[k] = n2_min(b.mn[k], nodes[f + i].min[k]);
            b.mx[k] = n2_max(b.mx[k], nodes[f + i].max[k]);
        }
    return b;
}

enum { KIND_LEAF = 1, KIND_NODE = 2 };

struct Child {
    int kind;
    uint32_t f, c;  // the Node1 list it stands for
    B6 box;
    uint32_t tris;  // leaf: triangle count
};

// A list of Node1 nodes as a child of a Node2 (or the root).
int resolve(const sqo_node* nodes, uint32_t node_count, uint32_t f, uint32_t c, const B6& box, uint32_t max_leaf,
            Child& out) {
    for (uint32_t guard = 0; guard < 64 && c == 1; guard++) {
        if (f >= node_count) return -1;
        const sqo_node& x = nodes[f];
        if (x.flags & 1) break;
        f = x.child;
        c = x.count;
        if (c == 0) return -1;
    }
    if (c == 0 || f >= node_count || c > node_count - f) return -1;
    bool all_leaves = true;
    uint32_t tris = 0;
    for (uint32_t i = 0; i < c && all_leaves; i++) {
        const sqo_node& x = nodes[f + i];
        if (!(x.flags & 1)) all_leaves = false;
        else tris += x.count;
    }
    out.f = f;
    out.c = c;
    out.box = box;
    if (all_leaves && tris <= (c == 1 ? 7u : max_leaf)) {
        out.kind = KIND_LEAF;
        out.tris = tris;
        return 0;
    }
    if (c == 1) return -2;  // a single leaf of more than 7 triangles
    out.kind = KIND_NODE;
    out.tris = 0;
    return 0;
}

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline float dec(float c, uint32_t q, float step) { return c + (float)((int)q - 128) * step; }

// Exponent of the quantisation step for a box with centre c.
int step_exponent(const B6& b, const float c[3]) {
    float h = 0.0f;
    for (int k = 0; k < 3; k++) {
        h = n2_max(h, c[k] - b.mn[k]);
        h = n2_max(h, b.mx[k] - c[k]);
    }
    const float t = h * (1.0f / 120.0f);
    const uint32_t bits = f2u(t);
    const uint32_t E = (bits >> 23) & 0xFF, M = bits & 0x7FFFFF;
    int e = E == 0 ? 1 : (int)E + (M != 0 ? 1 : 0);
    return e;
}

int quantise(float c, float step, float inv, float lo, float hi, uint8_t* qlo_out, uint8_t* qhi_out) {
    float gl = floorf((lo - c) * inv) + 128.0f, gh = ceilf((hi - c) * inv) + 128.0f;
    int qlo = gl < 0.0f ? 0 : (gl > 255.0f ? 255 : (int)gl);
    int qhi = gh < 0.0f ? 0 : (gh > 255.0f ? 255 : (int)gh);
    while (qlo > 0 && dec(c, (uint32_t)qlo, step) > lo) qlo--;
    while (qlo < 255 && dec(c, (uint32_t)qlo + 1, step) <= lo) qlo++;
    while (qhi < 255 && dec(c, (uint32_t)qhi, step) < hi) qhi++;
    while (qhi > 0 && dec(c, (uint32_t)qhi - 1, step) >= hi) qhi--;
    if (dec(c, (uint32_t)qlo, step) > lo || dec(c, (uint32_t)qhi, step) < hi) return -3;
    *qlo_out = (uint8_t)qlo;
    *qhi_out = (uint8_t)qhi;
    return 0;
}

}  // namespace

extern "C" {

void sqo_node2_decode(const sqo_node2* n, float a_min[3], float a_max[3], float b_min[3], float b_max[3]) {
    const uint32_t e = n->w0 >> 24;
    const float step = u2f(e << 23);
    for (int k = 0; k < 3; k++) {
        a_min[k] = dec(n->c[k], n->q[4 * k + 0], step);
        a_max[k] = dec(n->c[k], n->q[4 * k + 1], step);
        b_min[k] = dec(n->c[k], n->q[4 * k + 2], step);
        b_max[k] = dec(n->c[k], n->q[4 * k + 3], step);
    }
}

int sqo_node2_convert(const sqo_node* nodes, uint32_t node_count, uint32_t root, const sqo_tri* tris,
                      uint32_t ntris, uint32_t max_leaf, sqo_node2* out, uint32_t node_cap, uint32_t* out_nodes,
                      sqo_tri* tris2, uint32_t* perm, uint32_t tri_cap, uint32_t* out_tris,
                      sqo_node2_root* root_out) {
    if (root >= node_count) return -1;
    max_leaf = max_leaf == 0 ? 1u : (max_leaf > 7 ? 7u : max_leaf);
    uint32_t tri_next = 0;
    auto emit_leaf = [&](const Child& ch, uint32_t* first) -> int {
        *first = tri_next;
        for (uint32_t i = 0; i < ch.c; i++) {
            const sqo_node& x = nodes[ch.f + i];
            for (uint32_t j = 0; j < x.count; j++) {
                const uint32_t tid = x.child + j;
                if (tid >= ntris) return -1;
                if (tri_next >= tri_cap) return -4;
                tris2[tri_next] = tris[tid];
                perm[tri_next] = tid;
                tri_next++;
            }
        }
        return 0;
    };
    const B6 rbox = node_box(nodes[root]);
    Child rc;
    int rc_ = resolve(nodes, node_count, root, 1, rbox, max_leaf, rc);
    if (rc_) return rc_;
    std::memset(root_out, 0, sizeof *root_out);
    for (int k = 0; k < 3; k++) { root_out->min[k] = rbox.mn[k]; root_out->max[k] = rbox.mx[k]; }
    uint32_t nn = 0;
    std::vector<Child> level, next;
    if (rc.kind == KIND_LEAF) {
        root_out->kind = 1;
        root_out->count = rc.tris;
        if ((rc_ = emit_leaf(rc, &root_out->child))) return rc_;
    } else {
        root_out->kind = 2;
        root_out->child = 0;
        level.push_back(rc);
    }
    while (!level.empty()) {
        const uint32_t base = nn, next_base = nn + (uint32_t)level.size();
        if (next_base > node_cap) return -5;
        next.clear();
        for (size_t i = 0; i < level.size(); i++) {
            const Child& it = level[i];
            const uint32_t h = it.c / 2;
            Child half[2];
            const uint32_t hf[2] = {it.f, it.f + h}, hc[2] = {h, it.c - h};
            for (int s = 0; s < 2; s++) {
                const B6 hb = list_box(nodes, hf[s], hc[s]);
                if ((rc_ = resolve(nodes, node_count, hf[s], hc[s], hb, max_leaf, half[s]))) return rc_;
            }
            sqo_node2& o = out[base + i];
            std::memset(&o, 0, sizeof o);
            float c[3];
            for (int k = 0; k < 3; k++) {
                if (!std::isfinite(it.box.mn[k]) || !std::isfinite(it.box.mx[k])) return -6;
                c[k] = (it.box.mn[k] + it.box.mx[k]) * 0.5f;
                o.c[k] = c[k];
            }
            const int e = step_exponent(it.box, c);
            if (e > 254) return -6;
            const float step = u2f((uint32_t)e << 23), inv = u2f((uint32_t)(254 - e) << 23);
            uint32_t child[2], cnt[2], leaf[2];
            for (int s = 0; s < 2; s++) {
                for (int k = 0; k < 3; k++)
                    if ((rc_ = quantise(c[k], step, inv, half[s].box.mn[k], half[s].box.mx[k], &o.q[4 * k + 2 * s],
                                        &o.q[4 * k + 2 * s + 1])))
                        return rc_;
                if (half[s].kind == KIND_LEAF) {
                    leaf[s] = 1;
                    cnt[s] = half[s].tris;
                    if ((rc_ = emit_leaf(half[s], &child[s]))) return rc_;
                } else {
                    leaf[s] = 0;
                    cnt[s] = 0;
                    child[s] = next_base + (uint32_t)next.size();
                    next.push_back(half[s]);
                }
                if (child[s] >= (1u << 24)) return -7;
            }
            o.w0 = child[0] | ((uint32_t)e << 24);
            o.w1 = child[1] | (cnt[0] << 24) | (leaf[0] << 27) | (cnt[1] << 28) | (leaf[1] << 31);
        }
        nn = next_base;
        level.swap(next);
    }
    *out_nodes = nn;
    *out_tris = tri_next;
    return 0;
}

}  // extern "C"

namespace {

struct Hit2 {
    uint32_t count = 0, id = 0xFFFFFFFFu;
    float t = INFINITY;
};

inline bool slab6(const float mn[3], const float mx[3], const sqo_ray& r, float* front_out) {
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

inline void leaf2(const sqo_tri* tris2, const uint32_t* perm, uint32_t first, uint32_t count, const sqo_ray& r,
                  Hit2& h) {
    for (uint32_t j = 0; j < count; j++) {
        float t;
        if (sqo_tri_test(&tris2[first + j], &r, &t) && t > r.tmin && t < r.tmax) {
            const uint32_t id = perm[first + j];
            if (t < h.t || (t == h.t && id < h.id)) { h.t = t; h.id = id; }
        }
    }
}

}  // namespace

extern "C" {

void sqo_node2_trace(const sqo_node2* nodes, const sqo_node2_root* root, const sqo_tri* tris2, const uint32_t* perm,
                     const sqo_ray* rays, uint32_t nrays, uint32_t closest, uint32_t* counts, uint32_t* hit_id,
                     float* hit_t) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        const sqo_ray& r = rays[i];
        Hit2 h;
        float front;
        h.count = 1;
        if (root->kind != 0 && slab6(root->min, root->max, r, &front)) {
            if (root->kind == 1) {
                leaf2(tris2, perm, root->child, root->count, r, h);
            } else {
                std::vector<uint32_t> st;
                st.push_back(root->child);
                while (!st.empty()) {
                    const sqo_node2& n = nodes[st.back()];
                    st.pop_back();
                    float amn[3], amx[3], bmn[3], bmx[3], fa, fb;
                    sqo_node2_decode(&n, amn, amx, bmn, bmx);
                    h.count += 2;
                    const float cull = h.t + h.t * (1.0f / 1048576.0f);
                    bool ha = slab6(amn, amx, r, &fa), hb = slab6(bmn, bmx, r, &fb);
                    if (closest) {
                        ha = ha && !(fa > cull);
                        hb = hb && !(fb > cull);
                    }
                    const uint32_t ca = n.w0 & 0xFFFFFF, cb = n.w1 & 0xFFFFFF;
                    const uint32_t na = (n.w1 >> 24) & 7, nb = (n.w1 >> 28) & 7;
                    const bool la = (n.w1 >> 27) & 1, lb = (n.w1 >> 31) & 1;
                    // nearer child first (closest); leaves are tested at once
                    const bool b_first = closest && fb < fa;
                    for (int s = 0; s < 2; s++) {
                        const bool is_a = (s == 0) != b_first;
                        const bool hit = is_a ? ha : hb;
                        if (!hit) continue;
                        if (is_a ? la : lb) leaf2(tris2, perm, is_a ? ca : cb, is_a ? na : nb, r, h);
                    }
                    // interior children: the nearer one is pushed last
                    for (int s = 1; s >= 0; s--) {
                        const bool is_a = (s == 0) != b_first;
                        const bool hit = is_a ? ha : hb;
                        if (hit && !(is_a ? la : lb)) st.push_back(is_a ? ca : cb);
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
