// trace.hip — wave64 ray traversal over Node1 trees.
//
// Semantics follow traverse / trace (tracer.cpp:47-65, 195-230):
//   * every visited node costs one box test (intersect_ray_aabb,
//     tracer.h:13-35: inv = 1/d, slab with the reference's ternary min/max,
//     hit = back >= front && back <= tmax);
//   * EXHAUSTIVE visits the children of every hit interior node, so the
//     per-ray test count equals the reference's num_tests for any visiting
//     order;
//   * triangles of hit leaves are tested with the reference Moller-Trumbore
//     (tracer.h:83-116, same fused-multiply-add pattern as its x86 build);
//     the closest valid hit (t > max(1e-9, tmin), t < tmax; ties -> lower
//     triangle id) is reported, which the reference computes but discards;
//   * CLOSEST culls lists whose entry distance exceeds the best hit by more
//     than 2^-20 relative and descends nearer children first.
//
// Execution: persistent workgroups; each wave pulls a 64-item tile (8x8
// pixels or 64 rays) from a global counter. Each lane keeps a stack of child
// lists: the newest `S` entries live in LDS (slot-major, bank-conflict
// free), older ones in a per-thread spill area in HBM. An entry packs
// (first child << 4 | count-1) or, for lists longer than 15, the parent
// node index with tag 15 (refetched on pop).
#include "trace_common.h"

namespace sq {

// Per-lane traversal state. One loop iteration tests up to two nodes of the
// lane's current child list, so lanes advance in lock step through binary
// lists; a lane whose ray is finished takes the next ray of the wave's tile
// (dynamic refill), which keeps lanes busy while rays of very different
// length share a wave.
struct LaneRay {
    RayF r;
    float ix, iy, iz;
    bool fast;
    uint32_t first, cnt;  // remaining part of the current child list
    uint32_t tests;
    uint32_t best_id;
    float best_t;
};

__device__ __forceinline__ void lane_begin(LaneRay& L, const RayF& r, uint32_t root, Stack& st) {
    L.r = r;
    L.ix = 1.0f / r.dx; L.iy = 1.0f / r.dy; L.iz = 1.0f / r.dz;
    L.fast = isfinite(L.ix) && isfinite(L.iy) && isfinite(L.iz) && isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz);
    L.first = root;
    L.cnt = 1;
    L.tests = 0;
    L.best_id = SQ_MISS;
    L.best_t = __int_as_float(0x7f800000);
    st.sp = 0;
}

template <class TA>
__device__ __forceinline__ void leaf_tris(const TA& a, const NodeV& n, LaneRay& L, uint32_t& tri_tests,
                                          bool& bad, uint32_t j0 = 0) {
    for (uint32_t j = j0; j < n.count; j++) {
        const uint32_t tid = n.child + j;
        if (tid >= a.tri_count) { bad = true; break; }
        float v[9];
        float t;
        tri_tests++;
        bool m;
        if (a.plk) {
            m = plucker_test(a.plk, tid, L.r, t);
        } else {
            fetch_tri(a, tid, v);
            m = tri_test(v, L.r, t);
        }
        if (m && t > L.r.tmin && t < L.r.tmax && (t < L.best_t || (t == L.best_t && tid < L.best_id))) {
            L.best_t = t;
            L.best_id = tid;
        }
    }
}

// Leaf triangles of up to N hit leaves (nodes nd[j] with hit[j]). The
// lane's hit leaves are first compacted in slot order, so that the k-th test
// block runs only when some lane has more than k hit leaves in this step
// (rather than whenever some lane hit a leaf in slot k); the first triangle of
// every hit leaf is fetched before any is tested, so the fetches are in flight
// together. The reference's leaves hold one triangle; the rest of a longer
// leaf is tested after the others.
template <int N, class TA>
__device__ __forceinline__ void leaf_batch(const TA& a, const NodeV* nd, const bool* hit, LaneRay& L,
                                           uint32_t& tri_tests, bool& bad) {
    if (a.plk) {
#pragma unroll
        for (int j = 0; j < N; j++)
            if (hit[j] && nd[j].leaf) leaf_tris(a, nd[j], L, tri_tests, bad);
        return;
    }
    bool lf[N];
    uint32_t pos[N], c = 0;
    bool longer = false;
#pragma unroll
    for (int j = 0; j < N; j++) {
        lf[j] = hit[j] && nd[j].leaf && nd[j].count > 0;
        if (lf[j] && nd[j].child >= a.tri_count) bad = true;
        pos[j] = c;
        c += lf[j] ? 1u : 0u;
        longer |= lf[j] && nd[j].count > 1;
    }
    if (bad) return;
    uint32_t ids[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        uint32_t v = 0;
#pragma unroll
        for (int j = k; j < N; j++) v = (lf[j] && pos[j] == (uint32_t)k) ? nd[j].child : v;
        ids[k] = v;
    }
    float tv[N][9];
#pragma unroll
    for (int k = 0; k < N; k++)
        if ((uint32_t)k < c) fetch_tri(a, ids[k], tv[k]);
#pragma unroll
    for (int k = 0; k < N; k++)
        if ((uint32_t)k < c) {
            const uint32_t tid = ids[k];
            float t;
            tri_tests++;
            if (tri_test(tv[k], L.r, t) && t > L.r.tmin && t < L.r.tmax &&
                (t < L.best_t || (t == L.best_t && tid < L.best_id))) {
                L.best_t = t;
                L.best_id = tid;
            }
        }
    if (longer) {
#pragma unroll
        for (int j = 0; j < N; j++)
            if (lf[j] && nd[j].count > 1) leaf_tris(a, nd[j], L, tri_tests, bad, 1);
    }
}

// One step: pop a list if the current one is exhausted, then test up to two
// of its nodes. Returns false when the ray is finished.
template <bool CLOSEST, class TA>
__device__ __forceinline__ bool lane_step(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    const float cull = L.best_t + L.best_t * (1.0f / 1048576.0f);
    while (L.cnt == 0) {
        if (st.sp == 0) return false;
        float te = 0.f;
        uint32_t e = pop<CLOSEST>(st, te);
        if (CLOSEST && te > cull) continue;
        if ((e & 15u) != 15u) {
            L.first = e >> 4;
            L.cnt = (e & 15u) + 1;
        } else {
            if ((e >> 4) >= a.node_count) { st.fault |= kFaultNode; return false; }
            NodeV p = fetch_node(a, e >> 4);
            L.first = p.child;
            L.cnt = p.count;
        }
    }
    const uint32_t i0 = L.first;
    const bool two = L.cnt >= 2;
    if (i0 >= a.node_count || (two && i0 + 1 >= a.node_count)) { st.fault |= kFaultNode; return false; }
    const NodeV n0 = fetch_node(a, i0);
    const NodeV n1 = fetch_node(a, i0 + (two ? 1u : 0u));  // unconditional: no branch between fetches
    L.first += two ? 2 : 1;
    L.cnt -= two ? 2 : 1;
    L.tests += two ? 2 : 1;
    float f0, f1 = 0.f;
    bool h0 = slab(n0, L.r, L.ix, L.iy, L.iz, L.fast, f0);
    bool h1 = two && slab(n1, L.r, L.ix, L.iy, L.iz, L.fast, f1);
    if (CLOSEST) {
        h0 = h0 && !(f0 > cull);
        h1 = h1 && !(f1 > cull);
    }
    bool bad = false;
    {
        const NodeV nd[2] = {n0, n1};
        const bool hh[2] = {h0, h1};
        leaf_batch<2>(a, nd, hh, L, tri_tests, bad);
    }
    if (bad) { st.fault |= kFaultTri; return false; }
    const bool c0 = h0 && !n0.leaf, c1 = h1 && !n1.leaf;
    if (c0 || c1) {
        uint32_t e0 = c0 ? encode_list(i0, n0) : 0u, e1 = c1 ? encode_list(i0 + 1, n1) : 0u;
        if (c0 && c1) {
            // the nearer list is pushed last and popped first
            bool swap = CLOSEST && f1 < f0;
            uint32_t ea = swap ? e0 : e1, eb = swap ? e1 : e0;
            float fa = swap ? f0 : f1, fb = swap ? f1 : f0;
            push<CLOSEST>(st, ea, fa);
            push<CLOSEST>(st, eb, fb);
        } else {
            push<CLOSEST>(st, c0 ? e0 : e1, c0 ? f0 : f1);
        }
        if (st.fault) return false;
    }
    return true;
}

// Exhaustive semantics: one step tests up to WL x 2 nodes, the next two of
// the current list and the next two after them (from the same list or, once it
// is used up, from the list on top of the stack when that names its nodes
// directly), so a lane has several independent node fetches in flight instead
// of one. Counts and the closest hit do not depend on the order.
//
// Stack bound: a step pops at most WL lists, all from the top (the deepest
// level), and pushes at most 2 WL lists of the next level, so at most WL
// lists per level stay behind: WL x depth + 2 WL entries.
//
// WL = 2 by default (three or four waves per SIMD); WL = 4 (199 VGPRs, two
// waves per SIMD) on request. With 8x8-pixel camera tiles, whose last costly
// tiles left half the waves idle, four lists at 8 waves per CU ran best
// (1080p x 4 on a 1M-triangle Morton tree: 10.35 ms at 4 lists, 11.28 at 2);
// with 4x4 tiles two lists at 12 waves per CU are faster (6.35 against
// 7.07 ms).
constexpr int kWideLists = 4, kListLists = 2;

template <int WL, class TA>
__device__ __forceinline__ bool lane_step_wide(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    while (L.cnt == 0) {
        if (st.sp == 0) return false;
        float te = 0.f;
        const uint32_t e = pop<false>(st, te);
        if ((e & 15u) != 15u) {
            L.first = e >> 4;
            L.cnt = (e & 15u) + 1;
        } else {
            if ((e >> 4) >= a.node_count) { st.fault |= kFaultNode; return false; }
            const NodeV p = fetch_node(a, e >> 4);
            L.first = p.child;
            L.cnt = p.count;
        }
    }
    uint32_t idx[WL], num[WL];
#pragma unroll
    for (int g = 0; g < WL; g++) {
        if (g > 0 && L.cnt == 0 && st.sp > 0) {
            const uint32_t top = st.lds[((st.sp - 1) & st.mask) * st.stride];
            if ((top & 15u) != 15u) {
                float te = 0.f;
                (void)pop<false>(st, te);
                L.first = top >> 4;
                L.cnt = (top & 15u) + 1;
            }
        }
        idx[g] = L.first;
        num[g] = L.cnt >= 2 ? 2u : L.cnt;
        L.first += num[g];
        L.cnt -= num[g];
        if (num[g] && (idx[g] >= a.node_count || num[g] > a.node_count - idx[g])) {
            st.fault |= kFaultNode;
            return false;
        }
        if (num[g] == 0) idx[g] = idx[0];  // in-bounds repeat, never tested
    }
    // all fetches unconditional (in-bounds indices; unused ones repeat a node)
    // so that they issue back to back without branches between them
    NodeV n[WL][2];
#pragma unroll
    for (int g = 0; g < WL; g++) {
        n[g][0] = fetch_node(a, idx[g]);
        n[g][1] = fetch_node(a, idx[g] + (num[g] == 2 ? 1u : 0u));
    }
    bool h[WL][2];
    float f;
#pragma unroll
    for (int g = 0; g < WL; g++) {
        L.tests += num[g];
        h[g][0] = num[g] >= 1 && slab(n[g][0], L.r, L.ix, L.iy, L.iz, L.fast, f);
        h[g][1] = num[g] == 2 && slab(n[g][1], L.r, L.ix, L.iy, L.iz, L.fast, f);
    }
    bool bad = false;
    {
        NodeV nd[WL * 2];
        bool hh[WL * 2];
#pragma unroll
        for (int g = 0; g < WL; g++)
#pragma unroll
            for (int k = 0; k < 2; k++) {
                nd[2 * g + k] = n[g][k];
                hh[2 * g + k] = h[g][k];
            }
        leaf_batch<WL * 2>(a, nd, hh, L, tri_tests, bad);
    }
    if (bad) { st.fault |= kFaultTri; return false; }
#pragma unroll
    for (int g = 0; g < WL; g++)
#pragma unroll
        for (int k = 0; k < 2; k++)
            if (h[g][k] && !n[g][k].leaf) push<false>(st, encode_list(idx[g] + k, n[g][k]), 0.f);
    return st.fault == 0;
}

// Closest semantics, two lists per step: the current list and, once it is
// used up, the list on top of the stack when that names its nodes directly;
// lists whose entry distance lies beyond the best hit (2^-20 relative) are
// dropped when popped, as in lane_step. The four node fetches issue back to
// back; the hit leaves are tested first, then the surviving child lists are
// pushed farthest first (the nearest is popped next), culled against the
// best hit those leaves left. Same stack bound as the two-list exhaustive
// step: 2 x depth + 4 entries.
__device__ __forceinline__ void cswap(float& fa, uint32_t& ea, float& fb, uint32_t& eb) {
    const bool s = fb > fa;  // larger distance first
    const float tf = s ? fb : fa;
    const uint32_t te = s ? eb : ea;
    fb = s ? fa : fb;
    eb = s ? ea : eb;
    fa = tf;
    ea = te;
}

template <class TA>
__device__ __forceinline__ bool lane_step_cwide(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    const float cull = L.best_t + L.best_t * (1.0f / 1048576.0f);
    while (L.cnt == 0) {
        if (st.sp == 0) return false;
        float te = 0.f;
        const uint32_t e = pop<true>(st, te);
        if (te > cull) continue;
        if ((e & 15u) != 15u) {
            L.first = e >> 4;
            L.cnt = (e & 15u) + 1;
        } else {
            if ((e >> 4) >= a.node_count) { st.fault |= kFaultNode; return false; }
            const NodeV p = fetch_node(a, e >> 4);
            L.first = p.child;
            L.cnt = p.count;
        }
    }
    uint32_t idx[2], num[2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
        if (g > 0 && L.cnt == 0 && st.sp > 0) {
            const uint32_t top = st.lds[((st.sp - 1) & st.mask) * st.stride];
            if ((top & 15u) != 15u) {
                float te = 0.f;
                (void)pop<true>(st, te);
                if (!(te > cull)) {
                    L.first = top >> 4;
                    L.cnt = (top & 15u) + 1;
                }
            }
        }
        idx[g] = L.first;
        num[g] = L.cnt >= 2 ? 2u : L.cnt;
        L.first += num[g];
        L.cnt -= num[g];
        if (num[g] && (idx[g] >= a.node_count || num[g] > a.node_count - idx[g])) {
            st.fault |= kFaultNode;
            return false;
        }
        if (num[g] == 0) idx[g] = idx[0];  // in-bounds repeat, never tested
    }
    NodeV n[2][2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
        n[g][0] = fetch_node(a, idx[g]);
        n[g][1] = fetch_node(a, idx[g] + (num[g] == 2 ? 1u : 0u));
    }
    bool h[2][2];
    float f[2][2];
#pragma unroll
    for (int g = 0; g < 2; g++) {
        L.tests += num[g];
        f[g][0] = f[g][1] = 0.f;
        h[g][0] = num[g] >= 1 && slab(n[g][0], L.r, L.ix, L.iy, L.iz, L.fast, f[g][0]) && !(f[g][0] > cull);
        h[g][1] = num[g] == 2 && slab(n[g][1], L.r, L.ix, L.iy, L.iz, L.fast, f[g][1]) && !(f[g][1] > cull);
    }
    bool bad = false;
    {
        const NodeV nd[4] = {n[0][0], n[0][1], n[1][0], n[1][1]};
        const bool hh[4] = {h[0][0], h[0][1], h[1][0], h[1][1]};
        leaf_batch<4>(a, nd, hh, L, tri_tests, bad);
    }
    if (bad) { st.fault |= kFaultTri; return false; }
    const float cull2 = L.best_t + L.best_t * (1.0f / 1048576.0f);
    const float inf = __int_as_float(0x7f800000);
    float fk[4];
    uint32_t ek[4];
#pragma unroll
    for (int g = 0; g < 2; g++)
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const bool c = h[g][k] && !n[g][k].leaf && !(f[g][k] > cull2);
            fk[2 * g + k] = c ? f[g][k] : -inf;  // absent entries sort last and are not pushed
            ek[2 * g + k] = c ? encode_list(idx[g] + k, n[g][k]) : 0u;
        }
    // sorting network, largest distance first
    cswap(fk[0], ek[0], fk[1], ek[1]);
    cswap(fk[2], ek[2], fk[3], ek[3]);
    cswap(fk[0], ek[0], fk[2], ek[2]);
    cswap(fk[1], ek[1], fk[3], ek[3]);
    cswap(fk[1], ek[1], fk[2], ek[2]);
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (fk[k] != -inf) push<true>(st, ek[k], fk[k]);
    return st.fault == 0;
}

// Node2 trees: triangles of a leaf group; hit ids are original indices.
template <class TA>
__device__ __forceinline__ void leaf_tris2(const TA& a, uint32_t first, uint32_t count, LaneRay& L,
                                           uint32_t& tri_tests, bool& bad) {
    for (uint32_t j = 0; j < count; j++) {
        const uint32_t k = first + j;
        if (k >= a.tri_count) { bad = true; break; }
        float v[9];
        fetch_tri(a, k, v);
        float t;
        tri_tests++;
        if (tri_test(v, L.r, t) && t > L.r.tmin && t < L.r.tmax && t <= L.best_t) {
            const uint32_t id = a.perm[k];
            if (t < L.best_t || id < L.best_id) {
                L.best_t = t;
                L.best_id = id;
            }
        }
    }
}

// Decoded child boxes of a Node2 (c + (q - 128) * 2^(size - 127)).
struct Node2V {
    NodeV a, b;
    uint32_t ca, cb, cnta, cntb;
    bool la, lb;
};

__device__ __forceinline__ Node2V load_node2(const sq_node2* nodes, uint32_t e) {
    const float4* p = (const float4*)(nodes + e);
    const float4 A = p[0], B = p[1];
    const uint32_t w0 = __float_as_uint(A.w), w1 = __float_as_uint(B.x);
    const uint32_t qw[3] = {__float_as_uint(B.y), __float_as_uint(B.z), __float_as_uint(B.w)};
    const float c[3] = {A.x, A.y, A.z};
    const float step = __uint_as_float((w0 >> 24) << 23);
    Node2V n;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        n.a.mn[k] = c[k] + (float)((int)(qw[k] & 0xFFu) - 128) * step;
        n.a.mx[k] = c[k] + (float)((int)((qw[k] >> 8) & 0xFFu) - 128) * step;
        n.b.mn[k] = c[k] + (float)((int)((qw[k] >> 16) & 0xFFu) - 128) * step;
        n.b.mx[k] = c[k] + (float)((int)(qw[k] >> 24) - 128) * step;
    }
    n.ca = w0 & 0xFFFFFFu;
    n.cb = w1 & 0xFFFFFFu;
    n.cnta = (w1 >> 24) & 7u;
    n.cntb = (w1 >> 28) & 7u;
    n.la = (w1 >> 27) & 1u;
    n.lb = (w1 >> 31) & 1u;
    return n;
}

// Leaf groups of a Node2 step (first triangle first[j], cnt[j] triangles,
// hit[j]): the lane's hit groups are compacted, the first triangle of each is
// fetched before any is tested, and the rest of longer groups follow.
template <int N, class TA>
__device__ __forceinline__ void leaf_batch2(const TA& a, const uint32_t* first, const uint32_t* cnt,
                                            const bool* hit, LaneRay& L, uint32_t& tri_tests, bool& bad) {
    bool lf[N];
    uint32_t pos[N], c = 0;
    bool longer = false;
#pragma unroll
    for (int j = 0; j < N; j++) {
        lf[j] = hit[j] && cnt[j] > 0;
        if (lf[j] && first[j] >= a.tri_count) bad = true;
        pos[j] = c;
        c += lf[j] ? 1u : 0u;
        longer |= lf[j] && cnt[j] > 1;
    }
    if (bad) return;
    uint32_t ids[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        uint32_t v = 0;
#pragma unroll
        for (int j = k; j < N; j++) v = (lf[j] && pos[j] == (uint32_t)k) ? first[j] : v;
        ids[k] = v;
    }
    float tv[N][9];
#pragma unroll
    for (int k = 0; k < N; k++)
        if ((uint32_t)k < c) fetch_tri(a, ids[k], tv[k]);
#pragma unroll
    for (int k = 0; k < N; k++)
        if ((uint32_t)k < c) {
            float t;
            tri_tests++;
            if (tri_test(tv[k], L.r, t) && t > L.r.tmin && t < L.r.tmax && t <= L.best_t) {
                const uint32_t id = a.perm[ids[k]];
                if (t < L.best_t || id < L.best_id) {
                    L.best_t = t;
                    L.best_id = id;
                }
            }
        }
    if (longer) {
#pragma unroll
        for (int j = 0; j < N; j++)
            if (lf[j] && cnt[j] > 1) leaf_tris2(a, first[j] + 1, cnt[j] - 1, L, tri_tests, bad);
    }
}

// Exhaustive Node2 step: up to WL Node2s (independent 32-byte fetches) per
// step, the current one and the next ones on the stack; a step pops at most
// WL entries and pushes at most 2 WL, so at most WL per level stay behind.
template <int WL, class TA>
__device__ __forceinline__ bool lane_step2_wide(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    float te = 0.f;
    if (st.sp == 0) return false;
    uint32_t e[WL];
    bool use[WL];
#pragma unroll
    for (int g = 0; g < WL; g++) {
        use[g] = g == 0 || st.sp > 0;
        e[g] = use[g] ? pop<false>(st, te) : 0u;
        if (use[g] && e[g] >= a.node_count) { st.fault |= kFaultNode; return false; }
        if (!use[g]) e[g] = e[0];  // in-bounds repeat, never tested
    }
    Node2V n[WL];
#pragma unroll
    for (int g = 0; g < WL; g++) n[g] = load_node2(a.nodes2, e[g]);  // unconditional: no branch between fetches
    bool h[2 * WL];
    uint32_t first[2 * WL], cnt[2 * WL];
    float f;
#pragma unroll
    for (int g = 0; g < WL; g++) {
        L.tests += use[g] ? 2u : 0u;
        h[2 * g] = use[g] && slab(n[g].a, L.r, L.ix, L.iy, L.iz, L.fast, f);
        h[2 * g + 1] = use[g] && slab(n[g].b, L.r, L.ix, L.iy, L.iz, L.fast, f);
        first[2 * g] = n[g].ca;
        first[2 * g + 1] = n[g].cb;
        cnt[2 * g] = n[g].la ? n[g].cnta : 0u;
        cnt[2 * g + 1] = n[g].lb ? n[g].cntb : 0u;
    }
    bool bad = false;
    {
        bool hl[2 * WL];
#pragma unroll
        for (int g = 0; g < WL; g++) {
            hl[2 * g] = h[2 * g] && n[g].la;
            hl[2 * g + 1] = h[2 * g + 1] && n[g].lb;
        }
        leaf_batch2<2 * WL>(a, first, cnt, hl, L, tri_tests, bad);
    }
    if (bad) { st.fault |= kFaultTri; return false; }
#pragma unroll
    for (int g = 0; g < WL; g++) {
        if (h[2 * g] && !n[g].la) push<false>(st, n[g].ca, 0.f);
        if (h[2 * g + 1] && !n[g].lb) push<false>(st, n[g].cb, 0.f);
    }
    return st.fault == 0;
}

// Node2 step: the root box on the first step, then one Node2 (both children
// from one 32-byte fetch) per step; exhaustive semantics take two.
// Node2 trees, closest semantics, camera rays: two Node2s per step (four
// child boxes, both fetches issued together), leaf groups tested first, then
// the surviving children pushed farthest first, as in lane_step_cwide.
template <class TA>
__device__ __forceinline__ bool lane_step2_cwide(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    const float cull = L.best_t + L.best_t * (1.0f / 1048576.0f);
    uint32_t e[2];
    uint32_t got = 0;
    while (got < 2 && st.sp > 0) {
        float te = 0.f;
        const uint32_t x = pop<true>(st, te);
        if (te > cull) continue;
        if (x >= a.node_count) { st.fault |= kFaultNode; return false; }
        e[got++] = x;
        if (got == 1 && st.sp == 0) break;
    }
    if (got == 0) return false;
    if (got == 1) e[1] = e[0];  // in-bounds repeat, never tested
    const Node2V n0 = load_node2(a.nodes2, e[0]);
    const Node2V n1 = load_node2(a.nodes2, e[1]);
    const bool two = got == 2;
    L.tests += two ? 4 : 2;
    float f[4];
    bool h[4];
    h[0] = slab(n0.a, L.r, L.ix, L.iy, L.iz, L.fast, f[0]) && !(f[0] > cull);
    h[1] = slab(n0.b, L.r, L.ix, L.iy, L.iz, L.fast, f[1]) && !(f[1] > cull);
    h[2] = two && slab(n1.a, L.r, L.ix, L.iy, L.iz, L.fast, f[2]) && !(f[2] > cull);
    h[3] = two && slab(n1.b, L.r, L.ix, L.iy, L.iz, L.fast, f[3]) && !(f[3] > cull);
    const uint32_t cc[4] = {n0.ca, n0.cb, n1.ca, n1.cb}, cn[4] = {n0.cnta, n0.cntb, n1.cnta, n1.cntb};
    const bool lf[4] = {n0.la, n0.lb, n1.la, n1.lb};
    bool bad = false;
    {
        const uint32_t cs[4] = {lf[0] ? cn[0] : 0u, lf[1] ? cn[1] : 0u, lf[2] ? cn[2] : 0u, lf[3] ? cn[3] : 0u};
        const bool hl[4] = {h[0] && lf[0], h[1] && lf[1], h[2] && lf[2], h[3] && lf[3]};
        leaf_batch2<4>(a, cc, cs, hl, L, tri_tests, bad);
    }
    if (bad) { st.fault |= kFaultTri; return false; }
    const float cull2 = L.best_t + L.best_t * (1.0f / 1048576.0f);
    const float inf = __int_as_float(0x7f800000);
    float fk[4];
    uint32_t ek[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool c = h[k] && !lf[k] && !(f[k] > cull2);
        fk[k] = c ? f[k] : -inf;  // absent entries sort last and are not pushed
        ek[k] = c ? cc[k] : 0u;
    }
    cswap(fk[0], ek[0], fk[1], ek[1]);
    cswap(fk[2], ek[2], fk[3], ek[3]);
    cswap(fk[0], ek[0], fk[2], ek[2]);
    cswap(fk[1], ek[1], fk[3], ek[3]);
    cswap(fk[1], ek[1], fk[2], ek[2]);
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (fk[k] != -inf) push<true>(st, ek[k], fk[k]);
    return st.fault == 0;
}

template <bool CLOSEST, int WL, class TA>
__device__ __forceinline__ bool lane_step2(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    if (L.cnt) {
        L.cnt = 0;
        L.tests += 1;
        NodeV rb;
        for (int k = 0; k < 3; k++) { rb.mn[k] = a.rb[k]; rb.mx[k] = a.rb[3 + k]; }
        float f;
        if (!slab(rb, L.r, L.ix, L.iy, L.iz, L.fast, f)) return false;
        if (a.root_kind == 1) {
            bool bad = false;
            leaf_tris2(a, a.root, a.root_count, L, tri_tests, bad);
            if (bad) st.fault |= kFaultTri;
            return false;
        }
        if (a.root_kind != 2) return false;
        push<CLOSEST>(st, a.root, f);
        return st.fault == 0;
    }
    if (!CLOSEST) return lane_step2_wide<WL>(a, L, st, tri_tests);
    if (WL >= 2 && a.camera) return lane_step2_cwide(a, L, st, tri_tests);
    const float cull = L.best_t + L.best_t * (1.0f / 1048576.0f);
    uint32_t e = 0;
    while (true) {
        if (st.sp == 0) return false;
        float te = 0.f;
        e = pop<CLOSEST>(st, te);
        if (CLOSEST && te > cull) continue;
        break;
    }
    if (e >= a.node_count) { st.fault |= kFaultNode; return false; }
    const Node2V n = load_node2(a.nodes2, e);
    L.tests += 2;
    float fa, fb;
    bool ha = slab(n.a, L.r, L.ix, L.iy, L.iz, L.fast, fa);
    bool hb = slab(n.b, L.r, L.ix, L.iy, L.iz, L.fast, fb);
    if (CLOSEST) {
        ha = ha && !(fa > cull);
        hb = hb && !(fb > cull);
    }
    const uint32_t ca = n.ca, cb = n.cb, cnta = n.cnta, cntb = n.cntb;
    const bool la = n.la, lb = n.lb;
    bool bad = false;
    {
        const uint32_t fs[2] = {ca, cb}, cs[2] = {la ? cnta : 0u, lb ? cntb : 0u};
        const bool hl[2] = {ha && la, hb && lb};
        leaf_batch2<2>(a, fs, cs, hl, L, tri_tests, bad);
    }
    if (bad) { st.fault |= kFaultTri; return false; }
    const bool pa = ha && !la, pb = hb && !lb;
    if (pa && pb) {
        const bool swap = CLOSEST && fb < fa;  // the nearer child is pushed last
        push<CLOSEST>(st, swap ? ca : cb, swap ? fa : fb);
        push<CLOSEST>(st, swap ? cb : ca, swap ? fb : fa);
    } else if (pa) {
        push<CLOSEST>(st, ca, fa);
    } else if (pb) {
        push<CLOSEST>(st, cb, fb);
    }
    return st.fault == 0;
}

// Node<K> trees (K = 4, 8): a node is K slots in structure-of-arrays order
// (min x/y/z, words, max x/y/z, children: K / 4 16-byte loads per field), so
// one fetch brings K boxes. Stack entries: node << 3 | (used slots - 1).
template <int K>
struct NodeKV {
    float mn[3][K], mx[3][K];
    uint32_t w[K], child[K];
};

template <int K>
__device__ __forceinline__ void load_nodek(const uint32_t* __restrict__ nodes, uint32_t i, NodeKV<K>& n) {
    constexpr int Q = K / 4;
    const float4* p = (const float4*)(nodes + (size_t)i * (8 * K));
#pragma unroll
    for (int f = 0; f < 8; f++)
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const float4 v = p[f * Q + q];
            const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int j = 4 * q + t;
                if (f < 3) n.mn[f][j] = e[t];
                else if (f == 3) n.w[j] = __float_as_uint(e[t]);
                else if (f < 7) n.mx[f - 4][j] = e[t];
                else n.child[j] = __float_as_uint(e[t]);
            }
        }
}

// Sorting networks, largest distance first (Batcher's odd-even merge sort).
template <int N>
__device__ __forceinline__ void sort_desc(float* f, uint32_t* e) {
    static_assert(N == 4 || N == 8, "4 or 8 entries");
    if constexpr (N == 4) {
        cswap(f[0], e[0], f[1], e[1]);
        cswap(f[2], e[2], f[3], e[3]);
        cswap(f[0], e[0], f[2], e[2]);
        cswap(f[1], e[1], f[3], e[3]);
        cswap(f[1], e[1], f[2], e[2]);
    } else {
        constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                    {0, 4}, {1, 5}, {2, 6}, {3, 7}, {2, 4}, {3, 5}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
        for (int c = 0; c < 19; c++) cswap(f[net[c][0]], e[net[c][0]], f[net[c][1]], e[net[c][1]]);
    }
}

// Node<K> steps: the root box on the first step, then WL nodes per step
// (the entries on top of the stack; closest: entries beyond the best hit are
// dropped as they are popped), their fetches issued together; the used slots
// are tested, hit leaves tested (compacted, first triangles fetched
// together), hit interior slots pushed (closest: culled against the best hit
// after the leaves, farthest first). A step pops WL entries and pushes at
// most WL K: WL (K - 1) x depth + WL K entries bound the stack.
template <bool CLOSEST, int K, int WL, class TA>
__device__ __forceinline__ bool lane_stepk(const TA& a, LaneRay& L, Stack& st, uint32_t& tri_tests) {
    static_assert(WL == 1 || (WL == 2 && K == 4), "one node per step, or two Node<4>s");
    constexpr int N = K * WL;
    if (L.cnt) {
        L.cnt = 0;
        L.tests += 1;
        NodeV rb;
        for (int k = 0; k < 3; k++) { rb.mn[k] = a.rb[k]; rb.mx[k] = a.rb[3 + k]; }
        float f;
        if (!slab(rb, L.r, L.ix, L.iy, L.iz, L.fast, f)) return false;
        if (a.root_kind == 1) {
            NodeV lf;
            lf.child = a.root;
            lf.count = a.root_count;
            lf.leaf = 1;
            bool bad = false;
            leaf_tris(a, lf, L, tri_tests, bad);
            if (bad) st.fault |= kFaultTri;
            return false;
        }
        if (a.root_kind != 2 || a.root_count == 0) return false;
        push<CLOSEST>(st, (a.root << 3) | (a.root_count - 1), f);
        return st.fault == 0;
    }
    const float cull = L.best_t + L.best_t * (1.0f / 1048576.0f);
    uint32_t e[WL];
    int got = 0;
    while (got < WL && st.sp > 0) {
        float te = 0.f;
        const uint32_t x = pop<CLOSEST>(st, te);
        if (CLOSEST && te > cull) continue;
        e[got++] = x;
    }
    if (got == 0) return false;
    uint32_t idx[WL], used[WL];
#pragma unroll
    for (int g = 0; g < WL; g++) {
        const bool have = g < got;
        idx[g] = have ? e[g] >> 3 : idx[0];  // in-bounds repeat, never tested
        used[g] = have ? (e[g] & 7u) + 1 : 0u;
        if (idx[g] >= a.node_count) { st.fault |= kFaultNode; return false; }
    }
    NodeKV<K> n[WL];
#pragma unroll
    for (int g = 0; g < WL; g++) load_nodek<K>(a.nodesk, idx[g], n[g]);
    NodeV nd[N];
    bool h[N];
    float f[N];
#pragma unroll
    for (int g = 0; g < WL; g++) {
        L.tests += used[g];
#pragma unroll
        for (int j = 0; j < K; j++) {
            NodeV b;
            for (int k = 0; k < 3; k++) { b.mn[k] = n[g].mn[k][j]; b.mx[k] = n[g].mx[k][j]; }
            const int q = g * K + j;
            f[q] = 0.f;
            h[q] = (uint32_t)j < used[g] && slab(b, L.r, L.ix, L.iy, L.iz, L.fast, f[q]);
            if (CLOSEST) h[q] = h[q] && !(f[q] > cull);
            nd[q].leaf = (n[g].w[j] >> 16) & 1u;
            nd[q].count = n[g].w[j] & 0xFFFFu;
            nd[q].child = n[g].child[j];
        }
    }
    bool bad = false;
    leaf_batch<N>(a, nd, h, L, tri_tests, bad);
    if (bad) { st.fault |= kFaultTri; return false; }
    if (!CLOSEST) {
#pragma unroll
        for (int q = 0; q < N; q++)
            if (h[q] && !nd[q].leaf && nd[q].count) push<false>(st, (nd[q].child << 3) | (nd[q].count - 1), 0.f);
        return st.fault == 0;
    }
    const float cull2 = L.best_t + L.best_t * (1.0f / 1048576.0f);
    const float inf = __int_as_float(0x7f800000);
    float fk[N];
    uint32_t ek[N];
#pragma unroll
    for (int q = 0; q < N; q++) {
        const bool c = h[q] && !nd[q].leaf && nd[q].count && !(f[q] > cull2);
        fk[q] = c ? f[q] : -inf;  // absent entries sort last and are not pushed
        ek[q] = c ? ((nd[q].child << 3) | (nd[q].count - 1)) : 0u;
    }
    sort_desc<N>(fk, ek);
#pragma unroll
    for (int q = 0; q < N; q++)
        if (fk[q] != -inf) push<true>(st, ek[q], fk[q]);
    return st.fault == 0;
}

constexpr uint32_t kRayTile = 256;  // rays per tile in ray-list mode (fewer for short lists)

// One tile of job J: the wave's lanes take the tile's rays one after
// another (per-lane refill) until all are traced.
//
// Camera tiles are dealt a row of tiles at a time from the middle row
// outwards (rows c, c - 1, c + 1, c - 2, ... with c = tiles_y / 2): the
// camera looks at the scene-box centre, so the costly tiles come first and
// the cheap outer rows fill the end of the launch, when waves run out of
// work at different times.
template <bool CLOSEST, int FMT, int WL, class TA>
__device__ __forceinline__ void trace_tile(const TA& a, uint32_t tile, uint32_t lane, Stack& st,
                                           uint32_t* s_cnt, uint32_t& tests_total, uint32_t& tris_total,
                                           uint32_t& faults, uint32_t* job_fail) {
    uint32_t tx = 0, ty = 0, tile_rays;
    const uint32_t ts = a.tile_shift ? a.tile_shift : 3u;  // tile edge 1 << ts pixels
    const uint32_t tpix = 1u << (2 * ts);                  // pixels per tile (at most 64)
    if (a.camera) {
        const uint32_t rank = tile / a.tiles_x;
        const uint32_t half = (rank + 1) >> 1, mid = (a.num_tiles / a.tiles_x) >> 1;
        tx = tile - rank * a.tiles_x;
        ty = (rank & 1) ? mid - half : mid + half;
        tile_rays = tpix * a.spp;
        s_cnt[lane] = 0;
    } else {
        const uint32_t rt = a.ray_tile ? a.ray_tile : kRayTile;
        size_t base = (size_t)tile * rt;
        tile_rays = (uint32_t)((a.nrays - base) < rt ? (a.nrays - base) : rt);
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t next = 0;
    bool active = false;
    uint32_t my_ray = 0;
    LaneRay L;
    while (true) {
        const uint64_t need = __ballot(!active);
        if (need != 0 && next < tile_rays) {
            const uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
            const uint32_t r = next + rank;
            if (!active && r < tile_rays) {
                if (a.camera) {
                    const uint32_t p = r & (tpix - 1u), smp = r >> (2 * ts);
                    const uint32_t px = (tx << ts) + (p & ((1u << ts) - 1u)), py = (ty << ts) + (p >> ts);
                    if (px < a.w && py < a.h) {
                        lane_begin(L, camera_ray(a, px, py, smp), a.root, st);
                        my_ray = r;
                        active = true;
                    }
                } else {
                    const size_t ri = (size_t)tile * (a.ray_tile ? a.ray_tile : kRayTile) + r;
                    const sq_ray q = a.rays[ri];
                    RayF rf;
                    rf.ox = q.origin.x; rf.oy = q.origin.y; rf.oz = q.origin.z;
                    rf.dx = q.dir.x; rf.dy = q.dir.y; rf.dz = q.dir.z;
                    rf.tmin = q.tmin; rf.tmax = q.tmax;
                    lane_begin(L, rf, a.root, st);
                    my_ray = r;
                    active = true;
                }
            }
            const uint32_t taken = (uint32_t)__popcll(need);
            next = (next + taken < tile_rays) ? next + taken : tile_rays;
        }
        if (__ballot(active) == 0) {
            if (next >= tile_rays) break;
            continue;
        }
        if (active) {
            bool more;
            if constexpr (FMT >= 4) {
                more = lane_stepk<CLOSEST, FMT, WL>(a, L, st, tris_total);
            } else if constexpr (FMT == 2) {
                more = lane_step2<CLOSEST, WL>(a, L, st, tris_total);
            } else {
                more = CLOSEST ? (WL >= 2 && a.camera ? lane_step_cwide(a, L, st, tris_total)
                                                      : lane_step<CLOSEST>(a, L, st, tris_total))
                               : lane_step_wide<WL>(a, L, st, tris_total);
            }
            if (!more) {
                // ray finished
                if (st.fault) {
                    faults |= st.fault;
                    if (job_fail) atomicOr(job_fail, st.fault);
                    st.fault = 0;
                }
                tests_total += L.tests;
                size_t ri;
                if (a.camera) {
                    const uint32_t p = my_ray & (tpix - 1u), smp = my_ray >> (2 * ts);
                    const uint32_t px = (tx << ts) + (p & ((1u << ts) - 1u)), py = (ty << ts) + (p >> ts);
                    ri = ((size_t)py * a.w + px) * a.spp + smp;
                    atomicAdd(&s_cnt[p], L.tests);
                } else {
                    ri = (size_t)tile * (a.ray_tile ? a.ray_tile : kRayTile) + my_ray;
                    if (a.counts) a.counts[ri] = L.tests;
                }
                if (a.hit_id) a.hit_id[ri] = L.best_id;
                if (a.hit_t) a.hit_t[ri] = L.best_t;
                active = false;
            }
        }
    }
    if (a.camera) {
        __builtin_amdgcn_wave_barrier();
        const uint32_t px = (tx << ts) + (lane & ((1u << ts) - 1u)), py = (ty << ts) + (lane >> ts);
        if (lane < tpix && px < a.w && py < a.h) {
            const uint32_t sum = s_cnt[lane];
            const size_t p = (size_t)py * a.w + px;
            if (a.counts) a.counts[p] = sum;
            if (a.rgb) {
                uint8_t b = heat_byte(sum, a.spp);
                a.rgb[3 * p] = b; a.rgb[3 * p + 1] = b; a.rgb[3 * p + 2] = b;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// MULTI: the launch covers a.njobs jobs (tree + rays + outputs each); tile t
// of the launch is tile t - job_tiles[j] of the job j with job_tiles[j] <= t
// < job_tiles[j + 1]. Waves steal tiles from one counter across all jobs.
template <bool CLOSEST, bool MULTI, int FMT = 1, int WL = kListLists, bool NARROW = false>
__global__ __launch_bounds__(128) void k_trace(TraceArgs a) {
    using TA = typename std::conditional<NARROW, TraceArgsN, TraceArgs>::type;
    extern __shared__ uint32_t s_dyn[];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Stack st;
    st.stride = blockDim.x;
    st.S = a.slots;
    st.mask = a.slots - 1;
    st.lds = s_dyn + threadIdx.x;
    st.lds_t = CLOSEST ? (float*)(s_dyn + a.slots * blockDim.x) + threadIdx.x : nullptr;
    uint32_t* s_cnt = s_dyn + a.slots * blockDim.x * (CLOSEST ? 2 : 1) + wave * 64;
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    st.cap = a.spill_cap;
    st.spill = a.spill + gtid * (size_t)a.spill_cap * (CLOSEST ? 2 : 1);
    st.fault = 0;
    uint32_t tests_total = 0, tris_total = 0, faults = 0;
    while (true) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(&a.work[0], 1u);
        tile = __shfl(tile, 0, 64);
        if (tile >= a.num_tiles) break;
        if (MULTI) {
            // The job table is read-only here: constant-address-space loads
            // keep the wave-uniform job in scalar registers.
#if defined(__HIP_DEVICE_COMPILE__)
            typedef const __attribute__((address_space(4))) TraceArgs CJob;
            typedef const __attribute__((address_space(4))) uint32_t CU32;
#else
            typedef const TraceArgs CJob;  // host pass: never executed
            typedef const uint32_t CU32;
#endif
            const CJob* jobs = (const CJob*)a.jobs;
            const CU32* jt = (const CU32*)a.job_tiles;
            tile = __builtin_amdgcn_readfirstlane(tile);
            uint32_t lo = 0, hi = a.njobs;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (jt[mid] <= tile) lo = mid;
                else hi = mid;
            }
            TA J;
            static_cast<TraceArgs&>(J) = jobs[lo];
            trace_tile<CLOSEST, FMT, WL>(J, tile - jt[lo], lane, st, s_cnt, tests_total, tris_total, faults,
                                         a.job_fail ? a.job_fail + lo : nullptr);
        } else {
            trace_tile<CLOSEST, FMT, WL>(static_cast<const TA&>(a), tile, lane, st, s_cnt, tests_total, tris_total,
                                         faults, nullptr);
        }
    }
    if (faults) atomicOr(&a.work[1], faults);
    if (a.stats) {
        unsigned long long t0 = tests_total, t1 = tris_total;
        for (int off = 32; off > 0; off >>= 1) {
            t0 += __shfl_xor(t0, off, 64);
            t1 += __shfl_xor(t1, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&a.stats[0], t0);
            atomicAdd(&a.stats[1], t1);
        }
    }
}

__global__ void k_trace_init(uint32_t* work, unsigned long long* stats) {
    if (threadIdx.x < 4) work[threadIdx.x] = 0;
    if (stats && threadIdx.x < 2) stats[threadIdx.x] = 0;
}

int trace_fault(uint32_t bits, const char* what) {
    SQ_RET(SQ_ERR_OVERFLOW, "%s: rays abandoned:%s%s%s", what, (bits & kFaultStack) ? " stack bound exceeded" : "",
           (bits & kFaultNode) ? " child index outside the node array" : "",
           (bits & kFaultTri) ? " triangle index outside the scene" : "");
}

constexpr uint32_t kCameraWavesPerCu = 12, kCameraClosestWavesPerCu = 12, kWavesPerCu = 16;

int trace_plan(sq_ctx* c, const sq_trace_params* p, uint32_t depth, bool coherent, TracePlan* out, int fmt,
               bool ray_list) {
    TracePlan L;
    const bool node2 = fmt == 2, wide = fmt >= 4;
    L.closest = p && p->semantics == SQ_TRACE_CLOSEST;
    // closest hits: camera rays take two lists per step (134 VGPRs, three
    // waves per SIMD), ray lists one. A launch of one ray list uses the
    // one-list kernel, which fits 16 waves per CU (1M incoherent closest rays:
    // 772 ms at 16 waves, 1,100 ms at 12 with one or two lists per step);
    // many-job launches use the two-list kernel, which steps each job's rays
    // the way a launch of that job alone would.
    L.lists = L.closest && ray_list ? 1 : kListLists;
    if (wide) {
        // Node<K> trees: nodes per step (two only for K = 4)
        L.lists = 1;
        if (p && p->lists_per_step) {
            if (p->lists_per_step > (fmt == 4 ? 2u : 1u))
                SQ_RET(SQ_ERR_INVALID, "trace lists_per_step %u (Node<%d>: 1%s)", p->lists_per_step, fmt,
                       fmt == 4 ? " or 2" : "");
            L.lists = p->lists_per_step;
        }
    } else if (p && p->lists_per_step) {
        if (p->lists_per_step != kListLists && p->lists_per_step != kWideLists)
            SQ_RET(SQ_ERR_INVALID, "trace lists_per_step %u (2 or 4)", p->lists_per_step);
        if (!L.closest) L.lists = p->lists_per_step;
    }
    // Node<K> closest hits of camera rays take 16 LDS entries per lane, other
    // Node<K> launches 32 (1M incoherent closest rays on 10M clustered
    // triangles, Node<4>: 614 ms at 32, 675 at 16)
    uint32_t slots = p && p->stack_slots    ? p->stack_slots
                     : wide                 ? (L.closest && coherent ? 16 : 32)
                     : L.lists == kWideLists ? 32
                                             : 16;
    uint32_t pow2 = 1;
    while (pow2 < slots) pow2 <<= 1;
    L.slots = pow2 > 64 ? 64 : pow2;
    L.block = p && p->block_size ? p->block_size : 64;
    if (L.block != 64 && L.block != 128) SQ_RET(SQ_ERR_INVALID, "trace block size %u (64 or 128)", L.block);
    L.lds_bytes = ((size_t)L.slots * (L.closest ? 2 : 1) + 1) * L.block * 4;
    int per_cu = 0;
    if (fmt == 4 && L.lists == 2)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, L.closest ? k_trace<true, false, 4, 2> : k_trace<false, false, 4, 2>, L.block, L.lds_bytes);
    else if (fmt == 4)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, L.closest ? k_trace<true, false, 4, 1> : k_trace<false, false, 4, 1>, L.block, L.lds_bytes);
    else if (fmt == 8)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, L.closest ? k_trace<true, false, 8, 1> : k_trace<false, false, 8, 1>, L.block, L.lds_bytes);
    else if (node2 && L.closest && L.lists == 1)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<true, false, 2, 1>, L.block, L.lds_bytes);
    else if (node2 && L.closest)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<true, false, 2>, L.block, L.lds_bytes);
    else if (L.closest && L.lists == 1)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<true, false, 1, 1>, L.block, L.lds_bytes);
    else if (node2 && L.lists == kWideLists)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false, 2, kWideLists>, L.block,
                                                           L.lds_bytes);
    else if (node2)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false, 2>, L.block, L.lds_bytes);
    else if (L.closest)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<true, false>, L.block, L.lds_bytes);
    else if (L.lists == kWideLists)
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false, 1, kWideLists>, L.block,
                                                           L.lds_bytes);
    else
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false>, L.block, L.lds_bytes);
    if (per_cu < 1) per_cu = 8;
    // Default waves per CU, measured on 1080p x 4 jittered over the 1M soup
    // (Morton63 tree; 4x4-pixel tiles): exhaustive Node1 steps run best at 12
    // (two lists per step: 6.35 ms at 12, 6.59 at 16, 7.25 at 10, 7.77 at 8;
    // four lists per step reach only 8 waves, 7.07 ms), Node2 steps at 16
    // (5.78 ms against 5.91 at 12); closest hits on 8x8 tiles, two lists
    // per step, at 12, all that fits (Morton 1.56 ms against 1.68 at 8, SAH
    // 1.55 / 1.54; one list per step: 2.34 / 1.97 ms at 8; Node2 trees, two
    // Node2s per step: 1.67 ms at 12, 1.77 at 8, 2.01 at 16); ray lists
    // (incoherent secondary rays on 10M clustered
    // triangles: 0.78 s at 16 against 1.17 s at 8) and many-scene launches
    // (256 scenes: 30 ms at 16 against 49 ms at 8) want 16.
    const uint32_t waves = !coherent ? kWavesPerCu
                           : L.closest ? kCameraClosestWavesPerCu
                           : node2 || wide ? kWavesPerCu
                                           : kCameraWavesPerCu;
    const int target = (int)(waves * 64 / L.block);
    if (per_cu > target) per_cu = target;
    if (p && p->blocks_per_cu) {
        if (p->blocks_per_cu > 64) SQ_RET(SQ_ERR_INVALID, "trace blocks_per_cu %u (1..64)", p->blocks_per_cu);
        per_cu = (int)p->blocks_per_cu;
    }
    L.grid = (uint32_t)c->num_cus * (uint32_t)per_cu;
    if (depth == 0) depth = 64;
    // stack entries a ray can need: WL x depth + 2 WL for the exhaustive step,
    // depth + 2 for the closest step, WL ((K - 1) x depth + K) for Node<K> steps
    uint32_t cap = wide ? L.lists * ((uint32_t)(fmt - 1) * depth + (uint32_t)fmt) + 16
                        : (L.closest ? 2 : L.lists) * depth + 16;
    L.spill_cap = cap > L.slots ? cap - L.slots + 8 : 8;
    L.spill_words = (size_t)L.grid * L.block * L.spill_cap * (L.closest ? 2 : 1);
    *out = L;
    return SQ_OK;
}

template <bool CLOSEST, bool MULTI, int FMT, int WL>
static void launch_trace(hipStream_t stream, const TracePlan& L, const TraceArgs& a, bool narrow) {
    if (narrow)
        hipLaunchKernelGGL((k_trace<CLOSEST, MULTI, FMT, WL, true>), dim3(L.grid), dim3(L.block), L.lds_bytes, stream,
                           a);
    else
        hipLaunchKernelGGL((k_trace<CLOSEST, MULTI, FMT, WL, false>), dim3(L.grid), dim3(L.block), L.lds_bytes,
                           stream, a);
}

int trace_launch_plan(hipStream_t stream, const TracePlan& L, const TraceArgs& a, bool multi, bool narrow) {
    hipLaunchKernelGGL(k_trace_init, dim3(1), dim3(64), 0, stream, a.work, a.stats);
    if (multi) {
        if (L.closest && L.lists == 1) launch_trace<true, true, 1, 1>(stream, L, a, narrow);
        else if (L.closest) launch_trace<true, true, 1, kListLists>(stream, L, a, narrow);
        else if (L.lists == kWideLists) launch_trace<false, true, 1, kWideLists>(stream, L, a, narrow);
        else launch_trace<false, true, 1, kListLists>(stream, L, a, narrow);
    } else {
        if (L.closest && L.lists == 1) launch_trace<true, false, 1, 1>(stream, L, a, narrow);
        else if (L.closest) launch_trace<true, false, 1, kListLists>(stream, L, a, narrow);
        else if (L.lists == kWideLists) launch_trace<false, false, 1, kWideLists>(stream, L, a, narrow);
        else launch_trace<false, false, 1, kListLists>(stream, L, a, narrow);
    }
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

static int trace_launch_plank(hipStream_t stream, const TracePlan& L, const TraceArgs& a, int fmt) {
    hipLaunchKernelGGL(k_trace_init, dim3(1), dim3(64), 0, stream, a.work, a.stats);
    if (fmt == 4 && L.lists == 2) {
        if (L.closest) launch_trace<true, false, 4, 2>(stream, L, a, false);
        else launch_trace<false, false, 4, 2>(stream, L, a, false);
    } else if (fmt == 4) {
        if (L.closest) launch_trace<true, false, 4, 1>(stream, L, a, false);
        else launch_trace<false, false, 4, 1>(stream, L, a, false);
    } else {
        if (L.closest) launch_trace<true, false, 8, 1>(stream, L, a, false);
        else launch_trace<false, false, 8, 1>(stream, L, a, false);
    }
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

int trace_launch_plan2(hipStream_t stream, const TracePlan& L, const TraceArgs& a) {
    hipLaunchKernelGGL(k_trace_init, dim3(1), dim3(64), 0, stream, a.work, a.stats);
    // Node2 trees keep 64-bit addressing (their nodes come through load_node2)
    if (L.closest && L.lists == 1) launch_trace<true, false, 2, 1>(stream, L, a, false);
    else if (L.closest) launch_trace<true, false, 2, kListLists>(stream, L, a, false);
    else if (L.lists == kWideLists) launch_trace<false, false, 2, kWideLists>(stream, L, a, false);
    else launch_trace<false, false, 2, kListLists>(stream, L, a, false);
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

bool use_packets(const sq_trace_params* p) {
    const bool exhaustive = !p || p->semantics == SQ_TRACE_EXHAUSTIVE;
    const uint32_t k = p ? p->kernel : 0u;
    return exhaustive && k == 2u;
}

// Tree fields of the launch arguments (Node1 tree or Node2 tree).
// The tree a trace runs over: a Node1 tree, a Node2 tree or a Node<K> tree.
struct Target {
    const sq_bvh* b = nullptr;
    const sq_bvh2* t2 = nullptr;
    const sq_bvhk* tk = nullptr;
    int fmt() const { return b ? 1 : t2 ? 2 : (int)tk->k; }
};

static int target_args(const Target& t, const sq_trace_params* p, TraceArgs& a) {
    const sq_bvh* b = t.b;
    const sq_bvh2* t2 = t.t2;
    if (p && p->tri_test > 1) SQ_RET(SQ_ERR_INVALID, "trace: tri_test %u", p->tri_test);
    if (t.tk) {
        const sq_bvhk* k = t.tk;
        if (p && p->tri_test) SQ_RET(SQ_ERR_UNSUPPORTED, "trace: Node<K> trees use the reference triangle test");
        if (!k->scene) SQ_RET(SQ_ERR_INVALID, "trace: the tree has no scene");
        a.nodesk = k->d_nodes;
        a.tris = k->scene->d_tris;
        a.root = k->root.child;
        a.root_kind = k->root.kind;
        a.root_count = k->root.count;
        for (int j = 0; j < 3; j++) {
            a.rb[j] = k->root.min[j];
            a.rb[3 + j] = k->root.max[j];
        }
        a.node_count = k->nodes;
        a.tri_count = k->scene->n;
        return SQ_OK;
    }
    if (b) {
        if (!b->scene) SQ_RET(SQ_ERR_INVALID, "trace: the tree has no scene");
        if (b->capacity > kMaxTraceNodes)
            SQ_RET(SQ_ERR_UNSUPPORTED, "trace: %u nodes; stack entries hold node indices below 2^28", b->capacity);
        a.nodes = b->d_nodes;
        a.tris = b->scene->d_tris;
        a.root = b->root;
        a.node_count = b->capacity;
        a.tri_count = b->scene->n;
        if (p && p->tri_test == 1) {
            if (!b->d_plucker) SQ_RET(SQ_ERR_INVALID, "trace: tri_test 1 needs sq_bvh_plucker first");
            a.plk = b->d_plucker;
        }
        return SQ_OK;
    }
    if (p && p->tri_test) SQ_RET(SQ_ERR_UNSUPPORTED, "trace: Node2 trees use the reference triangle test");
    a.nodes2 = t2->d_nodes;
    a.perm = t2->d_perm;
    a.tris = t2->d_tris;
    a.root = t2->root.child;
    a.root_kind = t2->root.kind;
    a.root_count = t2->root.count;
    for (int k = 0; k < 3; k++) {
        a.rb[k] = t2->root.min[k];
        a.rb[3 + k] = t2->root.max[k];
    }
    a.node_count = t2->nodes;
    a.tri_count = t2->tris;
    return SQ_OK;
}

static uint32_t target_depth(const Target& t) {
    return t.b ? t.b->depth : t.t2 ? t.t2->levels + 2 : t.tk->levels + 2;
}

static int trace_launch(sq_ctx* c, uint32_t depth, int fmt, TraceArgs& a, const sq_trace_params* p,
                        const TraceScratch* ext = nullptr) {
    if (fmt == 1 && a.camera && use_packets(p)) {
        if (ext) SQ_RET(SQ_ERR_UNSUPPORTED, "trace: wave packets need the context workspace");
        SQ_TRY(c->counters.ensure(64));
        SQ_TRY(c->stats.ensure(64));
        a.work = c->counters.as<uint32_t>();
        a.stats = (p && p->count_work) ? c->stats.as<unsigned long long>() : nullptr;
        const uint32_t grid = trace_packet_grid(c, false);
        phase_begin(c, PH_TRACE);
        int rc = trace_launch_packets(c->stream, grid, a, false);
        phase_end(c, PH_TRACE);
        return rc;
    }
    TracePlan L;
    SQ_TRY(trace_plan(c, p, depth, a.camera != 0u, &L, fmt, a.camera == 0u));
    if (ext) {
        // Caller scratch: counters first, then per-lane spill areas. A
        // persistent launch stays correct with fewer workgroups (tiles are
        // dealt from one counter), so the grid shrinks to what the spill
        // area holds, down to one workgroup per CU.
        const size_t per_block = (size_t)L.block * L.spill_cap * (L.closest ? 2 : 1);
        const size_t avail = ext->words > kScratchSpillWord ? ext->words - kScratchSpillWord : 0;
        const size_t fit = avail / per_block;
        if (fit < (size_t)c->num_cus)
            SQ_RET(SQ_ERR_BUDGET,
                   "trace: a tree of depth %u needs %zu bytes of stack spill per workgroup; the %zu-byte scratch "
                   "holds %zu workgroups, fewer than one per CU (%d)",
                   depth, per_block * 4, ext->words * 4, fit, c->num_cus);
        if (L.grid > fit) L.grid = (uint32_t)fit;
        L.spill_words = (size_t)L.grid * per_block;
    }
    if (!a.camera) {
        // Tiles of 64 k rays, k = min(4, ceil(rays / (64 x waves))): a list
        // too short to give every wave a 256-ray tile is cut into at most one
        // tile per wave, each lane taking as many rays as the list allows.
        // Secondary rays on 10M clustered triangles: 100k exhaustive 301 ms
        // with 64-ray tiles against 583 ms with 256 (390 waves of 4,096
        // busy); 1M closest 793 ms with 256-ray tiles against 1,160 with 64.
        const uint64_t lanes = (uint64_t)L.grid * L.block;
        const uint64_t k = ((uint64_t)a.nrays + lanes - 1) / lanes;
        a.ray_tile = 64u * (uint32_t)(k < 1 ? 1 : (k > kRayTile / 64 ? kRayTile / 64 : k));
        a.num_tiles = (uint32_t)(((uint64_t)a.nrays + a.ray_tile - 1) / a.ray_tile);
    }
    if (ext) {
        a.work = ext->base;
        a.stats = (p && p->count_work) ? (unsigned long long*)(ext->base + kScratchStatsWord) : nullptr;
        a.spill = ext->base + kScratchSpillWord;
    } else {
        SQ_TRY(c->spill.ensure(L.spill_words * 4));
        SQ_TRY(c->counters.ensure(64));
        SQ_TRY(c->stats.ensure(64));
        a.work = c->counters.as<uint32_t>();
        a.stats = (p && p->count_work) ? c->stats.as<unsigned long long>() : nullptr;
        a.spill = c->spill.as<uint32_t>();
    }
    a.slots = L.slots;
    a.spill_cap = L.spill_cap;
    phase_begin(c, PH_TRACE);
    int rc = fmt >= 4   ? trace_launch_plank(c->stream, L, a, fmt)
             : fmt == 2 ? trace_launch_plan2(c->stream, L, a)
                        : trace_launch_plan(c->stream, L, a, false, !c->wide_fetch && narrow_ok(a.node_count, a.tri_count));
    phase_end(c, PH_TRACE);
    return rc;
}

// Waits for the trace and reads its fault bits (and work totals) back from
// a.work / a.stats through the context's pinned words.
static int trace_finish(sq_ctx* c, const sq_trace_params* p, const TraceArgs& a) {
    if (!c->pinned || c->pinned_bytes < 64) SQ_RET(SQ_ERR_HIP, "trace: the context has no pinned status words");
    uint32_t* h = (uint32_t*)c->pinned;
    SQ_CHECK_HIP(hipMemcpyAsync(h, a.work, 16, hipMemcpyDeviceToHost, c->stream));
    if (p && p->count_work && a.stats)
        SQ_CHECK_HIP(hipMemcpyAsync(h + 8, a.stats, 16, hipMemcpyDeviceToHost, c->stream));
    SQ_TRY(phases_collect(c));
    if (p && p->count_work && a.stats) {
        unsigned long long* s = (unsigned long long*)(h + 8);
        c->last.node_tests = s[0];
        c->last.tri_tests = s[1];
    }
    if (h[1]) return trace_fault(h[1], "trace");
    return SQ_OK;
}

static int rays_device(sq_ctx* c, const Target& t, const sq_ray* rays, uint32_t n, const sq_trace_params* p,
                       const sq_trace_out* out, const TraceScratch* ext) {
    SQ_CHECK_HIP(hipSetDevice(c->device));
    phases_reset(c);
    TraceArgs a{};
    SQ_TRY(target_args(t, p, a));
    a.camera = 0;
    a.nrays = n;
    a.rays = rays;
    a.hit_id = out->hit_id; a.hit_t = out->hit_t; a.counts = out->counts;
    SQ_TRY(trace_launch(c, target_depth(t), t.fmt(), a, p, ext));
    return trace_finish(c, p, a);
}

int trace_rays_device(sq_ctx* c, const sq_bvh* b, const sq_bvh2* t2, const sq_ray* rays, uint32_t n,
                      const sq_trace_params* p, const sq_trace_out* out, const TraceScratch* ext) {
    Target t;
    t.b = b;
    t.t2 = t2;
    return rays_device(c, t, rays, n, p, out, ext);
}

}  // namespace sq

using namespace sq;

extern "C" {

static int camera_impl(sq_ctx* c, const Target& t, const sq_camera* cam, const sq_trace_params* p,
                       const sq_trace_out* out, int out_on_device) {
    const sq_bvh* b = t.b;
    const bool other = t.t2 || t.tk;  // a Node2 or Node<K> tree
    if (!c || !(b || other) || !cam || !out) SQ_RET(SQ_ERR_INVALID, "sq_trace_camera: NULL argument");
    if (cam->width == 0 || cam->height == 0 || cam->spp == 0) SQ_RET(SQ_ERR_INVALID, "sq_trace_camera: empty image");
    if (p && p->heatmap > SQ_HEATMAP_BUNDLE) SQ_RET(SQ_ERR_INVALID, "sq_trace_camera: heatmap mode %u", p->heatmap);
    if (p && p->kernel > 2) SQ_RET(SQ_ERR_INVALID, "sq_trace_camera: kernel %u", p->kernel);
    const bool heat = p && p->heatmap != SQ_HEATMAP_RAY;
    if (heat && (cam->spp != 1 || out->hit_id || out->hit_t))
        SQ_RET(SQ_ERR_UNSUPPORTED, "packet/bundle heatmaps take one sample per pixel and produce counts/rgb only");
    if (other && (heat || (p && p->kernel == SQ_KERNEL_PACKETS)))
        SQ_RET(SQ_ERR_UNSUPPORTED, "Node2 and Node<K> trees are traced by the per-ray kernel with ray heatmaps only");
    if (p && p->tri_test && (heat || p->kernel == SQ_KERNEL_PACKETS))
        SQ_RET(SQ_ERR_UNSUPPORTED, "precomputed triangles are tested by the per-ray kernel only");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    phases_reset(c);
    const size_t npix = (size_t)cam->width * cam->height, nray = npix * cam->spp;
    TraceArgs a{};
    SQ_TRY(target_args(t, p, a));
    a.camera = 1;
    a.ox = cam->origin.x; a.oy = cam->origin.y; a.oz = cam->origin.z;
    a.tanh = cam->tan_half_fov;
    a.w = cam->width; a.h = cam->height; a.spp = cam->spp; a.jitter = cam->jitter; a.seed = cam->seed;
    a.tmin = cam->tmin; a.tmax = cam->tmax;
    // 4x4-pixel tiles for the per-ray kernel's exhaustive traversal: four
    // times as many tiles as 8x8 ones, so the last tiles of a frame are
    // short. 1080p x 4 over the 1M soup: 7.08 ms, against 9.55 ms with 8x8
    // tiles, whose ~2,500 costly tiles fell on 2,048 waves (half the waves
    // idle for the last fifth of the launch).
    const bool exhaustive = !p || p->semantics == SQ_TRACE_EXHAUSTIVE;
    if (camera_tiling(cam->width, cam->height, cam->spp, exhaustive && (other || !use_packets(p)), a) >= (1ull << 31))
        SQ_RET(SQ_ERR_INVALID, "sq_trace_camera: %ux%u image has too many tiles", cam->width, cam->height);
    if (out_on_device) {
        a.hit_id = out->hit_id; a.hit_t = out->hit_t; a.counts = out->counts; a.rgb = out->rgb;
        if (heat) SQ_TRY(heat_launch(c, b, a, p));
        else SQ_TRY(trace_launch(c, target_depth(t), t.fmt(), a, p));
        return trace_finish(c, heat ? nullptr : p, a);
    }
    // Host outputs: stage through workspace buffers.
    size_t off_id = 0, off_t = off_id + (out->hit_id ? nray * 4 : 0);
    size_t off_c = off_t + (out->hit_t ? nray * 4 : 0), off_rgb = off_c + (out->counts ? npix * 4 : 0);
    size_t total = off_rgb + (out->rgb ? npix * 3 : 0);
    SQ_TRY(c->scratch.ensure(total + 256));
    char* base = c->scratch.as<char>();
    a.hit_id = out->hit_id ? (uint32_t*)(base + off_id) : nullptr;
    a.hit_t = out->hit_t ? (float*)(base + off_t) : nullptr;
    a.counts = out->counts ? (uint32_t*)(base + off_c) : nullptr;
    a.rgb = out->rgb ? (uint8_t*)(base + off_rgb) : nullptr;
    if (heat) SQ_TRY(heat_launch(c, b, a, p));
    else SQ_TRY(trace_launch(c, target_depth(t), t.fmt(), a, p));
    if (out->hit_id) SQ_CHECK_HIP(hipMemcpyAsync(out->hit_id, a.hit_id, nray * 4, hipMemcpyDeviceToHost, c->stream));
    if (out->hit_t) SQ_CHECK_HIP(hipMemcpyAsync(out->hit_t, a.hit_t, nray * 4, hipMemcpyDeviceToHost, c->stream));
    if (out->counts) SQ_CHECK_HIP(hipMemcpyAsync(out->counts, a.counts, npix * 4, hipMemcpyDeviceToHost, c->stream));
    if (out->rgb) SQ_CHECK_HIP(hipMemcpyAsync(out->rgb, a.rgb, npix * 3, hipMemcpyDeviceToHost, c->stream));
    return trace_finish(c, heat ? nullptr : p, a);
}

static int rays_impl(sq_ctx* c, const Target& t, const sq_ray* rays, uint32_t n, const sq_trace_params* p,
                     const sq_trace_out* out, int io_on_device) {
    if (!c || !(t.b || t.t2 || t.tk) || !rays || !out) SQ_RET(SQ_ERR_INVALID, "sq_trace_rays: NULL argument");
    if (p && p->heatmap != SQ_HEATMAP_RAY) SQ_RET(SQ_ERR_UNSUPPORTED, "sq_trace_rays: heatmap mode %u", p->heatmap);
    if (n == 0) return SQ_OK;
    if (io_on_device) return rays_device(c, t, rays, n, p, out, nullptr);
    SQ_CHECK_HIP(hipSetDevice(c->device));
    phases_reset(c);
    TraceArgs a{};
    SQ_TRY(target_args(t, p, a));
    a.camera = 0;
    a.nrays = n;
    const size_t nr = n;
    size_t off_r = 0, off_id = nr * sizeof(sq_ray), off_t = off_id + nr * 4, off_c = off_t + nr * 4;
    SQ_TRY(c->scratch.ensure(off_c + nr * 4 + 256));
    char* base = c->scratch.as<char>();
    SQ_CHECK_HIP(hipMemcpyAsync(base + off_r, rays, nr * sizeof(sq_ray), hipMemcpyHostToDevice, c->stream));
    a.rays = (const sq_ray*)(base + off_r);
    a.hit_id = out->hit_id ? (uint32_t*)(base + off_id) : nullptr;
    a.hit_t = out->hit_t ? (float*)(base + off_t) : nullptr;
    a.counts = out->counts ? (uint32_t*)(base + off_c) : nullptr;
    SQ_TRY(trace_launch(c, target_depth(t), t.fmt(), a, p));
    if (out->hit_id) SQ_CHECK_HIP(hipMemcpyAsync(out->hit_id, a.hit_id, nr * 4, hipMemcpyDeviceToHost, c->stream));
    if (out->hit_t) SQ_CHECK_HIP(hipMemcpyAsync(out->hit_t, a.hit_t, nr * 4, hipMemcpyDeviceToHost, c->stream));
    if (out->counts) SQ_CHECK_HIP(hipMemcpyAsync(out->counts, a.counts, nr * 4, hipMemcpyDeviceToHost, c->stream));
    return trace_finish(c, p, a);
}

int sq_trace_camera(sq_ctx* c, const sq_bvh* b, const sq_camera* cam, const sq_trace_params* p,
                    const sq_trace_out* out, int out_on_device) {
    if (!b) SQ_RET(SQ_ERR_INVALID, "sq_trace_camera: NULL tree");
    Target t;
    t.b = b;
    return camera_impl(c, t, cam, p, out, out_on_device);
}

int sq_trace_rays(sq_ctx* c, const sq_bvh* b, const sq_ray* rays, uint32_t n, const sq_trace_params* p,
                  const sq_trace_out* out, int io_on_device) {
    if (!b) SQ_RET(SQ_ERR_INVALID, "sq_trace_rays: NULL tree");
    Target t;
    t.b = b;
    return rays_impl(c, t, rays, n, p, out, io_on_device);
}

int sq_trace2_camera(sq_ctx* c, const sq_bvh2* t, const sq_camera* cam, const sq_trace_params* p,
                     const sq_trace_out* out, int out_on_device) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_trace2_camera: NULL tree");
    Target g;
    g.t2 = t;
    return camera_impl(c, g, cam, p, out, out_on_device);
}

int sq_trace2_rays(sq_ctx* c, const sq_bvh2* t, const sq_ray* rays, uint32_t n, const sq_trace_params* p,
                   const sq_trace_out* out, int io_on_device) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_trace2_rays: NULL tree");
    Target g;
    g.t2 = t;
    return rays_impl(c, g, rays, n, p, out, io_on_device);
}

int sq_tracek_camera(sq_ctx* c, const sq_bvhk* t, const sq_camera* cam, const sq_trace_params* p,
                     const sq_trace_out* out, int out_on_device) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_tracek_camera: NULL tree");
    Target g;
    g.tk = t;
    return camera_impl(c, g, cam, p, out, out_on_device);
}

int sq_tracek_rays(sq_ctx* c, const sq_bvhk* t, const sq_ray* rays, uint32_t n, const sq_trace_params* p,
                   const sq_trace_out* out, int io_on_device) {
    if (!t) SQ_RET(SQ_ERR_INVALID, "sq_tracek_rays: NULL tree");
    Target g;
    g.tk = t;
    return rays_impl(c, g, rays, n, p, out, io_on_device);
}

}  // extern "C"
