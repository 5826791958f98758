// plucker.cpp — CPU restatement of the precomputed-triangle intersection
// (PluckerTriangle) used as an alternative to the reference's Möller–Trumbore
// test.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP precompute and traversal.
// Built with -ffp-contract=off; no fused multiply-adds anywhere here.
//
// The reference declares PluckerTriangle (aabb.h:182-195: nu, nv, np, pu, pv,
// ci, e0u, e0v, e1u, e1v, two pads; 48 bytes) and leaves its conversion
// commented out (aabb.h:208-241) with no intersection routine. This is the
// projection method that layout belongs to (I. Wald, "Realtime Ray Tracing and
// Interactive Global Illumination", 2004, section 7.1), written out:
//   e1 = v1 - v0, e2 = v2 - v0, n = e1 x e2; w = the axis of the largest |n|
//   (x on ties with y or z, then y), (u, v) = the other two in order;
//   nu = n[u] / n[w], nv = n[v] / n[w] (the plane is P[w] + nu P[u] + nv P[v]
//   = np with np = nu v0[u] + nv v0[v] + v0[w]); pu, pv = v0[u], v0[v];
//   det = n[w], negated for w = y (where (u, v, w) is not cyclic);
//   e0u, e0v = e1[u] / det, e1[v] / det; e1u, e1v = e2[u] / det, e2[v] / det;
//   ci = w, or 3 for a triangle whose n[w] is 0 or whose quotients are not
//   finite (it is never hit). The commented-out reference conversion divides
//   nu, nv by |n[w]|, which breaks the plane equation when n[w] < 0; the
//   signed n[w] is used here.
// Test: den = (d[w] + nu d[u]) + nv d[v]; t = (np - ((o[w] + nu o[u]) +
// nv o[v])) / den; hu = (o[u] + t d[u]) - pu, hv = (o[v] + t d[v]) - pv;
// beta = hu e1v - hv e1u, gamma = hv e0u - hu e0v; hit when beta >= 0,
// gamma >= 0, beta + gamma <= 1 and t > 1e-9 (the reference's epsilon).
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

extern "C" {

void sqo_plucker_precompute(const sqo_tri* tris, uint32_t n, sqo_plucker* out) {
    for (uint32_t i = 0; i < n; i++) {
        const float* a = &tris[i].v0.x;
        const float* b = &tris[i].v1.x;
        const float* c = &tris[i].v2.x;
        float e1[3], e2[3], nn[3];
        for (int k = 0; k < 3; k++) { e1[k] = b[k] - a[k]; e2[k] = c[k] - a[k]; }
        nn[0] = e1[1] * e2[2] - e1[2] * e2[1];
        nn[1] = e1[2] * e2[0] - e1[0] * e2[2];
        nn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        const float ax = fabsf(nn[0]), ay = fabsf(nn[1]), az = fabsf(nn[2]);
        const int w = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
        const int u = w == 0 ? 1 : 0, v = w == 2 ? 1 : 2;
        sqo_plucker p;
        std::memset(&p, 0, sizeof p);
        const float det = w == 1 ? -nn[w] : nn[w];
        p.nu = nn[u] / nn[w];
        p.nv = nn[v] / nn[w];
        p.np = (p.nu * a[u] + p.nv * a[v]) + a[w];
        p.pu = a[u];
        p.pv = a[v];
        p.e0u = e1[u] / det;
        p.e0v = e1[v] / det;
        p.e1u = e2[u] / det;
        p.e1v = e2[v] / det;
        const bool ok = nn[w] != 0.0f && std::isfinite(p.nu) && std::isfinite(p.nv) && std::isfinite(p.np) &&
                        std::isfinite(p.e0u) && std::isfinite(p.e0v) && std::isfinite(p.e1u) && std::isfinite(p.e1v);
        p.ci = ok ? w : 3;
        out[i] = p;
    }
}

int sqo_plucker_test(const sqo_plucker* p, const sqo_ray* r, float* t_out) {
    *t_out = INFINITY;
    if (p->ci > 2) return 0;
    const int w = p->ci, u = w == 0 ? 1 : 0, v = w == 2 ? 1 : 2;
    const float o[3] = {r->ox, r->oy, r->oz}, d[3] = {r->dx, r->dy, r->dz};
    const float den = (d[w] + p->nu * d[u]) + p->nv * d[v];
    if (den == 0.0f) return 0;
    const float t = (p->np - ((o[w] + p->nu * o[u]) + p->nv * o[v])) / den;
    const float hu = (o[u] + t * d[u]) - p->pu, hv = (o[v] + t * d[v]) - p->pv;
    const float beta = hu * p->e1v - hv * p->e1u, gamma = hv * p->e0u - hu * p->e0v;
    *t_out = t;
    return (beta >= 0.0f) & (gamma >= 0.0f) & ((beta + gamma) <= 1.0f) & (t > 0.000000001f);
}

// Exhaustive traversal (as sqo_trace_rays) with the precomputed test.
void sqo_trace_rays_plucker(const sqo_node* nodes, uint32_t root, const sqo_plucker* pl, const sqo_ray* rays,
                            uint32_t nrays, uint32_t* counts, uint32_t* hit_id, float* hit_t) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)nrays; i++) {
        const sqo_ray& r = rays[i];
        uint32_t cnt = 0, best_id = 0xFFFFFFFFu;
        float best_t = INFINITY;
        std::vector<uint32_t> st;
        st.push_back(root);
        st.push_back(1);
        while (!st.empty()) {
            const uint32_t c = st.back(); st.pop_back();
            const uint32_t f = st.back(); st.pop_back();
            for (uint32_t j = 0; j < c; j++) {
                const sqo_node& nd = nodes[f + j];
                cnt++;
                if (!sqo_slab(&nd, &r)) continue;
                if (nd.flags & 1) {
                    for (uint32_t k = 0; k < nd.count; k++) {
                        const uint32_t tid = nd.child + k;
                        float t;
                        if (sqo_plucker_test(&pl[tid], &r, &t) && t > r.tmin && t < r.tmax &&
                            (t < best_t || (t == best_t && tid < best_id))) {
                            best_t = t;
                            best_id = tid;
                        }
                    }
                } else {
                    st.push_back(nd.child);
                    st.push_back(nd.count);
                }
            }
        }
        if (counts) counts[i] = cnt;
        if (hit_id) hit_id[i] = best_id;
        if (hit_t) hit_t[i] = best_t;
    }
}

}  // extern "C"
