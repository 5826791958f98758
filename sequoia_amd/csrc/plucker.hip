// plucker.hip — precomputed triangles (the reference's PluckerTriangle
// layout, aabb.h:182-195; conversion commented out at aabb.h:208-241) for the
// projection-method intersection test (Wald 2004, 7.1):
//   e1 = v1 - v0, e2 = v2 - v0, n = e1 x e2, w = axis of the largest |n| (x on
//   ties with y or z, then y), (u, v) = the other two in order;
//   nu = n[u] / n[w], nv = n[v] / n[w], np = (nu v0[u] + nv v0[v]) + v0[w],
//   pu = v0[u], pv = v0[v]; det = n[w], negated for w = y;
//   e0u, e0v = e1[u] / det, e1[v] / det; e1u, e1v = e2[u] / det, e2[v] / det;
//   ci = w, or 3 when n[w] is 0 or a quotient is not finite.
// The reference's commented-out conversion divides by |n[w]|, which breaks
// the plane equation for n[w] < 0; the signed n[w] is used. Plain operations
// only (no fused multiply-adds), so the CPU restatement gives the same bits.
#include "trace_common.h"

namespace sq {

__global__ __launch_bounds__(256) void k_plucker(const sq_triangle* __restrict__ tris, uint32_t n,
                                                 sq_plucker* __restrict__ out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float v[9];
        sqd::load_tri(tris, i, v);
        const float e1[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
        const float e2[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
        const float nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                             e1[0] * e2[1] - e1[1] * e2[0]};
        const float ax = fabsf(nn[0]), ay = fabsf(nn[1]), az = fabsf(nn[2]);
        const int w = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
        const int u = w == 0 ? 1 : 0, vv = w == 2 ? 1 : 2;
        const float nw = w == 0 ? nn[0] : (w == 1 ? nn[1] : nn[2]);
        const float nuu = u == 0 ? nn[0] : nn[1], nvv = vv == 1 ? nn[1] : nn[2];
        const float au = u == 0 ? v[0] : v[1], av = vv == 1 ? v[1] : v[2];
        const float aw = w == 0 ? v[0] : (w == 1 ? v[1] : v[2]);
        const float e1u_ = u == 0 ? e1[0] : e1[1], e1v_ = vv == 1 ? e1[1] : e1[2];
        const float e2u_ = u == 0 ? e2[0] : e2[1], e2v_ = vv == 1 ? e2[1] : e2[2];
        const float det = w == 1 ? -nw : nw;
        sq_plucker p;
        p.nu = nuu / nw;
        p.nv = nvv / nw;
        p.np = (p.nu * au + p.nv * av) + aw;
        p.pu = au;
        p.pv = av;
        p.e0u = e1u_ / det;
        p.e0v = e1v_ / det;
        p.e1u = e2u_ / det;
        p.e1v = e2v_ / det;
        const bool ok = nw != 0.0f && isfinite(p.nu) && isfinite(p.nv) && isfinite(p.np) && isfinite(p.e0u) &&
                        isfinite(p.e0v) && isfinite(p.e1u) && isfinite(p.e1v);
        p.ci = ok ? w : 3;
        p.pad0 = 0;
        p.pad1 = 0;
        out[i] = p;
    }
}

}  // namespace sq

using namespace sq;

extern "C" {

int sq_bvh_plucker(sq_ctx* c, sq_bvh* b, const sq_plucker** out) {
    if (!c || !b) SQ_RET(SQ_ERR_INVALID, "sq_bvh_plucker: NULL argument");
    if (b->view || !b->scene) SQ_RET(SQ_ERR_UNSUPPORTED, "sq_bvh_plucker: trees of a batch are not supported");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    const uint32_t n = b->scene->n;
    if (!b->d_plucker) {
        SQ_CHECK_HIP(hipMalloc(&b->d_plucker, (size_t)(n ? n : 1) * sizeof(sq_plucker)));
        if (n) {
            uint32_t blocks = (n + 255) / 256, cap = (uint32_t)c->num_cus * 8;
            if (blocks > cap) blocks = cap;
            hipLaunchKernelGGL(k_plucker, dim3(blocks), dim3(256), 0, c->stream, b->scene->d_tris, n, b->d_plucker);
            SQ_CHECK_HIP(hipGetLastError());
        }
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    }
    if (out) *out = b->d_plucker;
    return SQ_OK;
}

}  // extern "C"
