// heatmap.hip — packet and bundle heatmap tracers.
//
// trace_packet<K> (tracer.cpp:88-125, 232-281): a packet of K rays over a
// floor(sqrt K) x K/floor(sqrt K) pixel tile descends into a node's children
// when any of its rays hits the node; one count per node test of the packet;
// pixel value = uint8(min(1, (count/K)/200) * 255).
//
// trace_bundle<K> (tracer.h:39-80, tracer.cpp:128-156, 292-362): K rays over a
// sqrt K x sqrt K tile (Morton pixel order) form a bundle tested with the
// interval directions [dir_min, dir_max] and the mid direction; the bundle
// descends while the mid ray hits; where only "any" holds, each packet of 8
// rays runs the packet traversal from that node. Pixel value =
// uint8(min(1, (bundle_tests/K + packet_tests/8)/200) * 255).
//
// Mapping: a packet of K rays is K adjacent lanes of a wave and decides with
// a masked ballot; a bundle is 16 or 64 adjacent lanes, its interval test is
// evaluated identically by every lane, and the fallback packets are 8-lane
// groups that traverse concurrently. Counts do not depend on visiting order,
// so the stack order of the reference's recursion is not reproduced.
// Triangles are not intersected: they do not affect these heatmaps.
#include "trace_common.h"

namespace sq {

__constant__ uint32_t kMortonLut[64] = {  // tracer.cpp:18-23
    0x00, 0x01, 0x04, 0x05, 0x02, 0x03, 0x06, 0x07, 0x08, 0x09, 0x0C, 0x0D, 0x0A, 0x0B, 0x0E, 0x0F,
    0x10, 0x11, 0x14, 0x15, 0x12, 0x13, 0x16, 0x17, 0x18, 0x19, 0x1C, 0x1D, 0x1A, 0x1B, 0x1E, 0x1F,
    0x20, 0x21, 0x24, 0x25, 0x22, 0x23, 0x26, 0x27, 0x28, 0x29, 0x2C, 0x2D, 0x2A, 0x2B, 0x2E, 0x2F,
    0x30, 0x31, 0x34, 0x35, 0x32, 0x33, 0x36, 0x37, 0x38, 0x39, 0x3C, 0x3D, 0x3A, 0x3B, 0x3E, 0x3F};

struct HeatArgs {
    TraceArgs t;
    uint32_t K;          // packet or bundle width
    uint32_t kx, ky;     // tile shape of one unit (packet or bundle)
    uint32_t units_x;    // units per image row
    uint32_t units;      // total units
    uint32_t bundle;     // 1: bundle mode
};

__device__ __forceinline__ uint64_t group_bits(uint64_t bal, uint32_t group, uint32_t K) {
    if (K >= 64) return bal;
    return (bal >> (group * K)) & ((1ull << K) - 1ull);
}

__device__ __forceinline__ bool decode_entry(const TraceArgs& a, uint32_t e, uint32_t& first, uint32_t& cnt) {
    if ((e & 15u) != 15u) {
        first = e >> 4;
        cnt = (e & 15u) + 1;
        return true;
    }
    if ((e >> 4) >= a.node_count) return false;
    NodeV p = load_node(a.nodes, e >> 4);
    first = p.child;
    cnt = p.count;
    return true;
}

// traverse_packet_stack for the K-lane group of this lane, over the entries
// above base_sp. Returns the packet's node tests.
__device__ uint32_t packet_walk(const TraceArgs& a, const RayF& r, float ix, float iy, float iz, bool fast,
                                Stack& st, uint32_t base_sp, uint32_t group, uint32_t K) {
    uint32_t tests = 0;
    while (st.sp > base_sp) {
        float te;
        uint32_t e = pop<false>(st, te), first, cnt;
        if (!decode_entry(a, e, first, cnt)) { st.fault |= kFaultNode; return tests; }
        for (uint32_t i = 0; i < cnt; i++) {
            const uint32_t ni = first + i;
            if (ni >= a.node_count) { st.fault |= kFaultNode; return tests; }
            NodeV n = load_node(a.nodes, ni);
            float f;
            bool h = slab(n, r, ix, iy, iz, fast, f);
            uint64_t g = group_bits(__ballot(h), group, K);
            tests++;
            if (g != 0 && !n.leaf) {
                push<false>(st, encode_list(ni, n), 0.f);
                if (st.fault) return tests;
            }
        }
    }
    return tests;
}

struct BundleBox {
    float imn[3], imx[3], imd[3], pos[3];
};

// intersect_bundle_aabb (tracer.h:39-80): returns any_hit, sets mid_hit.
__device__ __forceinline__ bool bundle_test(const BundleBox& b, const NodeV& n, bool& mid) {
    float tmin_mid[3], tmin_any[3], tmax_mid[3], tmax_any[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float lo = n.mn[k] - b.pos[k], hi = n.mx[k] - b.pos[k];
        float t1a = lo * b.imn[k], t1b = lo * b.imx[k], t2a = hi * b.imn[k], t2b = hi * b.imx[k];
        float t1 = lo * b.imd[k], t2 = hi * b.imd[k];
        float t1min = sqd::simd_min(t1a, t1b), t1max = sqd::simd_max(t1a, t1b);
        float t2min = sqd::simd_min(t2a, t2b), t2max = sqd::simd_max(t2a, t2b);
        tmin_mid[k] = sqd::simd_min(t1, t2);
        tmin_any[k] = sqd::simd_min(t1min, t2min);
        tmax_mid[k] = sqd::simd_max(t1, t2);
        tmax_any[k] = sqd::simd_max(t1max, t2max);
    }
    float front_mid = sqd::std_max(sqd::std_max(tmin_mid[0], tmin_mid[1]), tmin_mid[2]);
    float front_any = sqd::std_max(sqd::std_max(tmin_any[0], tmin_any[1]), tmin_any[2]);
    float back_mid = sqd::std_min(sqd::std_min(tmax_mid[0], tmax_mid[1]), tmax_mid[2]);
    float back_any = sqd::std_min(sqd::std_min(tmax_any[0], tmax_any[1]), tmax_any[2]);
    mid = back_mid >= front_mid;
    return back_any >= front_any;
}

__global__ void k_zero_work(uint32_t* work) {
    if (threadIdx.x < 4) work[threadIdx.x] = 0;
}

__global__ __launch_bounds__(128) void k_heat(HeatArgs h) {
    extern __shared__ uint32_t s_dyn[];
    const TraceArgs& a = h.t;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t K = h.K, group = lane / K, gl = lane % K, groups = 64 / K;
    Stack st;
    st.stride = blockDim.x;
    st.S = a.slots;
    st.mask = a.slots - 1;
    st.lds = s_dyn + threadIdx.x;
    st.lds_t = nullptr;
    const size_t gtid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    st.cap = a.spill_cap;
    st.spill = a.spill + gtid * (size_t)a.spill_cap;
    uint32_t overflow = 0;
    while (true) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(&a.work[0], 1u);
        tile = __shfl(tile, 0, 64);
        if (tile * groups >= h.units) break;
        const uint32_t unit = tile * groups + group;
        const bool valid = unit < h.units;
        const uint32_t ux = valid ? unit % h.units_x : 0, uy = valid ? unit / h.units_x : 0;
        uint32_t mx, my;
        if (h.bundle) {
            mx = kMortonLut[gl] % h.kx;
            my = kMortonLut[gl] / h.kx;
        } else {
            mx = gl % h.kx;
            my = gl / h.kx;
        }
        const uint32_t px = ux * h.kx + mx, py = uy * h.ky + my;
        RayF r = camera_ray(a, px, py, 0);
        const float ix = 1.0f / r.dx, iy = 1.0f / r.dy, iz = 1.0f / r.dz;
        const bool fast = isfinite(ix) && isfinite(iy) && isfinite(iz) && isfinite(r.ox) && isfinite(r.oy) &&
                          isfinite(r.oz);
        st.sp = 0;
        st.fault = 0;
        float value = 0.f;
        uint32_t count = 0;
        if (valid) {
            if (!h.bundle) {
                push<false>(st, a.root < (1u << 28) ? (a.root << 4) : 15u, 0.f);
                count = packet_walk(a, r, ix, iy, iz, fast, st, 0, group, K);
                value = ((float)count / (float)K) / 200.0f;
            } else {
                // bundle interval: reference min/max over the K directions
                float dmn[3] = {r.dx, r.dy, r.dz}, dmx[3] = {r.dx, r.dy, r.dz};
                for (uint32_t off = 1; off < K; off <<= 1) {
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        float a1 = __shfl_xor(dmn[k], (int)off, 64), b1 = __shfl_xor(dmx[k], (int)off, 64);
                        dmn[k] = sqd::simd_min(dmn[k], a1);
                        dmx[k] = sqd::simd_max(dmx[k], b1);
                    }
                }
                BundleBox bb;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    bb.imn[k] = 1.0f / dmn[k];
                    bb.imx[k] = 1.0f / dmx[k];
                    bb.imd[k] = 1.0f / ((dmn[k] + dmx[k]) * 0.5f);
                }
                bb.pos[0] = r.ox; bb.pos[1] = r.oy; bb.pos[2] = r.oz;
                uint32_t btests = 0, ptests = 0;
                const uint32_t pgroup = lane / 8;
                push<false>(st, a.root < (1u << 28) ? (a.root << 4) : 15u, 0.f);
                while (st.sp > 0 && !st.fault) {
                    float te;
                    uint32_t e = pop<false>(st, te), first, cnt;
                    if (!decode_entry(a, e, first, cnt)) { st.fault |= kFaultNode; break; }
                    for (uint32_t i = 0; i < cnt && !st.fault; i++) {
                        const uint32_t ni = first + i;
                        if (ni >= a.node_count) { st.fault |= kFaultNode; break; }
                        NodeV n = load_node(a.nodes, ni);
                        bool mid;
                        bool any = bundle_test(bb, n, mid);
                        btests++;
                        if (mid) {
                            if (!n.leaf) push<false>(st, encode_list(ni, n), 0.f);
                        } else if (any) {
                            const uint32_t sp0 = st.sp;
                            push<false>(st, ni << 4, 0.f);
                            ptests += packet_walk(a, r, ix, iy, iz, fast, st, sp0, pgroup, 8);
                        }
                    }
                }
                count = btests;
                value = ((float)btests / (float)K + (float)ptests / 8.0f) / 200.0f;
            }
            overflow |= st.fault;
            if (px < a.w && py < a.h) {
                const size_t p = (size_t)py * a.w + px;
                float v = (value < 1.0f) ? value : 1.0f;
                uint8_t byte = (uint8_t)(v * 255.0f);
                if (a.rgb) { a.rgb[3 * p] = byte; a.rgb[3 * p + 1] = byte; a.rgb[3 * p + 2] = byte; }
                if (a.counts) a.counts[p] = count;
            }
        }
    }
    if (overflow) atomicOr(&a.work[1], overflow);
}

int heat_launch(sq_ctx* c, const sq_bvh* b, TraceArgs& t, const sq_trace_params* p) {
    const uint32_t K = p->width;
    const bool bundle = p->heatmap == SQ_HEATMAP_BUNDLE;
    if (bundle ? !(K == 16 || K == 64) : !(K == 1 || K == 2 || K == 4 || K == 8))
        SQ_RET(SQ_ERR_UNSUPPORTED, "heatmap width %u (packet: 1,2,4,8; bundle: 16,64)", K);
    if (b->root >= (1u << 28)) SQ_RET(SQ_ERR_UNSUPPORTED, "heatmap modes need fewer than 2^28 nodes");
    HeatArgs h{};
    h.t = t;
    h.K = K;
    h.bundle = bundle ? 1u : 0u;
    if (bundle) {
        h.kx = K == 64 ? 8 : 4;
        h.ky = K / h.kx;
    } else {
        uint32_t kx = 1;
        while ((kx + 1) * (kx + 1) <= K) kx++;
        h.kx = kx;
        h.ky = K / kx;
    }
    h.units_x = (t.w + h.kx - 1) / h.kx;
    h.units = h.units_x * ((t.h + h.ky - 1) / h.ky);
    uint32_t slots = p->stack_slots ? p->stack_slots : 16;
    uint32_t pow2 = 1;
    while (pow2 < slots) pow2 <<= 1;
    slots = pow2 > 64 ? 64 : pow2;
    const uint32_t block = 128;
    int per_cu = 0;
    const size_t lds = (size_t)slots * block * 4;
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_heat, block, lds);
    if (per_cu < 1) per_cu = 8;
    const uint32_t grid = (uint32_t)c->num_cus * (uint32_t)per_cu;
    const uint32_t depth = b->depth ? b->depth : 64;
    uint32_t cap = 4 * depth + 32;
    const size_t words = (size_t)grid * block * cap;
    SQ_TRY(c->spill.ensure(words * 4));
    SQ_TRY(c->counters.ensure(64));
    h.t.slots = slots;
    h.t.spill = c->spill.as<uint32_t>();
    h.t.spill_cap = cap;
    h.t.work = c->counters.as<uint32_t>();
    t.work = h.t.work;  // the caller reads the fault bits back from here
    h.t.stats = nullptr;
    h.t.hit_id = nullptr;
    h.t.hit_t = nullptr;
    hipLaunchKernelGGL(k_zero_work, dim3(1), dim3(64), 0, c->stream, h.t.work);
    phase_begin(c, PH_TRACE);
    hipLaunchKernelGGL(k_heat, dim3(grid), dim3(block), lds, c->stream, h);
    phase_end(c, PH_TRACE);
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

}  // namespace sq
