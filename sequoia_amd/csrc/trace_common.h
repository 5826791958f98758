// trace_common.h — device code shared by the traversal kernels: ray
// generation, the reference's slab and triangle tests, node decoding and
// the per-lane LDS stack with HBM spill.
#pragma once
#include "internal.h"

#include <type_traits>

namespace sq {

struct TraceArgs {
    const sq_node* nodes;
    const sq_triangle* tris;
    uint32_t root;
    uint32_t node_count;  // reads at or beyond these bounds are skipped
    uint32_t tri_count;
    uint32_t camera;  // 1: camera rays, 0: ray list
    float ox, oy, oz, tanh;
    uint32_t w, h, spp, jitter, seed;
    float tmin, tmax;
    const sq_ray* rays;
    uint32_t nrays;
    uint32_t* hit_id;
    float* hit_t;
    uint32_t* counts;
    uint8_t* rgb;
    uint32_t slots;       // LDS entries per lane (power of two)
    uint32_t* spill;      // per-thread spill areas
    uint32_t spill_cap;   // entries per thread in the spill area
    uint32_t* work;       // [0] tile counter, [1] fault bits (kFault*)
    unsigned long long* stats;  // [0] node tests, [1] triangle tests
    uint32_t num_tiles;
    uint32_t tiles_x;
    uint32_t tile_shift;  // camera tiles of (1 << tile_shift)^2 pixels, all samples; 0 = 8x8
    uint32_t ray_tile;    // rays per ray-list tile; 0 = 256
    // multi-job launches: job table and its tile prefix (njobs + 1 entries)
    const TraceArgs* jobs;
    const uint32_t* job_tiles;
    uint32_t njobs;
    uint32_t* job_fail;   // optional: fault bits per job (njobs entries, zeroed by the caller)
    // Node2 trees: nodes, original triangle ids of the reordered triangles,
    // root box (tested as stored) and root kind (1: leaf group at `root`,
    // root_count triangles; 2: Node2 `root`)
    const sq_node2* nodes2;
    const uint32_t* perm;
    float rb[6];
    uint32_t root_kind, root_count;
    const sq_plucker* plk;  // non-null: precomputed triangle test
    // Node<K> trees (K = 4, 8): 8 K words per node; root fields as for Node2
    // (root_kind 0: nothing below the root box, 1: leaf at `root`, root_count
    // triangles, 2: node `root` with root_count used slots)
    const uint32_t* nodesk;
};

// Camera tile geometry of a job. `quad`: one camera traced by the per-ray
// kernel with exhaustive semantics, which takes 4x4-pixel tiles once a tile
// holds 64 rays or more (more, shorter tiles even out the end of a launch);
// closest hits, wave packets and many-job launches take 8x8. Returns the
// tile count; callers reject 2^31 or more.
inline uint64_t camera_tiling(uint32_t w, uint32_t h, uint32_t spp, bool quad, TraceArgs& a) {
    a.tile_shift = (quad && spp >= 4) ? 2u : 3u;
    const uint64_t e = 1ull << a.tile_shift;
    const uint64_t tx = (w + e - 1) >> a.tile_shift, ty = (h + e - 1) >> a.tile_shift;
    a.tiles_x = (uint32_t)tx;
    a.num_tiles = (uint32_t)(tx * ty);
    return tx * ty;
}

// Why a ray was abandoned (its results are then invalid).
constexpr uint32_t kFaultStack = 1;  // deeper than LDS slots + spill capacity
constexpr uint32_t kFaultNode = 2;   // child index outside the node array
constexpr uint32_t kFaultTri = 4;    // leaf triangle index outside the scene

__device__ __forceinline__ uint32_t hash_u32(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

struct RayF {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
};

// tracer.cpp:202-216 generalised: ndc = fma((p + j)/dim, 2, -1), x scaled by
// w/h, (x, y) scaled by tan(fov/2), p = (x, y, 1), |p|^2 = fma(z,z,fma(x,x,y*y)).
__device__ __forceinline__ RayF camera_ray(const TraceArgs& a, uint32_t px, uint32_t py, uint32_t s) {
    float jx = 0.5f, jy = 0.5f;
    if (a.jitter) {
        uint32_t pix = py * a.w + px;
        uint32_t h1 = hash_u32(a.seed ^ hash_u32(pix * a.spp + s));
        uint32_t h2 = hash_u32(h1 ^ 0x9E3779B9u);
        jx = (float)(h1 >> 8) * (1.0f / 16777216.0f);
        jy = (float)(h2 >> 8) * (1.0f / 16777216.0f);
    }
    float w = (float)a.w, h = (float)a.h;
    float nx = __fmaf_rn(((float)px + jx) / w, 2.0f, -1.0f);
    float ny = __fmaf_rn(((float)py + jy) / h, 2.0f, -1.0f);
    nx = nx * (w / h);
    float tx = nx * a.tanh, ty = ny * a.tanh, tz = 1.0f;
    float len = sqrtf(__fmaf_rn(tz, tz, __fmaf_rn(tx, tx, ty * ty)));
    RayF r;
    r.ox = a.ox; r.oy = a.oy; r.oz = a.oz;
    r.dx = tx / len; r.dy = ty / len; r.dz = tz / len;
    r.tmin = a.tmin; r.tmax = a.tmax;
    return r;
}

// intersect_ray_triangle<1> (tracer.h:83-116).
__device__ __forceinline__ bool tri_test(const float v[9], const RayF& r, float& t) {
    const float eps = 0.000000001f;
    float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    float hx = __fmaf_rn(-e2y, r.dz, r.dy * e2z);
    float hy = __fmaf_rn(-e2z, r.dx, r.dz * e2x);
    float hz = __fmaf_rn(-e2x, r.dy, r.dx * e2y);
    float av = __fmaf_rn(e1z, hz, __fmaf_rn(e1y, hy, e1x * hx));
    bool miss = (av > -eps) & (av < eps);
    float f = 1.0f / av;
    float sx = r.ox - v[0], sy = r.oy - v[1], sz = r.oz - v[2];
    float u = f * __fmaf_rn(sz, hz, __fmaf_rn(sy, hy, sx * hx));
    miss |= (u < 0.0f) | (u > 1.0f);
    float qx = __fmaf_rn(-e1y, sz, sy * e1z);
    float qy = __fmaf_rn(-e1z, sx, sz * e1x);
    float qz = __fmaf_rn(-e1x, sy, sx * e1y);
    float vv = f * __fmaf_rn(r.dz, qz, __fmaf_rn(r.dy, qy, r.dx * qx));
    miss |= (vv < 0.0f) | ((u + vv) > 1.0f);
    t = f * __fmaf_rn(e2z, qz, __fmaf_rn(e2y, qy, e2x * qx));
    return (!miss) & (t > eps);
}

// Precomputed-triangle test (projection method; layout of the reference's
// PluckerTriangle): plane along the dominant axis w, edges scaled by the
// projected determinant. No fused multiply-adds: the CPU restatement and the
// precompute use the same plain operations.
__device__ __forceinline__ bool plucker_test(const sq_plucker* __restrict__ pl, uint32_t i, const RayF& r, float& t) {
    const float4* p = (const float4*)(pl + i);
    const float4 A = p[0], B = p[1], C = p[2];
    const float nu = A.x, nv = A.y, np = A.z, pu = A.w, pv = B.x;
    const int w = __float_as_int(B.y);
    const float e0u = B.z, e0v = B.w, e1u = C.x, e1v = C.y;
    t = __int_as_float(0x7f800000);
    if (w > 2 || w < 0) return false;
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    const int u = w == 0 ? 1 : 0, v = w == 2 ? 1 : 2;
    const float ow = w == 0 ? o[0] : (w == 1 ? o[1] : o[2]), dw = w == 0 ? d[0] : (w == 1 ? d[1] : d[2]);
    const float ou = u == 0 ? o[0] : o[1], du = u == 0 ? d[0] : d[1];
    const float ov = v == 1 ? o[1] : o[2], dv = v == 1 ? d[1] : d[2];
    const float den = (dw + nu * du) + nv * dv;
    if (den == 0.0f) return false;
    t = (np - ((ow + nu * ou) + nv * ov)) / den;
    const float hu = (ou + t * du) - pu, hv = (ov + t * dv) - pv;
    const float beta = hu * e1v - hv * e1u, gamma = hv * e0u - hu * e0v;
    return (beta >= 0.0f) & (gamma >= 0.0f) & ((beta + gamma) <= 1.0f) & (t > 0.000000001f);
}

struct NodeV {
    float mn[3], mx[3];
    uint32_t count, leaf, child;
};

__device__ __forceinline__ NodeV load_node(const sq_node* __restrict__ nodes, uint32_t i) {
    const float4* p = (const float4*)(nodes + i);
    float4 a = p[0], b = p[1];
    NodeV n;
    n.mn[0] = a.x; n.mn[1] = a.y; n.mn[2] = a.z;
    n.mx[0] = b.x; n.mx[1] = b.y; n.mx[2] = b.z;
    uint32_t cf = __float_as_uint(a.w);
    n.count = cf & 0xFFFFu;
    n.leaf = (cf >> 16) & 1u;
    n.child = __float_as_uint(b.w);
    return n;
}

// Launch arguments whose node array and triangle array each span less than
// 4 GiB (every index a fetch uses is bounds-checked first): the traversal
// then addresses them with 32-bit byte offsets from the scalar base pointer
// instead of 64-bit per-lane address arithmetic.
struct TraceArgsN : TraceArgs {};
inline bool narrow_ok(uint64_t node_count, uint64_t tri_count) {
    return node_count * 32 < (1ull << 32) && tri_count * 36 < (1ull << 32);
}
__device__ __forceinline__ NodeV fetch_node(const TraceArgs& a, uint32_t i) { return load_node(a.nodes, i); }
__device__ __forceinline__ NodeV fetch_node(const TraceArgsN& a, uint32_t i) {
    return load_node((const sq_node*)((const char*)a.nodes + (i << 5)), 0);
}
__device__ __forceinline__ void fetch_tri(const TraceArgs& a, uint32_t i, float v[9]) { sqd::load_tri(a.tris, i, v); }
__device__ __forceinline__ void fetch_tri(const TraceArgsN& a, uint32_t i, float v[9]) {
    sqd::load_tri((const sq_triangle*)((const char*)a.tris + i * 36u), 0, v);
}

// Slab test with the reference's operation order and selects. When every
// reciprocal direction component is finite no NaN can arise (|1/d| >= 1, so
// (b - o) * inv is finite or +-inf), and IEEE min/max then differ from the
// reference's ternaries only in the sign of a zero, which no comparison
// below can observe; that path uses the hardware min3/max3 forms.
__device__ __forceinline__ bool slab(const NodeV& n, const RayF& r, float ix, float iy, float iz, bool fast,
                                     float& front) {
    float t1x = (n.mn[0] - r.ox) * ix, t2x = (n.mx[0] - r.ox) * ix;
    float t1y = (n.mn[1] - r.oy) * iy, t2y = (n.mx[1] - r.oy) * iy;
    float t1z = (n.mn[2] - r.oz) * iz, t2z = (n.mx[2] - r.oz) * iz;
    if (fast) {
        front = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
        float back = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
        return (back >= front) & (back <= r.tmax);
    }
    float tnx = sqd::simd_min(t1x, t2x), tny = sqd::simd_min(t1y, t2y), tnz = sqd::simd_min(t1z, t2z);
    float txx = sqd::simd_max(t1x, t2x), txy = sqd::simd_max(t1y, t2y), txz = sqd::simd_max(t1z, t2z);
    front = sqd::simd_max(sqd::simd_max(tnx, tny), tnz);
    float back = sqd::simd_min(sqd::simd_min(txx, txy), txz);
    return (back >= front) & (back <= r.tmax);
}

struct Stack {
    uint32_t* lds;     // this lane's column: entry k at lds[(k & mask) * stride]
    float* lds_t;      // closest mode: entry distances
    uint32_t stride, mask, S;
    uint32_t* spill;   // this thread's spill area (closest: 2 words per entry)
    uint32_t cap;
    uint32_t sp;
    uint32_t fault;  // kFault* bits of the current ray
};

template <bool CLOSEST>
__device__ __forceinline__ void push(Stack& st, uint32_t e, float t) {
    const uint32_t slot = (st.sp & st.mask) * st.stride;
    if (st.sp >= st.S) {
        uint32_t k = st.sp - st.S;
        if (k >= st.cap) { st.fault |= kFaultStack; return; }
        if (CLOSEST) {
            st.spill[2 * k] = st.lds[slot];
            st.spill[2 * k + 1] = __float_as_uint(st.lds_t[slot]);
        } else {
            st.spill[k] = st.lds[slot];
        }
    }
    st.lds[slot] = e;
    if (CLOSEST) st.lds_t[slot] = t;
    st.sp++;
}

template <bool CLOSEST>
__device__ __forceinline__ uint32_t pop(Stack& st, float& t) {
    st.sp--;
    const uint32_t slot = (st.sp & st.mask) * st.stride;
    uint32_t e = st.lds[slot];
    if (CLOSEST) t = st.lds_t[slot];
    if (st.sp >= st.S) {
        uint32_t k = st.sp - st.S;
        if (CLOSEST) {
            st.lds[slot] = st.spill[2 * k];
            st.lds_t[slot] = __uint_as_float(st.spill[2 * k + 1]);
        } else {
            st.lds[slot] = st.spill[k];
        }
    }
    return e;
}

// A stack entry keeps a node index in its upper 28 bits, so Node1 trees
// are traced only when every node index is below 2^28 (node capacity
// checked at every launch: single trees, batches and served scenes).
constexpr uint32_t kMaxTraceNodes = 1u << 28;

__device__ __forceinline__ uint32_t encode_list(uint32_t parent, const NodeV& n) {
    if (n.count >= 1 && n.count <= 15 && n.child < (1u << 28)) return (n.child << 4) | (n.count - 1);
    return (parent << 4) | 15u;
}

__device__ __forceinline__ uint8_t heat_byte(uint32_t count, uint32_t spp) {
    float c = (float)count;
    if (spp > 1) c = c / (float)spp;
    float v = c / 200.0f;
    v = (v < 1.0f) ? v : 1.0f;
    return (uint8_t)(v * 255.0f);
}


// Caller-provided trace scratch (instead of the context's workspace): the
// tile counter and fault bits at word 0, work totals (two u64) at word
// kScratchStatsWord, per-lane stack spill areas from word kScratchSpillWord.
struct TraceScratch {
    uint32_t* base;
    size_t words;
};
constexpr size_t kScratchStatsWord = 16, kScratchSpillWord = 64;
// Traces n device rays into device outputs (io on the device); ext == NULL
// uses the context's grow-only workspace, otherwise only ext's memory.
int trace_rays_device(sq_ctx* c, const sq_bvh* b, const sq_bvh2* t2, const sq_ray* rays, uint32_t n,
                      const sq_trace_params* p, const sq_trace_out* out, const TraceScratch* ext);

// Launch geometry and stack sizing of the traversal kernel.
struct TracePlan {
    bool closest;
    uint32_t lists;  // child lists per exhaustive step (k_trace's WL)
    uint32_t slots, block, grid, spill_cap;
    size_t lds_bytes, spill_words;
};
// coherent: one camera's primary rays (a smaller default launch).
// fmt: 1 Node1, 2 Node2, 4 / 8 Node<K>.
int trace_plan(sq_ctx* c, const sq_trace_params* p, uint32_t depth, bool coherent, TracePlan* out,
               int fmt = 1, bool ray_list = false);
// SQ_ERR_OVERFLOW with a message naming the fault bits.
int trace_fault(uint32_t bits, const char* what);
// Zeroes a.work / a.stats and launches k_trace (multi: a.jobs / a.job_tiles).
// narrow: every job's node and triangle arrays span less than 4 GiB (narrow_ok).
int trace_launch_plan(hipStream_t stream, const TracePlan& L, const TraceArgs& a, bool multi, bool narrow);
// Same over a Node2 tree (a.nodes2, a.perm, root fields).
int trace_launch_plan2(hipStream_t stream, const TracePlan& L, const TraceArgs& a);

// Wave-packet traversal of camera rays, exhaustive semantics (trace_packet.hip):
// grid size for a launch, and the launch (zeroes a.work / a.stats first).
uint32_t trace_packet_grid(sq_ctx* c, bool multi);
int trace_launch_packets(hipStream_t stream, uint32_t grid, const TraceArgs& a, bool multi);
// Kernel choice for camera rays: wave packets when the semantics allow.
bool use_packets(const sq_trace_params* p);

// Packet / bundle heatmap launch (heatmap.hip).
int heat_launch(sq_ctx* c, const sq_bvh* b, TraceArgs& t, const sq_trace_params* p);

}  // namespace sq
