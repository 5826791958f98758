// sah.hip — binned-SAH builders (recursive/horizontal) and the grid builder.
//
// Same trees as recurse/recurse_omp/recurse_sse/recurse_m128 and
// build_hierarchy{,_horizontal,_grid,_grid_sse,_grid_m128}
// (builder.cpp:114-147, 334-628, 683-903, 1053-1663, 1665-2194):
//   * choose_axis: longest centroid axis (z if strictly longest, y if > x and
//     >= z, else x);
//   * bin = int(k1 * (c - cmin)), k1 = NB(1 - 2^-23)/extent (an id of NB,
//     possible after rounding and unchecked in the reference, is clamped);
//   * choose_plane: prefix/suffix bin unions, score = fma(sa_l, n_l, sa_r*n_r)
//     with sa = fma(ly,lz, fma(lx,ly, lx*lz)) (x2 for AABB1), strict < while
//     scanning planes NB-2 .. 0;
//   * partition: the reference's in-place Hoare swap (recursive, horizontal,
//     grid) or the stable split of partition_sse/_m128 (grid_sse, grid_avx);
//   * leaves when count <= 8 or sa(centroid box) <= 0; the slot allocator of
//     recurse(), right-singleton quirk included (or fixed).
//
// GPU structure: tasks are processed level by level, one workgroup per task.
// Bins are reduced with per-wave LDS atomics on order-preserving integer
// images of the floats (min/max are exact and order-free), the plane is chosen
// by one lane, and the partition is done out of place between two id buffers.
// The Hoare result is produced without the sequential swap loop: the r-th
// right-side element found left of the final midpoint is exchanged with the
// r-th left-side element counted from the right end, which is exactly the
// pairing the two-pointer loop performs.
#include "internal.h"

namespace sq {

constexpr int kSahThreads = 256;
constexpr int kSahWaves = kSahThreads / 64;
constexpr int kMaxBins = 64;
constexpr int kBinWords = 13;  // count, p.min[3], p.max[3], c.min[3], c.max[3]

struct SahTask {
    uint32_t b, e, free, self;
    float cb[6];  // centroid box: min xyz, max xyz
    uint32_t pad[2];
};

struct SahArgs {
    const float4* boxes;  // 2 per primitive: (min, -), (max, -)
    sq_node* nodes;
    uint32_t cap;
    uint32_t nb;
    uint32_t part;     // 0 Hoare, 1 stable
    uint32_t sa_x2;    // 1: AABB1::sa (x2)
    uint32_t layout;
    const SahTask* tasks;
    SahTask* next;
    uint32_t* counters;  // [0] next count, [1] max level
    const uint32_t* ids_in;
    uint32_t* ids_out;
    uint32_t* scratch;   // Hoare misplaced positions
};

__device__ __forceinline__ float box_sa(const float* mn, const float* mx, uint32_t x2) {
    float lx = mx[0] - mn[0], ly = mx[1] - mn[1], lz = mx[2] - mn[2];
    float s = __fmaf_rn(ly, lz, __fmaf_rn(lx, ly, lx * lz));
    return x2 ? s + s : s;
}

__device__ __forceinline__ int choose_axis(const float* cb) {
    float lx = cb[3] - cb[0], ly = cb[4] - cb[1], lz = cb[5] - cb[2];
    int r = 0;
    r += 2 * (lz > lx && lz > ly);
    r += 1 * (ly > lx && ly >= lz);
    return r;
}

__device__ __forceinline__ float bin_k1(uint32_t nb, const float* cb, int axis) {
    const float epsilon = 1.1920929e-7f;
    return (float)nb * (1 - epsilon) / (cb[3 + axis] - cb[axis]);
}

__device__ __forceinline__ void prim_box(const float4* boxes, uint32_t id, float mn[3], float mx[3]) {
    float4 a = boxes[2 * id], b = boxes[2 * id + 1];
    mn[0] = a.x; mn[1] = a.y; mn[2] = a.z;
    mx[0] = b.x; mx[1] = b.y; mx[2] = b.z;
}

__device__ __forceinline__ void write_leaf_box(const SahArgs& a, uint32_t slot, uint32_t id) {
    if (slot >= a.cap) return;
    float mn[3], mx[3];
    prim_box(a.boxes, id, mn, mx);
    sqd::Box3 bx{{mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}};
    sqd::store_node(a.nodes, slot, bx, id, 1, 1);
}

__device__ __forceinline__ void write_inner_box(const SahArgs& a, uint32_t slot, const float* mn, const float* mx,
                                                uint32_t child, uint32_t count) {
    if (slot >= a.cap) return;
    sqd::Box3 bx{{mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}};
    sqd::store_node(a.nodes, slot, bx, child, count, 0);
}

// Exclusive block scan of one value per thread; returns prefix, sets *total.
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (int w = 0; w < kSahWaves; w++) {
        uint32_t t = s_tmp[w];
        if (w < wave) wbase += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return wbase + x - v;
}

// One workgroup per task: bin, choose the plane, partition, emit children.
__global__ __launch_bounds__(kSahThreads) void k_sah_level(SahArgs a) {
    __shared__ uint32_t s_bins[kSahWaves][kMaxBins][kBinWords];
    __shared__ float s_lp[6], s_rp[6], s_lc[6], s_rc[6];
    __shared__ int s_plane;
    __shared__ uint32_t s_m, s_tmp[kSahWaves];

    const SahTask t = a.tasks[blockIdx.x];
    const uint32_t b = t.b, e = t.e, count = e - b;
    const int wave = threadIdx.x >> 6;
    const int axis = choose_axis(t.cb);
    const float k1 = bin_k1(a.nb, t.cb, axis);
    const float cmin = t.cb[axis];
    const uint32_t nb = a.nb;

    // ---- binning ----
    const uint32_t omax = sqd::f2ord(3.402823466e+38f), omin = sqd::f2ord(-3.402823466e+38f);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kSahWaves * nb * kBinWords; i += kSahThreads) {
        uint32_t w = i / (nb * kBinWords), r = i % (nb * kBinWords), bi = r / kBinWords, f = r % kBinWords;
        uint32_t v = f == 0 ? 0u : ((f - 1) % 6 < 3 ? omax : omin);
        s_bins[w][bi][f] = v;
    }
    __syncthreads();
    for (uint32_t i = b + threadIdx.x; i < e; i += kSahThreads) {
        uint32_t id = a.ids_in[i];
        float mn[3], mx[3], c[3];
        prim_box(a.boxes, id, mn, mx);
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = (mn[k] + mx[k]) * 0.5f;
        int bi = (int)(k1 * (c[axis] - cmin));
        bi = bi < 0 ? 0 : (bi >= (int)nb ? (int)nb - 1 : bi);
        uint32_t* B = s_bins[wave][bi];
        atomicAdd(&B[0], 1u);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            atomicMin(&B[1 + k], sqd::f2ord(mn[k]));
            atomicMax(&B[4 + k], sqd::f2ord(mx[k]));
            atomicMin(&B[7 + k], sqd::f2ord(c[k]));
            atomicMax(&B[10 + k], sqd::f2ord(c[k]));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb * kBinWords; i += kSahThreads) {
        uint32_t bi = i / kBinWords, f = i % kBinWords;
        uint32_t v = s_bins[0][bi][f];
        for (int w = 1; w < kSahWaves; w++) {
            uint32_t x = s_bins[w][bi][f];
            if (f == 0) v += x;
            else if ((f - 1) % 6 < 3) v = x < v ? x : v;
            else v = x > v ? x : v;
        }
        s_bins[0][bi][f] = v;
    }
    __syncthreads();

    // ---- choose_plane (one lane) ----
    if (threadIdx.x == 0) {
        // decoded bins: pmin pmax cmin cmax, count
        auto bin_box = [&](uint32_t bi, float* p, float* c, uint32_t& n) {
            const uint32_t* B = s_bins[0][bi];
            n = B[0];
            for (int k = 0; k < 3; k++) {
                p[k] = sqd::ord2f(B[1 + k]); p[3 + k] = sqd::ord2f(B[4 + k]);
                c[k] = sqd::ord2f(B[7 + k]); c[3 + k] = sqd::ord2f(B[10 + k]);
            }
        };
        auto grow = [](float* dst, const float* src) {
            for (int k = 0; k < 3; k++) {
                dst[k] = sqd::std_min(dst[k], src[k]);
                dst[3 + k] = sqd::std_max(dst[3 + k], src[3 + k]);
            }
        };
        // l2r prefix unions are rebuilt on demand: keep them in registers per plane via a small array
        float l2r_p[kMaxBins][6], l2r_c[kMaxBins][6];
        uint32_t l2r_n[kMaxBins];
        {
            float p[6], c[6];
            uint32_t n;
            bin_box(0, p, c, n);
            for (int k = 0; k < 6; k++) { l2r_p[0][k] = p[k]; l2r_c[0][k] = c[k]; }
            l2r_n[0] = n;
            for (uint32_t i = 1; i + 1 < nb; i++) {
                bin_box(i, p, c, n);
                for (int k = 0; k < 6; k++) { l2r_p[i][k] = l2r_p[i - 1][k]; l2r_c[i][k] = l2r_c[i - 1][k]; }
                grow(l2r_p[i], p);
                grow(l2r_c[i], c);
                l2r_n[i] = l2r_n[i - 1] + n;
            }
        }
        float rp[6], rc[6], p[6], c[6];
        uint32_t rn, n;
        bin_box(nb - 1, rp, rc, rn);
        int result = 0;
        float best = 3.402823466e+38f;
        bool set = false;
        for (int i = (int)nb - 2; i >= 0; i--) {
            float score = __fmaf_rn(box_sa(l2r_p[i], l2r_p[i] + 3, a.sa_x2), (float)(int)l2r_n[i],
                                    box_sa(rp, rp + 3, a.sa_x2) * (float)(int)rn);
            if (score < best) {
                best = score; result = i; set = true;
                for (int k = 0; k < 6; k++) { s_lp[k] = l2r_p[i][k]; s_lc[k] = l2r_c[i][k]; s_rp[k] = rp[k]; s_rc[k] = rc[k]; }
            }
            bin_box(i, p, c, n);
            grow(rp, p);
            grow(rc, c);
            rn += n;
        }
        if (!set) {
            float qp[6], qc[6];
            uint32_t qn;
            bin_box(nb - 1, qp, qc, qn);
            for (uint32_t i = nb - 2; i >= 1; i--) {
                bin_box(i, p, c, n);
                grow(qp, p);
                grow(qc, c);
            }
            for (int k = 0; k < 6; k++) { s_lp[k] = l2r_p[0][k]; s_lc[k] = l2r_c[0][k]; s_rp[k] = qp[k]; s_rc[k] = qc[k]; }
            result = 0;
        }
        s_plane = result;
        s_m = l2r_n[result];
    }
    __syncthreads();
    const int plane = s_plane;
    const uint32_t m = s_m;

    // ---- partition ids_in[b,e) -> ids_out[b,e) ----
    auto is_right = [&](uint32_t id) {
        float mn[3], mx[3];
        prim_box(a.boxes, id, mn, mx);
        float c = (mn[axis] + mx[axis]) * 0.5f;
        return (int)(k1 * (c - cmin)) > plane;
    };
    if (a.part == 1) {
        uint32_t run_l = 0;
        for (uint32_t base = b; base < e; base += kSahThreads) {
            uint32_t i = base + threadIdx.x;
            bool valid = i < e;
            uint32_t id = valid ? a.ids_in[i] : 0u;
            bool right = valid && is_right(id);
            uint32_t isl = (valid && !right) ? 1u : 0u, tot;
            uint32_t pre = block_scan(isl, s_tmp, &tot);
            if (valid) {
                uint32_t rank_l = run_l + pre;
                uint32_t pos = isl ? b + rank_l : b + m + ((i - b) - rank_l);
                if (pos < e) a.ids_out[pos] = id;
            }
            run_l += tot;
        }
    } else {
        uint32_t run_r = 0, run_l = 0;
        for (uint32_t base = b; base < e; base += kSahThreads) {
            uint32_t i = base + threadIdx.x;
            bool valid = i < e;
            uint32_t id = valid ? a.ids_in[i] : 0u;
            bool right = valid && is_right(id);
            bool mis_r = valid && right && i < b + m;
            bool mis_l = valid && !right && i >= b + m;
            uint32_t packed = (mis_r ? 1u : 0u) | (mis_l ? 0x10000u : 0u), tot;
            uint32_t pre = block_scan(packed, s_tmp, &tot);
            if (mis_r) {
                uint32_t r = run_r + (pre & 0xFFFFu);
                if (b + r < e) a.scratch[b + r] = i;
            } else if (mis_l) {
                uint32_t r = run_l + (pre >> 16);
                if (b + m + r < e) a.scratch[b + m + r] = i;
            } else if (valid) {
                a.ids_out[i] = id;
            }
            run_r += tot & 0xFFFFu;
            run_l += tot >> 16;
        }
        __syncthreads();
        const uint32_t k = run_r;  // == run_l
        for (uint32_t r = threadIdx.x; r < k; r += kSahThreads) {
            uint32_t pr = a.scratch[b + r], pl = a.scratch[b + m + (k - 1 - r)];
            a.ids_out[pr] = a.ids_in[pl];
            a.ids_out[pl] = a.ids_in[pr];
        }
    }
    __threadfence_block();
    __syncthreads();

    // ---- emit children (recurse(), builder.cpp:1112-1155) ----
    const uint32_t L = m, R = count - m, f = t.free;
    uint32_t left_index = f, fi = f + (L != 1);
    uint32_t right_index = fi;
    fi += (R != 1);
    uint32_t left_free = fi, right_free = fi + (L - 1) * 2 + (L == 1);
    if (a.layout == SQ_LAYOUT_FIXED && R == 1 && L != 1) {
        right_index = f + 1;
        left_free = f + 2;
        right_free = f + 1;
    }
    const bool left_leaf = L <= 8 || box_sa(s_lc, s_lc + 3, a.sa_x2) <= 0.0f;
    const bool right_leaf = R <= 8 || box_sa(s_rc, s_rc + 3, a.sa_x2) <= 0.0f;
    if (left_leaf)
        for (uint32_t i = threadIdx.x; i < L; i += kSahThreads) write_leaf_box(a, left_free + i, a.ids_out[b + i]);
    if (right_leaf)
        for (uint32_t i = threadIdx.x; i < R; i += kSahThreads) write_leaf_box(a, right_free + i, a.ids_out[b + L + i]);
    if (threadIdx.x == 0) {
        if (L != 1) write_inner_box(a, left_index, s_lp, s_lp + 3, left_free, left_leaf ? L : 2);
        if (R != 1) write_inner_box(a, right_index, s_rp, s_rp + 3, right_free, right_leaf ? R : 2);
        if (!left_leaf) {
            SahTask c;
            c.b = b; c.e = b + L; c.free = left_free; c.self = left_index;
            for (int k = 0; k < 6; k++) c.cb[k] = s_lc[k];
            a.next[atomicAdd(&a.counters[0], 1u)] = c;
        }
        if (!right_leaf) {
            SahTask c;
            c.b = b + L; c.e = e; c.free = right_free; c.self = right_index;
            for (int k = 0; k < 6; k++) c.cb[k] = s_rc[k];
            a.next[atomicAdd(&a.counters[0], 1u)] = c;
        }
    }
}

// Per-primitive boxes, identity ids and the scene/centroid boxes (setup_reduction,
// builder.cpp:114-147). red[0..11]: p.min, p.max, c.min, c.max (ordered ints).
__global__ __launch_bounds__(256) void k_sah_setup(const sq_triangle* __restrict__ tris, uint32_t n,
                                                   float4* __restrict__ boxes, uint32_t* __restrict__ ids,
                                                   uint32_t* __restrict__ red) {
    uint32_t lo[6], hi[6];
    for (int k = 0; k < 6; k++) { lo[k] = 0xFFFFFFFFu; hi[k] = 0u; }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float v[9];
        sqd::load_tri(tris, i, v);
        sqd::Box3 bx = sqd::tri_box(v);
        boxes[2 * i] = make_float4(bx.mn[0], bx.mn[1], bx.mn[2], 0.f);
        boxes[2 * i + 1] = make_float4(bx.mx[0], bx.mx[1], bx.mx[2], 0.f);
        ids[i] = i;
        for (int k = 0; k < 3; k++) {
            float c = (bx.mn[k] + bx.mx[k]) * 0.5f;
            uint32_t a0 = sqd::f2ord(bx.mn[k]), a1 = sqd::f2ord(bx.mx[k]), a2 = sqd::f2ord(c);
            lo[k] = a0 < lo[k] ? a0 : lo[k];
            hi[k] = a1 > hi[k] ? a1 : hi[k];
            lo[3 + k] = a2 < lo[3 + k] ? a2 : lo[3 + k];
            hi[3 + k] = a2 > hi[3 + k] ? a2 : hi[3 + k];
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 6; k++) {
            uint32_t x = __shfl_xor(lo[k], off, 64), y = __shfl_xor(hi[k], off, 64);
            lo[k] = x < lo[k] ? x : lo[k];
            hi[k] = y > hi[k] ? y : hi[k];
        }
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; k++) {
            atomicMin(&red[k], lo[k]);
            atomicMax(&red[3 + k], hi[k]);
            atomicMin(&red[6 + k], lo[3 + k]);
            atomicMax(&red[9 + k], hi[3 + k]);
        }
    }
}

__global__ void k_sah_red_init(uint32_t* red, uint32_t* counters) {
    int k = threadIdx.x;
    if (k < 12) red[k] = ((k / 3) % 2 == 0) ? 0xFFFFFFFFu : 0u;
    if (k < 8) counters[k] = 0;
}

__device__ __forceinline__ void decode_box(const uint32_t* r, float* mn, float* mx) {
    for (int k = 0; k < 3; k++) { mn[k] = sqd::ord2f(r[k]); mx[k] = sqd::ord2f(r[3 + k]); }
}

// Root of build_hierarchy (builder.cpp:1587-1612): recursion from slot 0,
// root node at 2n-1 holding the scene box.
__global__ void k_sah_root(SahArgs a, uint32_t n, const uint32_t* red, SahTask* q0) {
    float pmn[3], pmx[3], cb[6];
    decode_box(red, pmn, pmx);
    decode_box(red + 6, cb, cb + 3);
    const uint32_t ri = 2 * n - 1;
    if (n <= 8 || box_sa(cb, cb + 3, a.sa_x2) <= 0.0f) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) write_leaf_box(a, i, a.ids_in[i]);
        if (threadIdx.x == 0) write_inner_box(a, ri, pmn, pmx, 0, n);
        return;
    }
    if (threadIdx.x == 0) {
        write_inner_box(a, ri, pmn, pmx, 0, 2);
        SahTask t;
        t.b = 0; t.e = n; t.free = 0; t.self = ri;
        for (int k = 0; k < 6; k++) t.cb[k] = cb[k];
        q0[0] = t;
        a.counters[0] = 1;
    }
}

// ---- grid builder ----
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 10; i++) r |= ((x >> i) & 1u) << (3 * i);
    return r;
}

// Cell of each primitive in Morton cell order (compact_1_by_2 inverse) and the
// per-cell bins (bin_centroids_to_grid_sse, builder.cpp:1755-1839).
__global__ __launch_bounds__(256) void k_grid_cells(const float4* __restrict__ boxes, uint32_t n, uint32_t dim,
                                                    const uint32_t* __restrict__ red, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals, uint32_t* __restrict__ cellbins) {
    float cb[6];
    decode_box(red + 6, cb, cb + 3);
    float k1[3];
    const float epsilon = 1.1920929e-7f;
    for (int k = 0; k < 3; k++) k1[k] = (float)dim * (1 - epsilon) / (cb[3 + k] - cb[k]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 A = boxes[2 * i], B = boxes[2 * i + 1];
        float mn[3] = {A.x, A.y, A.z}, mx[3] = {B.x, B.y, B.z}, c[3];
        uint32_t q[3];
        for (int k = 0; k < 3; k++) {
            c[k] = (mn[k] + mx[k]) * 0.5f;
            int bi = (int)(k1[k] * (c[k] - cb[k]));
            bi = bi < 0 ? 0 : (bi >= (int)dim ? (int)dim - 1 : bi);
            q[k] = (uint32_t)bi;
        }
        uint32_t cell = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
        keys[i] = cell;
        vals[i] = i;
        uint32_t* C = cellbins + cell * kBinWords;
        atomicAdd(&C[0], 1u);
        for (int k = 0; k < 3; k++) {
            atomicMin(&C[1 + k], sqd::f2ord(mn[k]));
            atomicMax(&C[4 + k], sqd::f2ord(mx[k]));
            atomicMin(&C[7 + k], sqd::f2ord(c[k]));
            atomicMax(&C[10 + k], sqd::f2ord(c[k]));
        }
    }
}

__global__ void k_grid_bins_init(uint32_t* cellbins, uint32_t cells) {
    const uint32_t omax = sqd::f2ord(3.402823466e+38f), omin = sqd::f2ord(-3.402823466e+38f);
    for (uint32_t i = threadIdx.x; i < cells * kBinWords; i += blockDim.x) {
        uint32_t f = i % kBinWords;
        cellbins[i] = f == 0 ? 0u : ((f - 1) % 6 < 3 ? omax : omin);
    }
}

// Cell nodes at cells-1+i, per-cell subtrees from 2*cells + 2*begin, heap top.
__global__ __launch_bounds__(256) void k_grid_init(SahArgs a, uint32_t cells, uint32_t cell_tight,
                                                   const uint32_t* cellbins, const uint32_t* red, SahTask* q0) {
    __shared__ uint32_t s_begin[513];
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i < cells; i++) { s_begin[i] = run; run += cellbins[i * kBinWords]; }
        s_begin[cells] = run;
    }
    __syncthreads();
    float pmn[3], pmx[3];
    decode_box(red, pmn, pmx);
    for (uint32_t i = 0; i < cells; i++) {
        const uint32_t* C = cellbins + i * kBinWords;
        const uint32_t cnt = C[0], begin = s_begin[i];
        float p[6], cb[6];
        for (int k = 0; k < 3; k++) {
            p[k] = sqd::ord2f(C[1 + k]); p[3 + k] = sqd::ord2f(C[4 + k]);
            cb[k] = sqd::ord2f(C[7 + k]); cb[3 + k] = sqd::ord2f(C[10 + k]);
        }
        const uint32_t free = 2 * cells + 2 * begin;
        const bool leaf = cnt <= 8 || box_sa(cb, cb + 3, a.sa_x2) <= 0.0f;
        if (leaf)
            for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) write_leaf_box(a, free + j, a.ids_in[begin + j]);
        if (threadIdx.x == 0) {
            const float* bmn = cell_tight ? p : pmn;
            const float* bmx = cell_tight ? p + 3 : pmx;
            write_inner_box(a, cells - 1 + i, bmn, bmx, free, leaf ? cnt : 2);
            if (!leaf) {
                SahTask t;
                t.b = begin; t.e = begin + cnt; t.free = free; t.self = cells - 1 + i;
                for (int k = 0; k < 6; k++) t.cb[k] = cb[k];
                q0[a.counters[0]++] = t;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int i = (int)cells - 2; i >= 0; i--) {
            sqd::Box3 l = sqd::load_node_box(a.nodes, 2 * i + 1), r = sqd::load_node_box(a.nodes, 2 * i + 2);
            sqd::Box3 u;
            for (int k = 0; k < 3; k++) { u.mn[k] = sqd::std_min(l.mn[k], r.mn[k]); u.mx[k] = sqd::std_max(l.mx[k], r.mx[k]); }
            sqd::store_node(a.nodes, i, u, 2 * i + 1, 2, 0);
        }
    }
}

int sah_build(sq_ctx* c, const sq_triangle* tris, uint32_t n, const sq_build_params* p, sq_builder kind,
              sq_node* d_nodes, uint32_t* root, uint32_t* depth) {
    const uint32_t nb = p && p->num_bins ? p->num_bins : 8;
    if (nb < 2 || nb > (uint32_t)kMaxBins) SQ_RET(SQ_ERR_INVALID, "num_bins %u outside 2..%d", nb, kMaxBins);
    const bool grid = kind == SQ_BUILDER_GRID || kind == SQ_BUILDER_GRID_SSE || kind == SQ_BUILDER_GRID_AVX;
    const uint32_t dim = p && p->grid_dim ? p->grid_dim : 4;
    if (grid && (dim < 1 || dim > 8 || (dim & (dim - 1)))) SQ_RET(SQ_ERR_INVALID, "grid_dim %u", dim);
    const uint32_t cells = dim * dim * dim;
    SahArgs a{};
    a.nodes = d_nodes;
    a.cap = 2 * n + (grid ? 2 * cells : 0);
    a.nb = nb;
    a.part = (kind == SQ_BUILDER_GRID_SSE || kind == SQ_BUILDER_GRID_AVX) ? 1u : 0u;
    a.sa_x2 = (kind == SQ_BUILDER_GRID_SSE || kind == SQ_BUILDER_GRID_AVX) ? 0u : 1u;
    a.layout = p ? p->layout : SQ_LAYOUT_REFERENCE;

    const size_t qcap = n / 9 + 2 + cells;
    SQ_TRY(c->aux[3].ensure((size_t)n * 32));             // boxes
    SQ_TRY(c->vals[0].ensure((size_t)n * 4));              // ids ping
    SQ_TRY(c->vals[1].ensure((size_t)n * 4));              // ids pong
    SQ_TRY(c->keys[0].ensure((size_t)n * 4));
    SQ_TRY(c->keys[1].ensure((size_t)n * 4));
    SQ_TRY(c->scratch.ensure((size_t)n * 4 + 256));
    SQ_TRY(c->tasks[0].ensure(qcap * sizeof(SahTask)));
    SQ_TRY(c->tasks[1].ensure(qcap * sizeof(SahTask)));
    SQ_TRY(c->aux[0].ensure(64 * 4));
    SQ_TRY(c->counters.ensure(64));
    SQ_TRY(c->aux[2].ensure((size_t)cells * kBinWords * 4 + 64));
    float4* boxes = c->aux[3].as<float4>();
    uint32_t* red = c->aux[0].as<uint32_t>();
    uint32_t* counters = c->counters.as<uint32_t>();
    a.boxes = boxes;
    a.counters = counters;
    a.scratch = c->scratch.as<uint32_t>();

    phase_begin(c, grid ? PH_GRID : PH_SAH);
    hipLaunchKernelGGL(k_sah_red_init, dim3(1), dim3(64), 0, c->stream, red, counters);
    uint32_t blocks = (n + 255) / 256, capb = (uint32_t)c->num_cus * 8;
    if (blocks > capb) blocks = capb;
    hipLaunchKernelGGL(k_sah_setup, dim3(blocks), dim3(256), 0, c->stream, tris, n, boxes, c->vals[0].as<uint32_t>(), red);
    int in = 0;
    if (grid) {
        uint32_t* cellbins = c->aux[2].as<uint32_t>();
        hipLaunchKernelGGL(k_grid_bins_init, dim3(1), dim3(256), 0, c->stream, cellbins, cells);
        hipLaunchKernelGGL(k_grid_cells, dim3(blocks), dim3(256), 0, c->stream, boxes, n, dim, red,
                           c->keys[0].as<uint32_t>(), c->vals[0].as<uint32_t>(), cellbins);
        int slot = 0;
        uint32_t kbits = 0;
        while ((1u << kbits) < cells) kbits++;
        SQ_TRY(radix_sort(c, n, kbits ? kbits : 1, 8, false, 0, &slot));
        in = slot;
        a.ids_in = c->vals[in].as<uint32_t>();
        hipLaunchKernelGGL(k_grid_init, dim3(1), dim3(256), 0, c->stream, a, cells,
                           kind == SQ_BUILDER_GRID ? 0u : 1u, cellbins, red, c->tasks[0].as<SahTask>());
        *root = 0;
    } else {
        a.ids_in = c->vals[0].as<uint32_t>();
        hipLaunchKernelGGL(k_sah_root, dim3(1), dim3(256), 0, c->stream, a, n, red, c->tasks[0].as<SahTask>());
        *root = 2 * n - 1;
    }
    SQ_CHECK_HIP(hipGetLastError());
    SQ_TRY(pinned_ensure(c, 64));
    uint32_t* h = (uint32_t*)c->pinned;
    SQ_CHECK_HIP(hipMemcpyAsync(h, counters, 4, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    uint32_t ncur = h[0], level = 0;
    int q = 0;
    while (ncur > 0) {
        if (ncur > qcap) SQ_RET(SQ_ERR_OVERFLOW, "sah: task queue overflow (%u > %zu)", ncur, qcap);
        hipLaunchKernelGGL(k_sah_red_init, dim3(1), dim3(1), 0, c->stream, red + 32, counters);  // counters[0] = 0
        a.tasks = c->tasks[q].as<SahTask>();
        a.next = c->tasks[q ^ 1].as<SahTask>();
        a.ids_in = c->vals[in].as<uint32_t>();
        a.ids_out = c->vals[in ^ 1].as<uint32_t>();
        hipLaunchKernelGGL(k_sah_level, dim3(ncur), dim3(kSahThreads), 0, c->stream, a);
        SQ_CHECK_HIP(hipGetLastError());
        SQ_CHECK_HIP(hipMemcpyAsync(h, counters, 4, hipMemcpyDeviceToHost, c->stream));
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
        ncur = h[0];
        q ^= 1;
        in ^= 1;
        level++;
        if (level > 4 * 1024 * 1024) SQ_RET(SQ_ERR_OVERFLOW, "sah: runaway recursion");
    }
    phase_end(c, grid ? PH_GRID : PH_SAH);
    *depth = level + 2 + (grid ? 7 : 0);
    return SQ_OK;
}

int refit_tree(sq_ctx* c, sq_bvh* b) {
    (void)c; (void)b;
    SQ_RET(SQ_ERR_UNSUPPORTED, "refit is not available in this build");
}

}  // namespace sq
