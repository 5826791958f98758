// trace_packet.hip — wave-packet traversal of camera rays (exhaustive).
//
// One wave traces the 64 pixels of an 8x8 tile, one sample at a time. The
// wave walks the union of its rays' traversals: a stack entry is a child
// list plus the 64-bit mask of lanes whose ray hit the list's parent; each
// node of the list is fetched once for the whole wave (one uniform address),
// and every lane in the mask runs the reference slab
// test on it. A lane counts one test per node it is active for, which is the
// reference's num_tests (every child of every hit interior node, root
// included) whatever order the wave visits nodes in; triangle tests of a hit
// leaf run on exactly the lanes that hit it, so closest hits (t, then lowest
// id) equal the per-ray kernel's.
//
// Against the per-lane kernel this trades lane utilisation (the union is
// longer than each ray's walk) for 64x fewer node fetches: coherent primary
// rays share most of their nodes, so the walk is bound by one load chain per
// wave instead of by 64 divergent loads per step.
#include "trace_common.h"

namespace sq {

constexpr int kPkWaves = 2;        // waves per workgroup
constexpr int kPkStack = 192;      // stack entries per wave (LDS)

// Wave-uniform node and triangle fetches: every lane loads the same address,
// which the vector memory path serves as one cache-line request per
// instruction (the scalar cache keeps too few misses in flight for a
// pointer chase over a tree far larger than it).
__device__ __forceinline__ NodeV load_node_uniform(const sq_node* nodes, uint32_t i) { return load_node(nodes, i); }

__device__ __forceinline__ void load_tri_uniform(const sq_triangle* tris, uint32_t i, float v[9]) {
    sqd::load_tri(tris, i, v);
}

__device__ __forceinline__ uint32_t uni(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    return ((uint64_t)uni((uint32_t)(x >> 32)) << 32) | uni((uint32_t)x);
}

struct PkLane {
    RayF r;
    float ix, iy, iz;
    bool fast;
    uint32_t tests, tri_tests, best_id, fault;
    float best_t;
};

// Triangle tests of a hit leaf on the lanes in `m`.
__device__ __forceinline__ void pk_leaf(const TraceArgs& a, const NodeV& n, bool mine, PkLane& L) {
    for (uint32_t j = 0; j < n.count; j++) {
        const uint32_t tid = n.child + j;
        if (tid >= a.tri_count) {
            if (mine) L.fault |= kFaultTri;
            break;
        }
        float v[9];
        load_tri_uniform(a.tris, tid, v);
        if (mine) {
            float t;
            L.tri_tests++;
            const bool m = tri_test(v, L.r, t);
            if (m && t > L.r.tmin && t < L.r.tmax && (t < L.best_t || (t == L.best_t && tid < L.best_id))) {
                L.best_t = t;
                L.best_id = tid;
            }
        }
    }
}

// One 64-ray packet: the union walk of the tree for the lanes in `m0`.
__device__ __forceinline__ void pk_walk(const TraceArgs& a, uint64_t m0, uint32_t lane, uint32_t* s_enc,
                                        uint64_t* s_msk, PkLane& L) {
    uint32_t sp = 0;
    if (m0) {
        s_enc[0] = (a.root << 4) | 0u;  // list of one node: the root
        s_msk[0] = m0;
        sp = 1;
    }
    const uint64_t me = 1ull << lane;
    while (sp > 0) {
        sp--;
        __builtin_amdgcn_wave_barrier();
        const uint32_t enc = uni(s_enc[sp]);
        const uint64_t m = uni64(s_msk[sp]);
        const bool active = (m & me) != 0;
        uint32_t first, cnt;
        if ((enc & 15u) != 15u) {
            first = enc >> 4;
            cnt = (enc & 15u) + 1;
        } else {
            const uint32_t parent = enc >> 4;
            if (parent >= a.node_count) {
                if (active) L.fault |= kFaultNode;
                continue;
            }
            const NodeV p = load_node_uniform(a.nodes, parent);
            first = p.child;
            cnt = p.count;
        }
        if (cnt > a.node_count || first > a.node_count - cnt) {
            if (active) L.fault |= kFaultNode;
            continue;
        }
        for (uint32_t j = 0; j < cnt; j += 2) {
            const uint32_t i0 = first + j;
            const bool two = j + 1 < cnt;
            const NodeV n0 = load_node_uniform(a.nodes, i0);
            NodeV n1;
            if (two) n1 = load_node_uniform(a.nodes, i0 + 1);
            bool h0 = false, h1 = false;
            float f;
            if (active) {
                L.tests += two ? 2u : 1u;
                h0 = slab(n0, L.r, L.ix, L.iy, L.iz, L.fast, f);
                if (two) h1 = slab(n1, L.r, L.ix, L.iy, L.iz, L.fast, f);
            }
            const uint64_t hm0 = __ballot(h0), hm1 = two ? __ballot(h1) : 0ull;
            if (hm0 && n0.leaf) pk_leaf(a, n0, h0, L);
            if (hm1 && n1.leaf) pk_leaf(a, n1, h1, L);
            // interior children: push lists (second first, so the first pops first)
            if (hm1 && !n1.leaf) {
                if (sp < (uint32_t)kPkStack) {
                    if (lane == 0) {
                        s_enc[sp] = encode_list(i0 + 1, n1);
                        s_msk[sp] = hm1;
                    }
                    sp++;
                } else if (h1) {
                    L.fault |= kFaultStack;
                }
            }
            if (hm0 && !n0.leaf) {
                if (sp < (uint32_t)kPkStack) {
                    if (lane == 0) {
                        s_enc[sp] = encode_list(i0, n0);
                        s_msk[sp] = hm0;
                    }
                    sp++;
                } else if (h0) {
                    L.fault |= kFaultStack;
                }
            }
        }
    }
}

// All samples of one 8x8 tile of job `a`, as packets of 4x4 pixels x 4
// samples (spp a multiple of 4): the rays of one packet start within a few
// pixels of each other, so their walks share more of the tree than the 64
// pixels of one sample do. The per-pixel counts gather in LDS.
__device__ __forceinline__ void pk_tile4(const TraceArgs& a, uint32_t tile, uint32_t lane, uint32_t* s_enc,
                                         uint64_t* s_msk, uint32_t* s_cnt, uint32_t& tests_total,
                                         uint32_t& tris_total, uint32_t& faults, uint32_t* job_fail) {
    const uint32_t tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    s_cnt[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t q = 0; q < 4; q++) {
        const uint32_t col = ((q & 1u) << 2) + (lane & 3u), row = ((q >> 1) << 2) + ((lane >> 2) & 3u);
        const uint32_t px = tx * 8 + col, py = ty * 8 + row;
        const bool valid = px < a.w && py < a.h;
        const uint64_t m0 = __ballot(valid);
        if (!m0) continue;
        for (uint32_t g = 0; g < a.spp; g += 4) {
            const uint32_t smp = g + (lane >> 4);
            PkLane L;
            L.tests = 0;
            L.tri_tests = 0;
            L.fault = 0;
            L.best_id = SQ_MISS;
            L.best_t = __int_as_float(0x7f800000);
            if (valid) {
                L.r = camera_ray(a, px, py, smp);
            } else {
                L.r = RayF{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f};
            }
            L.ix = 1.0f / L.r.dx; L.iy = 1.0f / L.r.dy; L.iz = 1.0f / L.r.dz;
            L.fast = isfinite(L.ix) && isfinite(L.iy) && isfinite(L.iz) && isfinite(L.r.ox) && isfinite(L.r.oy) &&
                     isfinite(L.r.oz);
            pk_walk(a, m0, lane, s_enc, s_msk, L);
            if (valid) {
                const size_t ri = ((size_t)py * a.w + px) * a.spp + smp;
                if (a.hit_id) a.hit_id[ri] = L.best_id;
                if (a.hit_t) a.hit_t[ri] = L.best_t;
                atomicAdd(&s_cnt[row * 8 + col], L.tests);
                tests_total += L.tests;
                tris_total += L.tri_tests;
                if (L.fault) {
                    faults |= L.fault;
                    if (job_fail) atomicOr(job_fail, L.fault);
                }
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t px = tx * 8 + (lane & 7), py = ty * 8 + (lane >> 3);
    if (px < a.w && py < a.h) {
        const uint32_t pix_tests = s_cnt[lane];
        const size_t p = (size_t)py * a.w + px;
        if (a.counts) a.counts[p] = pix_tests;
        if (a.rgb) {
            const uint8_t b = heat_byte(pix_tests, a.spp);
            a.rgb[3 * p] = b; a.rgb[3 * p + 1] = b; a.rgb[3 * p + 2] = b;
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// All samples of one 8x8 tile of job `a`, one sample of the 64 pixels per packet.
__device__ __forceinline__ void pk_tile(const TraceArgs& a, uint32_t tile, uint32_t lane, uint32_t* s_enc,
                                        uint64_t* s_msk, uint32_t& tests_total, uint32_t& tris_total,
                                        uint32_t& faults, uint32_t* job_fail) {
    const uint32_t tx = tile % a.tiles_x, ty = tile / a.tiles_x;
    const uint32_t px = tx * 8 + (lane & 7), py = ty * 8 + (lane >> 3);
    const bool valid = px < a.w && py < a.h;
    const uint64_t m0 = __ballot(valid);
    uint32_t pix_tests = 0;
    for (uint32_t smp = 0; smp < a.spp; smp++) {
        PkLane L;
        L.tests = 0;
        L.tri_tests = 0;
        L.fault = 0;
        L.best_id = SQ_MISS;
        L.best_t = __int_as_float(0x7f800000);
        if (valid) {
            L.r = camera_ray(a, px, py, smp);
        } else {
            L.r = RayF{0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f};
        }
        L.ix = 1.0f / L.r.dx; L.iy = 1.0f / L.r.dy; L.iz = 1.0f / L.r.dz;
        L.fast = isfinite(L.ix) && isfinite(L.iy) && isfinite(L.iz) && isfinite(L.r.ox) && isfinite(L.r.oy) &&
                 isfinite(L.r.oz);
        pk_walk(a, m0, lane, s_enc, s_msk, L);
        if (valid) {
            const size_t ri = ((size_t)py * a.w + px) * a.spp + smp;
            if (a.hit_id) a.hit_id[ri] = L.best_id;
            if (a.hit_t) a.hit_t[ri] = L.best_t;
            pix_tests += L.tests;
            tests_total += L.tests;
            tris_total += L.tri_tests;
            if (L.fault) {
                faults |= L.fault;
                if (job_fail) atomicOr(job_fail, L.fault);
            }
        }
    }
    if (valid) {
        const size_t p = (size_t)py * a.w + px;
        if (a.counts) a.counts[p] = pix_tests;
        if (a.rgb) {
            const uint8_t b = heat_byte(pix_tests, a.spp);
            a.rgb[3 * p] = b; a.rgb[3 * p + 1] = b; a.rgb[3 * p + 2] = b;
        }
    }
}

template <bool MULTI>
__global__ __launch_bounds__(64 * kPkWaves) void k_trace_pk(TraceArgs a) {
    __shared__ uint32_t s_enc[kPkWaves][kPkStack];
    __shared__ uint64_t s_msk[kPkWaves][kPkStack];
    __shared__ uint32_t s_cnt[kPkWaves][64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t tests_total = 0, tris_total = 0, faults = 0;
    while (true) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(&a.work[0], 1u);
        tile = uni(__shfl(tile, 0, 64));
        if (tile >= a.num_tiles) break;
        if (MULTI) {
#if defined(__HIP_DEVICE_COMPILE__)
            typedef const __attribute__((address_space(4))) TraceArgs CJob;
            typedef const __attribute__((address_space(4))) uint32_t CU32;
#else
            typedef const TraceArgs CJob;
            typedef const uint32_t CU32;
#endif
            const CJob* jobs = (const CJob*)a.jobs;
            const CU32* jt = (const CU32*)a.job_tiles;
            uint32_t lo = 0, hi = a.njobs;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (jt[mid] <= tile) lo = mid;
                else hi = mid;
            }
            const TraceArgs J = jobs[lo];
            if ((J.spp & 3u) == 0)
                pk_tile4(J, tile - jt[lo], lane, s_enc[wave], s_msk[wave], s_cnt[wave], tests_total, tris_total,
                         faults, a.job_fail ? a.job_fail + lo : nullptr);
            else
                pk_tile(J, tile - jt[lo], lane, s_enc[wave], s_msk[wave], tests_total, tris_total, faults,
                        a.job_fail ? a.job_fail + lo : nullptr);
        } else if ((a.spp & 3u) == 0) {
            pk_tile4(a, tile, lane, s_enc[wave], s_msk[wave], s_cnt[wave], tests_total, tris_total, faults, nullptr);
        } else {
            pk_tile(a, tile, lane, s_enc[wave], s_msk[wave], tests_total, tris_total, faults, nullptr);
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

__global__ void k_pk_init(uint32_t* work, unsigned long long* stats) {
    if (threadIdx.x < 4) work[threadIdx.x] = 0;
    if (stats && threadIdx.x < 2) stats[threadIdx.x] = 0;
}

uint32_t trace_packet_grid(sq_ctx* c, bool multi) {
    int per_cu = 0;
    if (multi) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_pk<true>, 64 * kPkWaves, 0);
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_pk<false>, 64 * kPkWaves, 0);
    if (per_cu < 1) per_cu = 8;
    return (uint32_t)c->num_cus * (uint32_t)per_cu;
}

int trace_launch_packets(hipStream_t stream, uint32_t grid, const TraceArgs& a, bool multi) {
    hipLaunchKernelGGL(k_pk_init, dim3(1), dim3(64), 0, stream, a.work, a.stats);
    if (multi) hipLaunchKernelGGL((k_trace_pk<true>), dim3(grid), dim3(64 * kPkWaves), 0, stream, a);
    else hipLaunchKernelGGL((k_trace_pk<false>), dim3(grid), dim3(64 * kPkWaves), 0, stream, a);
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

}  // namespace sq
