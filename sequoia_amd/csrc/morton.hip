// morton.hip — scene bounds reduction and Morton-code generation.
//
// setup_morton / setup_morton2 (builder.cpp:2204-2279): the scene box is the
// union of per-triangle boxes; each triangle's box centre is normalised to
// the scene box and truncated to `bits` bits per axis, then interleaved
// x -> bit 3i, y -> 3i+1, z -> 3i+2 (bit_interleave_32/64, algorithms.cpp:8-22).
// Keys are written with values = triangle index for the stable sort.
//
// Algorithmic traffic: the bounds pass reads 36 B/triangle, the code pass
// reads 36 B and writes key + value (8 B for 30-bit, 12 B for wider codes).
#include "internal.h"

namespace sq {

constexpr int kBoundsBlock = 256;

// Block-reduces the six box extremes (ordered uints: min x,y,z then max
// x,y,z). Each block writes its six values to ord6[16 + 6 * block + k]; the
// block that takes the last ticket (a context word that is zero between
// launches: the last block resets it) folds all partials into ord6[0..5]. Atomics on the six result words from every
// block serialise at L2 and cost about 50 us whatever the triangle count.
// VEC: the triangle array is 16-byte aligned and read in chunks of 256
// triangles per wave (below); otherwise one triangle per thread and step.
template <bool VEC>
__global__ __launch_bounds__(kBoundsBlock) void k_scene_bounds(const sq_triangle* __restrict__ tris, uint32_t n,
                                                                uint32_t* __restrict__ ord6, uint32_t* ticket) {
    uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t hi[3] = {0u, 0u, 0u};
    auto take = [&](const float* v) {
        const sqd::Box3 b = sqd::tri_box(v);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t a = sqd::f2ord(b.mn[k]), c = sqd::f2ord(b.mx[k]);
            lo[k] = a < lo[k] ? a : lo[k];
            hi[k] = c > hi[k] ? c : hi[k];
        }
    };
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t done = 0;  // triangles covered by the chunked path
    if (VEC) {
        // A wave reads 256 consecutive triangles (576 float4) with coalesced
        // loads, passes them through LDS, and each lane then takes four whole
        // triangles from it: every load instruction covers 1 KB of contiguous
        // memory instead of 64 triangles 144 B apart.
        __shared__ float4 s_chunk[kBoundsBlock / 64][9 * 64];
        float4* sc = s_chunk[wave];
        const float4* q4 = (const float4*)tris;
        const uint32_t chunks = n / 256, waves = gridDim.x * (kBoundsBlock / 64);
        for (uint32_t ch = blockIdx.x * (kBoundsBlock / 64) + wave; ch < chunks; ch += waves) {
            float4 r[9];
#pragma unroll
            for (int j = 0; j < 9; j++) r[j] = q4[(size_t)ch * 576 + j * 64 + lane];
#pragma unroll
            for (int j = 0; j < 9; j++) sc[j * 64 + lane] = r[j];
            sqd::wave_lds_sync();
            float v[36];
#pragma unroll
            for (int j = 0; j < 9; j++) {
                const float4 f = sc[9 * lane + j];
                v[4 * j] = f.x; v[4 * j + 1] = f.y; v[4 * j + 2] = f.z; v[4 * j + 3] = f.w;
            }
            sqd::wave_lds_sync();  // the next chunk overwrites what this one read
#pragma unroll
            for (int t = 0; t < 4; t++) take(v + 9 * t);
        }
        done = chunks * 256;
    }
    for (uint32_t i = done + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float v[9];
        sqd::load_tri(tris, i, v);
        take(v);
    }
    __shared__ uint32_t s[6][kBoundsBlock / 64];
    __shared__ uint32_t s_last;
    auto block_reduce = [&]() {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
                uint32_t a = __shfl_xor(lo[k], off, 64), c = __shfl_xor(hi[k], off, 64);
                lo[k] = a < lo[k] ? a : lo[k];
                hi[k] = c > hi[k] ? c : hi[k];
            }
        }
        if (lane == 0) {
            for (int k = 0; k < 3; k++) { s[k][wave] = lo[k]; s[3 + k][wave] = hi[k]; }
        }
        __syncthreads();
    };
    block_reduce();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t r = s[k][0];
        for (int w = 1; w < kBoundsBlock / 64; w++) r = (k < 3) ? (s[k][w] < r ? s[k][w] : r) : (s[k][w] > r ? s[k][w] : r);
        ord6[16 + 6 * blockIdx.x + k] = r;
        __threadfence();
    }
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(ticket, 1u) == gridDim.x - 1 ? 1u : 0u;
    __syncthreads();
    if (!s_last) return;
    if (threadIdx.x == 0) *ticket = 0u;  // every other block has taken its ticket: ready for the next launch
    __threadfence();
    // last block: fold every block's partial. The fence after the ticket
    // (agent-scope acquire) invalidates this CU's L1, so plain loads read the
    // other blocks' partials from L2, and they can all be in flight at once
    // (volatile loads would each be waited for: ~50 us over 2048 blocks).
    const uint32_t* part = ord6 + 16;
    for (int k = 0; k < 3; k++) { lo[k] = 0xFFFFFFFFu; hi[k] = 0u; }
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t a = part[6 * b + k], c = part[6 * b + 3 + k];
            lo[k] = a < lo[k] ? a : lo[k];
            hi[k] = c > hi[k] ? c : hi[k];
        }
    }
    block_reduce();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        uint32_t r = s[k][0];
        for (int w = 1; w < kBoundsBlock / 64; w++) r = (k < 3) ? (s[k][w] < r ? s[k][w] : r) : (s[k][w] > r ? s[k][w] : r);
        ord6[k] = r;
    }
}

// Two workgroups per CU keep enough loads in flight to stream the triangles
// and leave the last block few partials to fold.
static uint32_t bounds_blocks(const sq_ctx* c, uint32_t n) {
    uint32_t blocks = (n / 4 + 1 + kBoundsBlock - 1) / kBoundsBlock;
    const uint32_t cap = (uint32_t)c->num_cus * 2;
    if (blocks > cap) blocks = cap;
    return blocks ? blocks : 1;
}

size_t scene_bounds_bytes(const sq_ctx* c, uint32_t n) { return (16 + 6 * (size_t)bounds_blocks(c, n)) * 4; }

int scene_bounds(sq_ctx* c, const sq_triangle* tris, uint32_t n, uint32_t* d_ord6) {
    if (((uintptr_t)tris & 15u) == 0)
        hipLaunchKernelGGL(k_scene_bounds<true>, dim3(bounds_blocks(c, n)), dim3(kBoundsBlock), 0, c->stream, tris, n,
                           d_ord6, c->d_bounds_ticket);
    else
        hipLaunchKernelGGL(k_scene_bounds<false>, dim3(bounds_blocks(c, n)), dim3(kBoundsBlock), 0, c->stream, tris,
                           n, d_ord6, c->d_bounds_ticket);
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

template <bool KEY64>
__global__ __launch_bounds__(256) void k_morton(const sq_triangle* __restrict__ tris, uint32_t n, uint32_t bits,
                                                const uint32_t* __restrict__ ord6, void* __restrict__ keys_out,
                                                uint32_t* __restrict__ vals_out) {
    float smn[3], ext[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        smn[k] = sqd::ord2f(ord6[k]);
        ext[k] = sqd::ord2f(ord6[3 + k]) - smn[k];
    }
    const float scale = (float)((1u << bits) - 1u);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float v[9];
        sqd::load_tri(tris, i, v);
        if (KEY64) ((uint64_t*)keys_out)[i] = sqd::morton_code<true>(v, smn, ext, scale);
        else ((uint32_t*)keys_out)[i] = (uint32_t)sqd::morton_code<false>(v, smn, ext, scale);
        vals_out[i] = i;
    }
}

int morton_codes(sq_ctx* c, const sq_triangle* tris, uint32_t n, uint32_t bits, const uint32_t* d_ord6, void* d_keys,
                 uint32_t* d_vals, bool key64) {
    uint32_t blocks = (n + 255) / 256;
    if (blocks == 0) blocks = 1;
    if (key64)
        hipLaunchKernelGGL(k_morton<true>, dim3(blocks), dim3(256), 0, c->stream, tris, n, bits, d_ord6, d_keys, d_vals);
    else
        hipLaunchKernelGGL(k_morton<false>, dim3(blocks), dim3(256), 0, c->stream, tris, n, bits, d_ord6, d_keys, d_vals);
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

}  // namespace sq
