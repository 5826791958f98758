// internal.h — shared definitions of the gfx950 BVH library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "sequoia.h"

namespace sq {

void set_error(const char* fmt, ...);

#define SQ_CHECK_HIP(call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            ::sq::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return e_ == hipErrorOutOfMemory ? SQ_ERR_NOMEM : SQ_ERR_HIP;                               \
        }                                                                                               \
    } while (0)

#define SQ_RET(code, ...)                    \
    do {                                     \
        ::sq::set_error(__VA_ARGS__);        \
        return (code);                       \
    } while (0)

#define SQ_TRY(expr)              \
    do {                          \
        int r_ = (expr);          \
        if (r_ != SQ_OK) return r_; \
    } while (0)

// Grow-only device buffer owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n);
    void release();
    template <typename T> T* as() const { return (T*)p; }
};

constexpr uint32_t kNoTask = 0xFFFFFFFFu;

// A range of the sorted primitive array that still has to be split.
struct Task {
    uint32_t begin, end;  // [begin, end) in sorted order
    uint32_t free;        // first free node slot for this range's children
    uint32_t self;        // node slot of the range's own interior node
};

// Events bracketing the phases of the last build/trace.
enum Phase { PH_BOUNDS, PH_MORTON, PH_SORT, PH_EMIT, PH_REFIT, PH_SAH, PH_GRID, PH_TRACE, PH_COUNT };

}  // namespace sq

struct sq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    // Workspace reused across calls (grow-only).
    sq::DevBuf keys[2], vals[2], hist, scratch, tasks[3], counters, spill, stats, aux[4];
    hipEvent_t ev_begin[sq::PH_COUNT] = {};
    hipEvent_t ev_end[sq::PH_COUNT] = {};
    bool phase_used[sq::PH_COUNT] = {};
    sq_timings last = {};
    void* pinned = nullptr;  // small pinned host staging area
    size_t pinned_bytes = 0;
};

struct sq_scene {
    sq_ctx* ctx = nullptr;
    sq_triangle* d_tris = nullptr;
    uint32_t n = 0;
    bool owned = false;
};

struct sq_bvh {
    sq_ctx* ctx = nullptr;
    const sq_scene* scene = nullptr;
    sq_node* d_nodes = nullptr;
    uint32_t capacity = 0;
    uint32_t root = 0;
    uint32_t depth = 0;  // levels from the root to the deepest node (root = 1)
};

namespace sq {

// Phase timing helpers (events recorded on the context stream).
void phase_begin(sq_ctx* c, Phase p);
void phase_end(sq_ctx* c, Phase p);
void phases_reset(sq_ctx* c);
int phases_collect(sq_ctx* c);
int pinned_ensure(sq_ctx* c, size_t bytes);

// Kernel launchers implemented across the .hip files.
int launch_fill_u32(sq_ctx* c, void* dst, uint32_t value, size_t words);
int scene_bounds(sq_ctx* c, const sq_triangle* tris, uint32_t n, uint32_t* d_ord6);
int morton_codes(sq_ctx* c, const sq_triangle* tris, uint32_t n, uint32_t bits, const uint32_t* d_ord6,
                 void* d_keys, uint32_t* d_vals, bool key64);
// Sorts (keys, vals) in c->keys/c->vals slot `in`; returns the slot holding the result.
int radix_sort(sq_ctx* c, uint32_t n, uint32_t key_bits, uint32_t digit_bits, bool key64, int in, int* out_slot);
int lbvh_emit(sq_ctx* c, const sq_triangle* tris, uint32_t n, const void* d_codes, const uint32_t* d_idx,
              bool key64, uint32_t layout, bool refit, uint32_t threshold, sq_node* d_nodes, uint32_t* depth);
int sah_build(sq_ctx* c, const sq_triangle* tris, uint32_t n, const sq_build_params* p, sq_builder kind,
              sq_node* d_nodes, uint32_t* root, uint32_t* depth);
int refit_tree(sq_ctx* c, sq_bvh* bvh);

}  // namespace sq

// ---------------------------------------------------------------------------
// Device helpers shared by the kernels.
// ---------------------------------------------------------------------------
namespace sqd {

// std::min(a,b) == (b < a) ? b : a ; std::max(a,b) == (a < b) ? b : a;
// the reference's simd max == (a > b) ? a : b. Spelled out so that NaN and
// +-0 ties resolve exactly as in the reference.
__device__ __forceinline__ float std_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float std_max(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float simd_max(float a, float b) { return (a > b) ? a : b; }
__device__ __forceinline__ float simd_min(float a, float b) { return (b < a) ? b : a; }

// Monotone float <-> uint32 map for atomic min/max reductions.
__device__ __forceinline__ uint32_t f2ord(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

struct Box3 {
    float mn[3], mx[3];
};

__device__ __forceinline__ void load_tri(const sq_triangle* __restrict__ tris, uint32_t i, float v[9]) {
    const float* p = (const float*)(tris + i);
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = p[k];
}

__device__ __forceinline__ Box3 tri_box(const float v[9]) {
    Box3 b;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        b.mn[k] = std_min(std_min(v[k], v[3 + k]), v[6 + k]);
        b.mx[k] = std_max(std_max(v[k], v[3 + k]), v[6 + k]);
    }
    return b;
}

__device__ __forceinline__ void box_grow(Box3& b, const Box3& a) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        b.mn[k] = std_min(b.mn[k], a.mn[k]);
        b.mx[k] = std_max(b.mx[k], a.mx[k]);
    }
}

__device__ __forceinline__ void store_node(sq_node* nodes, uint32_t i, const Box3& b, uint32_t child, uint32_t count,
                                           uint32_t flags) {
    float4* p = (float4*)(nodes + i);
    p[0] = make_float4(b.mn[0], b.mn[1], b.mn[2], __uint_as_float((count & 0xFFFFu) | (flags << 16)));
    p[1] = make_float4(b.mx[0], b.mx[1], b.mx[2], __uint_as_float(child));
}
__device__ __forceinline__ void store_node_links(sq_node* nodes, uint32_t i, uint32_t child, uint32_t count,
                                                 uint32_t flags) {
    uint32_t* p = (uint32_t*)(nodes + i);
    p[3] = (count & 0xFFFFu) | (flags << 16);
    p[7] = child;
}
__device__ __forceinline__ void store_node_box(sq_node* nodes, uint32_t i, const Box3& b) {
    float* p = (float*)(nodes + i);
    p[0] = b.mn[0]; p[1] = b.mn[1]; p[2] = b.mn[2];
    p[4] = b.mx[0]; p[5] = b.mx[1]; p[6] = b.mx[2];
}
__device__ __forceinline__ Box3 load_node_box(const sq_node* nodes, uint32_t i) {
    const float4* p = (const float4*)(nodes + i);
    float4 a = p[0], c = p[1];
    Box3 b;
    b.mn[0] = a.x; b.mn[1] = a.y; b.mn[2] = a.z;
    b.mx[0] = c.x; b.mx[1] = c.y; b.mx[2] = c.z;
    return b;
}

}  // namespace sqd
