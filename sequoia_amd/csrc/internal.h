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
    uint32_t begin, end;  // [begin, end) in sorted order (positions in the batch)
    uint32_t free;        // first free node slot for this range's children (scene-local)
    uint32_t self;        // node slot of the range's own interior node (scene-local)
    uint32_t scene;       // scene of the batch
};

// One scene of a batched Morton build. Positions index the batch's sorted
// key/value arrays; node slots and leaf triangle ids written into the
// scene's Node1[] are scene-local, as in a plain build.
struct SceneDesc {
    uint32_t seg_begin;   // first sorted position of the scene
    uint32_t n;           // triangles
    uint32_t tri_base;    // first triangle of the scene in the batch's triangle array
    uint32_t node_base;   // first node of the scene in the batch's node array
    uint32_t q_base;      // offset of the scene's top-level task queues
    uint32_t top_base;    // offset of the scene's top node list
    uint32_t lvl_base;    // offset of the scene's level table
    uint32_t top_cap;     // entries of the scene's top node list
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
    sq::DevBuf sort_os;  // single-pass radix sort area (sort_onesweep_bytes)
    sq::DevBuf sah[6];  // binned-SAH large-task queues, plans, partial bins, counts, bin ids
    hipEvent_t ev_begin[sq::PH_COUNT] = {};
    hipEvent_t ev_end[sq::PH_COUNT] = {};
    bool phase_used[sq::PH_COUNT] = {};
    sq_timings last = {};
    void* pinned = nullptr;  // small pinned host staging area
    uint32_t* d_bounds_ticket = nullptr;  // scene-bounds last-block ticket (zero between launches)
    size_t pinned_bytes = 0;
    bool phase_timing = false;       // build phases timed separately (sq_ctx_set_phase_timing)
    // End-of-build event of the running sq_build: a builder that waits for a
    // result copy records it right after queueing that copy (the build's last
    // device work), so the build time excludes the host's wake-up after the wait.
    hipEvent_t build_end = nullptr;
    bool build_end_recorded = false;
    // Test hook (SEQUOIA_TEST_POISON_NODES at creation): node arrays of
    // Morton builds that skip the zero fill are first filled with all-ones
    // words, so a slot the emission fails to write shows up in the tree.
    bool poison_nodes = false;
    // Test hook (SEQUOIA_TEST_WIDE_FETCH at creation): traversal launches
    // keep 64-bit fetch addressing even when every array is small enough for
    // 32-bit offsets, so tests can compare the two forms.
    bool wide_fetch = false;
    uint32_t* h_sort_err = nullptr;  // pinned: the last single-pass sort's error word
    // pinned, coherent: a level loop's task counts posted by the device
    // (words 0..7) with a sequence number (word 8), read without a copy or a
    // stream synchronisation
    uint32_t* h_mail = nullptr;
    uint32_t mail_seq = 0;
    bool sort_err_pending = false;
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
    // Node1 array. It starts one 32-byte slot into its allocation (d_alloc)
    // when most child lists begin at odd slots (Morton trees: root 0, first
    // list at 1), so that a sibling pair sits in one 64-byte segment instead
    // of straddling two.
    sq_node* d_nodes = nullptr;
    void* d_alloc = nullptr;
    uint32_t capacity = 0;
    uint32_t root = 0;
    uint32_t depth = 0;  // levels from the root to the deepest node (root = 1)
    bool view = false;   // part of a batch: storage owned by the batch
    sq_plucker* d_plucker = nullptr;  // precomputed triangles (sq_bvh_plucker)
};

// A Node2 tree (node2.hip): nodes, triangles in leaf-group order with their
// original indices, root record, breadth-first levels.
struct sq_bvh2 {
    sq_ctx* ctx = nullptr;
    sq_node2* d_nodes = nullptr;
    sq_triangle* d_tris = nullptr;
    uint32_t* d_perm = nullptr;
    uint32_t nodes = 0, tris = 0, levels = 0;
    sq_node2_root root{};
};

// A Node<K> tree (nodek.hip): K = 4 or 8, 8 K words per node, root record,
// breadth-first levels; leaves index the source tree's scene.
struct sq_bvhk {
    sq_ctx* ctx = nullptr;
    const sq_scene* scene = nullptr;
    uint32_t* d_nodes = nullptr;
    uint32_t k = 0, nodes = 0, levels = 0;
    sq_nodek_root root{};
};

namespace sq {

// Phase timing helpers (events recorded on the context stream).
void phase_begin(sq_ctx* c, Phase p);
void phase_end(sq_ctx* c, Phase p);
void phases_reset(sq_ctx* c);
int phases_collect(sq_ctx* c);
int pinned_ensure(sq_ctx* c, size_t bytes);
// Looks up, once per device (with that device current), the device address
// of the allocation-free entry points' scratch array (noalloc.hip).
int noalloc_setup(int device);

// Kernel launchers implemented across the .hip files.
int launch_fill_u32(sq_ctx* c, void* dst, uint32_t value, size_t words);
// Scene box of n triangles into d_ord6[0..5] (ordered uints); d_ord6 must hold
// scene_bounds_bytes(c, n) bytes (ticket and per-block partials follow).
size_t scene_bounds_bytes(const sq_ctx* c, uint32_t n);
int scene_bounds(sq_ctx* c, const sq_triangle* tris, uint32_t n, uint32_t* d_ord6);
int morton_codes(sq_ctx* c, const sq_triangle* tris, uint32_t n, uint32_t bits, const uint32_t* d_ord6,
                 void* d_keys, uint32_t* d_vals, bool key64);
// Sorts (keys, vals) in c->keys/c->vals slot `in`; returns the slot holding the result.
struct SortBufs {
    hipStream_t stream;
    void* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;  // sort_hist_bytes(n, digit_bits)
    // non-null: key passes run as single-pass look-back kernels in this area
    // (sort_onesweep_bytes); null: three kernels per pass
    uint32_t* os = nullptr;
    uint32_t* os_err = nullptr;  // pinned host word the look-back sets on a failed wait
    int num_cus = 256;
};
size_t sort_onesweep_bytes(uint32_t n, uint32_t key_bits, uint32_t digit_bits);
// Clears and returns the ctx's pinned sort error word; the next
// phases_collect (after its sync) reports a set word as a failed call.
int sort_error_word(sq_ctx* c, uint32_t** word);
size_t sort_hist_bytes(uint32_t n, uint32_t digit_bits);
uint32_t sort_segment_passes(uint32_t nseg, uint32_t digit_bits);
// Key passes, then (d_seg_off != NULL, nseg > 1) segment passes that group
// the pairs by the segment holding each value (nseg <= 2048).
int radix_sort_bufs(const SortBufs& b, uint32_t n, uint32_t key_bits, uint32_t digit_bits, bool key64,
                    const uint32_t* d_seg_off, uint32_t nseg, int in, int* out_slot, uint32_t* passes);
int radix_sort(sq_ctx* c, uint32_t n, uint32_t key_bits, uint32_t digit_bits, bool key64, int in, int* out_slot);
// Single-scene emission from sorted codes of code_bits bits.
int lbvh_emit(sq_ctx* c, const sq_triangle* tris, uint32_t n, const void* d_codes, const uint32_t* d_idx,
              bool key64, uint32_t code_bits, uint32_t layout, bool refit, uint32_t threshold, sq_node* d_nodes,
              uint32_t* depth);
// Batched emission: scenes[0..nscenes) on device; no host synchronisation.
// `ws` must hold lbvh_batch_workspace(...) bytes.
size_t lbvh_batch_workspace(uint32_t total_tris, uint32_t nscenes, uint32_t threshold);
void lbvh_fill_descs(SceneDesc* descs, const uint32_t* counts, uint32_t nscenes, uint32_t threshold);
int lbvh_emit_batch(sq_ctx* c, hipStream_t stream, const sq_triangle* tris, uint32_t total_tris, const void* d_codes,
                    const uint32_t* d_idx, bool key64, const SceneDesc* d_scenes, uint32_t nscenes, uint32_t layout,
                    uint32_t threshold, sq_node* d_nodes, void* ws, uint32_t* d_depth_out);
// Batched Morton pipeline over a concatenation of scenes (bounds, codes,
// sort by (scene, code), emission, refit), enqueued on `stream` with no host
// synchronisation. Buffers: keys 8 B x total x 2, vals 4 B x total x 2,
// hist sort_hist_bytes(total, digit), ws lbvh_batch_workspace(...), ord
// 6 x nscenes words, nodes 2 x total, depth nscenes words.
struct BatchBuild {
    const sq_triangle* tris;
    uint32_t total, nscenes;
    const uint32_t* d_off;  // nscenes + 1 triangle offsets
    const SceneDesc* d_desc;
    uint32_t* d_ord;
    void* keys[2];
    uint32_t* vals[2];
    uint32_t* hist;
    void* ws;
    uint32_t* d_depth;
    sq_node* nodes;
    uint32_t bits, digit, layout, threshold;
    // non-null: the key passes run as single-pass look-back kernels in this
    // area (sort_onesweep_bytes), a failed wait sets *os_err (pinned host word)
    uint32_t* os = nullptr;
    uint32_t* os_err = nullptr;
};
int enqueue_batch_build(sq_ctx* c, hipStream_t stream, const BatchBuild& j);
// Upper bound of a Morton tree's depth: at most 3*bits splits on a code bit,
// then midpoint splits of an all-equal range, a fan level and the root.
uint32_t morton_depth_bound(uint32_t bits, uint32_t n);
int launch_fill_u32_on(hipStream_t stream, int num_cus, void* dst, uint32_t value, size_t words);
// Host memcpy split over up to 16 threads for large copies.
void par_memcpy(void* dst, const void* src, size_t bytes);
// Binned SAH over the scenes of a batch (see sah.hip); eager (waits per level).
int sah_build_batch(sq_ctx* c, const sq_triangle* tris, uint32_t total, const uint32_t* d_off, uint32_t nscenes,
                    uint32_t nb, uint32_t layout, sq_node* d_nodes, uint32_t* depth);
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

// A workgroup barrier that orders LDS accesses only: unlike __syncthreads it
// does not wait for the thread's global loads and stores to complete, so
// loads issued before it stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders one wave's LDS accesses around it (the wave's LDS operations
// execute in program order; this keeps the compiler from moving any across).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

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
__device__ __forceinline__ void store_zero_node(sq_node* nodes, uint32_t i) {
    float4* p = (float4*)(nodes + i);
    p[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    p[1] = make_float4(0.f, 0.f, 0.f, 0.f);
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

// Spread the low 10 bits of x to bits 0,3,...,27.
__device__ __forceinline__ uint32_t expand10(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
// Spread the low 21 bits of x to bits 0,3,...,60.
__device__ __forceinline__ uint64_t expand21(uint64_t x) {
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Truncating float -> unsigned as the reference's x86-64 build does for the
// values the quantiser produces; NaN (zero-extent axis) -> 0.
__device__ __forceinline__ uint32_t quantise(float v) {
    if (!(v == v) || v <= 0.0f) return 0u;
    return (uint32_t)v;
}

// setup_morton's code of one triangle: box centre normalised to the scene
// box (smn, ext = max - min) and truncated to `scale` = 2^bits - 1.
template <bool KEY64>
__device__ __forceinline__ uint64_t morton_code(const float v[9], const float smn[3], const float ext[3], float scale) {
    Box3 b = tri_box(v);
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float cc = (b.mn[k] + b.mx[k]) * 0.5f;
        cc = (cc - smn[k]) / ext[k];
        q[k] = quantise(cc * scale);
    }
    if (KEY64) return expand21(q[0]) | (expand21(q[1]) << 1) | (expand21(q[2]) << 2);
    return expand10(q[0]) | (expand10(q[1]) << 1) | (expand10(q[2]) << 2);
}

}  // namespace sqd
