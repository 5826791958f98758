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
// GPU structure: tasks are processed level by level. A task of up to
// kBigTask primitives gets one workgroup; a larger one is cut into chunks,
// one workgroup each, whose bins are folded by a second kernel before the
// plane is chosen, and whose partition ranks start at the counts of earlier
// chunks. Ranges of at most sub_max primitives finish in one wave's LDS.
// Bins are reduced with LDS atomics on order-preserving integer images of the
// floats (min/max are exact and order-free; of +0 and -0, -0 is the smaller),
// the plane is chosen by lane-parallel prefix/suffix unions over the bins, and
// the partition is done out of place between two id buffers, so both paths
// give the same tree.
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
// Tasks of more than kBigTask primitives are split by many workgroups, one per
// chunk of at least SahArgs::chunk_min primitives (512..kChunkMin, at most
// kMaxChunks chunks a task).
constexpr uint32_t kBigTask = 2048;
constexpr uint32_t kChunkMin = 2048;
constexpr uint32_t kMaxChunks = 512;
constexpr int kBigShift = 40;  // large-queue counter: (tasks << 40) | chunks
// Tasks of at most kSubtree primitives are queued once and finished, their
// whole subtree, by one wave in LDS after the level loop.
constexpr uint32_t kSubtree = 256;
// counter words posted to the host's level mailbox (the sequence number
// follows them), and levels enqueued ahead per batch once only small
// tasks are left
constexpr uint32_t kMailWords = 16;
constexpr uint32_t kAheadLevels = 4;

struct SahTask {
    uint32_t b, e, free, self;
    float cb[6];               // centroid box: min xyz, max xyz
    uint32_t chunk0, nchunks;  // large tasks: first chunk of the level, chunks;
                               // subtree tasks: id buffer holding the range, level
};

// Split chosen for a task: child bounds and centroid boxes, plane, left count.
struct SahPlan {
    float lp[6], rp[6], lc[6], rc[6];
    uint32_t plane, m, pad[2];
};

struct SahArgs {
    const float4* boxes;  // 2 per primitive: (min, -), (max, -)
    sq_node* nodes;
    uint32_t cap;
    uint32_t nb;
    uint32_t part;     // 0 Hoare, 1 stable
    uint32_t sa_x2;    // 1: AABB1::sa (x2)
    uint32_t layout;
    uint32_t small_cap, big_cap;
    uint32_t chunk_min;     // primitives per chunk of a large task, at least
    uint32_t sub_max;       // tasks of at most this many primitives (<= kSubtree) go to the subtree queue
    uint32_t nbig;          // large tasks of this level
    const SahTask* tasks;   // small tasks of this level
    SahTask* next;          // small tasks of the next level
    const SahTask* big;     // large tasks of this level
    SahTask* next_big;      // large tasks of the next level
    SahTask* sub;           // subtree tasks (all levels)
    uint32_t sub_cap;
    uint32_t out_buf;       // id buffer (0/1) that holds the ranges of queued children
    uint32_t child_level;   // level of queued children
    const uint32_t* ids_buf[2];
    uint32_t* counters;     // [0] next small count, [2..3] next large (packed), [4] error flag,
                            // [5] 1 + deepest level split by a subtree or look-ahead small-task
                            // wave, [6] subtree tasks, [8..10] look-ahead small counts
    uint32_t* small_ctr;    // counter of the next level's small tasks (counters + 0, or 8..10)
    // look-ahead levels (small tasks only, enqueued before the host has seen
    // their counts): the word holding this level's task count, and the word
    // to clear for the level after next; null otherwise
    const uint32_t* small_n;
    uint32_t* small_zero;
    const uint32_t* ids_in;
    uint32_t* ids_out;
    uint32_t* scratch;      // Hoare misplaced positions
    SahPlan* plans;         // per large task
    uint32_t* part_bins;    // per chunk: nb * kBinWords words
    uint32_t* chunk_cnt;    // per chunk: 2 words
    uint8_t* bin_of;        // bin of the primitive at each position of a large task
    // batched builds: the scenes' first primitives (nseg + 1 ascending
    // offsets). Scene s owns positions [off[s], off[s+1]) and node slots
    // [2 off[s], 2 off[s+1]); its tree is built with batch-wide slots and
    // primitive ids, and stored child links are made scene-local as they are
    // written (leaf ids minus off[s], child slots minus 2 off[s]).
    const uint32_t* seg_off;
    uint32_t nseg;
};

typedef uint32_t BinsLds[kSahWaves][kMaxBins][kBinWords];

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

// v[k] for k in 0..2 by selects: an index known only at run time into a
// small array would put the array in scratch memory.
__device__ __forceinline__ float sel3(const float* v, int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

__device__ __forceinline__ float bin_k1(uint32_t nb, const float* cb, int axis) {
    const float epsilon = 1.1920929e-7f;
    return (float)nb * (1 - epsilon) / (sel3(cb + 3, axis) - sel3(cb, axis));
}

__device__ __forceinline__ void prim_box(const float4* boxes, uint32_t id, float mn[3], float mx[3]) {
    float4 a = boxes[2 * id], b = boxes[2 * id + 1];
    mn[0] = a.x; mn[1] = a.y; mn[2] = a.z;
    mx[0] = b.x; mx[1] = b.y; mx[2] = b.z;
}

// The scene of position t of a batch: off[s] <= t < off[s + 1].
__device__ __forceinline__ uint32_t seg_of_pos(const uint32_t* off, uint32_t nseg, uint32_t t) {
    uint32_t lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// First primitive of the scene that holds position t (0 outside batches).
__device__ __forceinline__ uint32_t scene_base(const SahArgs& a, uint32_t t) {
    return a.seg_off ? a.seg_off[seg_of_pos(a.seg_off, a.nseg, t)] : 0u;
}

// pb: the first primitive of the node's scene in a batched build (0 for one
// scene): leaf ids and child slots are stored scene-local (id - pb, slot -
// 2 pb) while nodes are written at their batch-wide slots.
__device__ __forceinline__ void write_leaf_box(const SahArgs& a, uint32_t slot, uint32_t id, uint32_t pb) {
    if (slot >= a.cap) return;
    float mn[3], mx[3];
    prim_box(a.boxes, id, mn, mx);
    sqd::Box3 bx{{mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}};
    sqd::store_node(a.nodes, slot, bx, id - pb, 1, 1);
}

__device__ __forceinline__ void write_inner_box(const SahArgs& a, uint32_t slot, const float* mn, const float* mx,
                                                uint32_t child, uint32_t count, uint32_t pb) {
    if (slot >= a.cap) return;
    sqd::Box3 bx{{mn[0], mn[1], mn[2]}, {mx[0], mx[1], mx[2]}};
    sqd::store_node(a.nodes, slot, bx, child - 2 * pb, count, 0);
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

// Sum over the block (every thread gets the total).
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* s_tmp) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < kSahWaves; w++) t += s_tmp[w];
    __syncthreads();
    return t;
}

// ---- bins: ordered-integer images of the floats, so min/max/sum are exact
// and independent of the order primitives are added in ----
__device__ __forceinline__ uint32_t bin_identity(uint32_t f) {
    return f == 0 ? 0u : ((f - 1) % 6 < 3 ? sqd::f2ord(3.402823466e+38f) : sqd::f2ord(-3.402823466e+38f));
}

__device__ __forceinline__ uint32_t bin_combine(uint32_t f, uint32_t v, uint32_t x) {
    if (f == 0) return v + x;
    if ((f - 1) % 6 < 3) return x < v ? x : v;
    return x > v ? x : v;
}

// Inclusive scan over lanes 0 .. nb-1 (lanes at or above nb hold the
// identity and may end with anything) with DPP: row shifts by 1, 2, 4, 8
// within each 16-lane row, then the row broadcasts of lanes 15 and 31 (the
// GFX9 scan sequence), each only while the bins reach that far.
// Lanes without a source lane combine with the operation's own identity
// (0 for sums and maxima, all ones for minima of the unsigned images), which
// leaves their value unchanged and lets the compiler fuse the lane move into
// the min / max / add.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_step(uint32_t v, uint32_t f) {
    const int id = (f != 0 && (f - 1) % 6 < 3) ? -1 : 0;
    const uint32_t y = (uint32_t)__builtin_amdgcn_update_dpp(id, (int)v, CTRL, ROWS, 0xF, false);
    return bin_combine(f, v, y);
}

// NBC: the bin count rounded up to a power of two (the steps it needs are
// fixed at compile time: no branch between the steps). Both arrays (13 bin
// words each) take each step together, so consecutive DPP reads never wait
// on the instruction just before them.
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_step_all(uint32_t* a, uint32_t* b) {
#pragma unroll
    for (uint32_t f = 0; f < (uint32_t)kBinWords; f++) {
        a[f] = dpp_step<CTRL, ROWS>(a[f], f);
        b[f] = dpp_step<CTRL, ROWS>(b[f], f);
    }
}

template <int NBC>
__device__ __forceinline__ void wave_scan_bins2(uint32_t* a, uint32_t* b) {
    dpp_step_all<0x111, 0xF>(a, b);
    if constexpr (NBC > 2) dpp_step_all<0x112, 0xF>(a, b);
    if constexpr (NBC > 4) dpp_step_all<0x114, 0xF>(a, b);
    if constexpr (NBC > 8) dpp_step_all<0x118, 0xF>(a, b);
    if constexpr (NBC > 16) dpp_step_all<0x142, 0xA>(a, b);
    if constexpr (NBC > 32) dpp_step_all<0x143, 0xC>(a, b);
}

__device__ __forceinline__ uint32_t lane_of(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Full-wave inclusive scan (every lane active); returns lane 63's total.
__device__ __forceinline__ uint32_t wave_total_bin(uint32_t v, uint32_t f) {
    v = dpp_step<0x111, 0xF>(v, f);
    v = dpp_step<0x112, 0xF>(v, f);
    v = dpp_step<0x114, 0xF>(v, f);
    v = dpp_step<0x118, 0xF>(v, f);
    v = dpp_step<0x142, 0xA>(v, f);
    v = dpp_step<0x143, 0xC>(v, f);
    return lane_of(v, 63);
}

// Adds each lane's primitive (act: the lane has one; bin bi; w: its 13 bin
// words, count 1 first) to the bins B. Called by every lane of the wave.
// Coherent primitives share bins, and LDS atomics on one address serialise:
// so, up to three times, the lanes that share the bin of the first lane not
// yet added are combined by a wave reduction and added by one lane (when they
// are 8 or more); the others use LDS atomics. The bins' operations (sum,
// min, max of ordered-integer images) make the result independent of the
// order and grouping.
__device__ __forceinline__ void bin_add(uint32_t (*B)[kBinWords], bool act, uint32_t bi, const uint32_t* w) {
    const uint32_t lane = __lane_id();
    uint64_t rem = __ballot(act);
    bool done = false;
    for (int it = 0; it < 3 && rem; it++) {
        const uint32_t first = (uint32_t)__ffsll((long long)rem) - 1u;
        const uint32_t lead = lane_of(bi, first);
        const bool in = act && !done && bi == lead;
        const uint64_t m = __ballot(in);
        rem &= ~m;
        if (__popcll(m) < 8) continue;
        uint32_t tot[kBinWords];
#pragma unroll
        for (uint32_t f = 0; f < (uint32_t)kBinWords; f++) tot[f] = wave_total_bin(in ? w[f] : bin_identity(f), f);
        if (lane == first) {
            uint32_t* W = B[lead];
            atomicAdd(&W[0], tot[0]);
#pragma unroll
            for (int k = 0; k < 3; k++) {
                atomicMin(&W[1 + k], tot[1 + k]);
                atomicMax(&W[4 + k], tot[4 + k]);
                atomicMin(&W[7 + k], tot[7 + k]);
                atomicMax(&W[10 + k], tot[10 + k]);
            }
        }
        done = done || in;
    }
    if (act && !done) {
        uint32_t* W = B[bi];
        atomicAdd(&W[0], 1u);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            atomicMin(&W[1 + k], w[1 + k]);
            atomicMax(&W[4 + k], w[4 + k]);
            atomicMin(&W[7 + k], w[7 + k]);
            atomicMax(&W[10 + k], w[10 + k]);
        }
    }
}

// The 13 bin words of one primitive: count 1, box min, box max, centroid
// (as min and as max), ordered-integer images.
__device__ __forceinline__ void prim_words(const float* mn, const float* mx, const float* c, uint32_t* w) {
    w[0] = 1u;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        w[1 + k] = sqd::f2ord(mn[k]);
        w[4 + k] = sqd::f2ord(mx[k]);
        w[7 + k] = sqd::f2ord(c[k]);
        w[10 + k] = sqd::f2ord(c[k]);
    }
}

__device__ __forceinline__ void bins_clear(BinsLds& s, uint32_t nb) {
    for (uint32_t i = threadIdx.x; i < (uint32_t)kSahWaves * nb * kBinWords; i += kSahThreads) {
        uint32_t w = i / (nb * kBinWords), r = i % (nb * kBinWords);
        s[w][r / kBinWords][r % kBinWords] = bin_identity(r % kBinWords);
    }
}

// Bins the positions [lo, hi) of ids_in into the block's per-wave bins;
// with `keep`, records each position's bin in bin_of.
__device__ __forceinline__ void bin_range(const SahArgs& a, uint32_t lo, uint32_t hi, int axis, float k1, float cmin,
                                          BinsLds& s, bool keep) {
    const int wave = threadIdx.x >> 6;
    const uint32_t nb = a.nb;
    // every thread of the block runs the same number of iterations (bin_add
    // needs whole waves)
    for (uint32_t base = lo; base < hi; base += 4 * kSahThreads) {
        uint32_t idv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = base + u * kSahThreads + threadIdx.x;
            idv[u] = a.ids_in[i < hi ? i : lo];
        }
        float mn[4][3], mx[4][3];
#pragma unroll
        for (int u = 0; u < 4; u++) prim_box(a.boxes, idv[u], mn[u], mx[u]);
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = base + u * kSahThreads + threadIdx.x;
            const bool act = i < hi;
            float c[3];
            uint32_t w[kBinWords];
#pragma unroll
            for (int k = 0; k < 3; k++) c[k] = (mn[u][k] + mx[u][k]) * 0.5f;
            int b = (int)(k1 * (sel3(c, axis) - cmin));
            const uint32_t bi = (uint32_t)(b < 0 ? 0 : (b >= (int)nb ? (int)nb - 1 : b));
            if (act && keep) a.bin_of[i] = (uint8_t)bi;
            prim_words(mn[u], mx[u], c, w);
            bin_add(s[wave], act, bi, w);
        }
    }
}

// Folds the per-wave bins into s[0].
__device__ __forceinline__ void bins_merge_waves(BinsLds& s, uint32_t nb) {
    for (uint32_t i = threadIdx.x; i < nb * kBinWords; i += kSahThreads) {
        uint32_t bi = i / kBinWords, f = i % kBinWords;
        uint32_t v = s[0][bi][f];
        for (int w = 1; w < kSahWaves; w++) v = bin_combine(f, v, s[w][bi][f]);
        s[0][bi][f] = v;
    }
}

// Queues a child task: above kBigTask primitives on the large queue with its
// chunk range (one packed atomic hands out the queue slot and the chunks, so
// chunk ranges ascend with queue slots), above sub_max on the small queue,
// else on the subtree queue with the buffer holding its ids and its level.
__device__ void push_task(const SahArgs& a, SahTask c) {
    const uint32_t cnt = c.e - c.b;
    if (cnt > kBigTask) {
        uint32_t nch = (cnt + a.chunk_min - 1) / a.chunk_min;
        nch = nch < kMaxChunks ? nch : kMaxChunks;
        const unsigned long long old =
            atomicAdd((unsigned long long*)(a.counters + 2), (1ull << kBigShift) | (unsigned long long)nch);
        const uint32_t idx = (uint32_t)(old >> kBigShift);
        if (idx >= a.big_cap) {
            atomicOr(&a.counters[4], 1u);
            return;
        }
        c.chunk0 = (uint32_t)(old & ((1ull << kBigShift) - 1));
        c.nchunks = nch;
        a.next_big[idx] = c;
    } else if (cnt > a.sub_max) {
        const uint32_t idx = atomicAdd(a.small_ctr, 1u);
        if (idx >= a.small_cap) {
            atomicOr(&a.counters[4], 1u);
            return;
        }
        c.chunk0 = 0;
        c.nchunks = 0;
        a.next[idx] = c;
    } else {
        const uint32_t idx = atomicAdd(&a.counters[6], 1u);
        if (idx >= a.sub_cap) {
            atomicOr(&a.counters[4], 1u);
            return;
        }
        c.chunk0 = a.out_buf;
        c.nchunks = a.child_level;
        a.sub[idx] = c;
    }
}

// Children of a split task (recurse(), builder.cpp:1112-1155): slots, leaf
// boxes from ids_out, interior nodes and child tasks, by `nthr` threads of
// which `tid` is this one; ids_out of the task's range must be complete and
// visible.
__device__ void emit_children(const SahArgs& a, const SahTask& t, const SahPlan& P, uint32_t tid, uint32_t nthr) {
    const uint32_t b = t.b, e = t.e, count = e - b;
    const uint32_t L = P.m, R = count - P.m, f = t.free;
    uint32_t left_index = f, fi = f + (L != 1);
    uint32_t right_index = fi;
    fi += (R != 1);
    uint32_t left_free = fi, right_free = fi + (L - 1) * 2 + (L == 1);
    if (a.layout == SQ_LAYOUT_FIXED && R == 1 && L != 1) {
        right_index = f + 1;
        left_free = f + 2;
        right_free = f + 1;
    }
    const bool left_leaf = L <= 8 || box_sa(P.lc, P.lc + 3, a.sa_x2) <= 0.0f;
    const bool right_leaf = R <= 8 || box_sa(P.rc, P.rc + 3, a.sa_x2) <= 0.0f;
    const uint32_t pb = scene_base(a, b);
    if (left_leaf)
        for (uint32_t i = tid; i < L; i += nthr) write_leaf_box(a, left_free + i, a.ids_out[b + i], pb);
    if (right_leaf)
        for (uint32_t i = tid; i < R; i += nthr) write_leaf_box(a, right_free + i, a.ids_out[b + L + i], pb);
    if (tid == 0) {
        if (L != 1) write_inner_box(a, left_index, P.lp, P.lp + 3, left_free, left_leaf ? L : 2, pb);
        if (R != 1) write_inner_box(a, right_index, P.rp, P.rp + 3, right_free, right_leaf ? R : 2, pb);
        if (!left_leaf) {
            SahTask c;
            c.b = b; c.e = b + L; c.free = left_free; c.self = left_index;
            for (int k = 0; k < 6; k++) c.cb[k] = P.lc[k];
            push_task(a, c);
        }
        if (!right_leaf) {
            SahTask c;
            c.b = b + L; c.e = e; c.free = right_free; c.self = right_index;
            for (int k = 0; k < 6; k++) c.cb[k] = P.rc[k];
            push_task(a, c);
        }
    }
}

// ---- plane choice by one wave (small tasks, large-task folds) ----
// Lane-parallel choose_plane: lane i holds bin i; inclusive prefix unions give
// the left side of plane i, and inclusive prefix unions over the bins in
// reverse order (lane i holding bin nb - 1 - i) give the right side of plane
// i at lane nb - 2 - i (exact on the ordered-integer images). The sequential
// scan from plane nb-2 down to 0 with a strict < keeps the smallest score
// and, among equal scores, the highest plane; scores that are NaN or not
// below FLT_MAX never win, and when none wins plane 0 is taken with the same
// boxes as the scan's fallback.
template <int NBC>
__device__ __forceinline__ void choose_plane_wave_t(const uint32_t (*bins)[kBinWords], uint32_t nb, uint32_t sa_x2,
                                                    uint32_t lane, SahPlan& P) {
    uint32_t pre[kBinWords], suf[kBinWords];
#pragma unroll
    for (uint32_t f = 0; f < (uint32_t)kBinWords; f++) {
        pre[f] = lane < nb ? bins[lane][f] : bin_identity(f);
        suf[f] = lane < nb ? bins[nb - 1 - lane][f] : bin_identity(f);
    }
    wave_scan_bins2<NBC>(pre, suf);
    const bool valid_lane = lane + 2 <= nb;
    const uint32_t src = valid_lane ? nb - 2 - lane : lane;
    uint32_t rgt[7];
#pragma unroll
    for (int f = 0; f < 7; f++) rgt[f] = (uint32_t)__shfl((int)suf[f], (int)src, 64);
    float lp[6], rp[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        lp[k] = sqd::ord2f(pre[1 + k]);
        rp[k] = sqd::ord2f(rgt[1 + k]);
    }
    const float score = __fmaf_rn(box_sa(lp, lp + 3, sa_x2), (float)(int)pre[0],
                                  box_sa(rp, rp + 3, sa_x2) * (float)(int)rgt[0]);
    const bool valid = valid_lane && score < 3.402823466e+38f;
    float best = valid ? score : 3.402823466e+38f;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = fminf(best, __shfl_xor(best, off, 64));
    const uint64_t win = __ballot(valid && score == best);
    const uint32_t plane = win ? 63u - (uint32_t)__clzll((long long)win) : 0u;
    const uint32_t rl = nb - 2 - plane;
    P.plane = plane;
    P.m = lane_of(pre[0], plane);
#pragma unroll
    for (int k = 0; k < 6; k++) {
        P.lp[k] = sqd::ord2f(lane_of(pre[1 + k], plane));
        P.lc[k] = sqd::ord2f(lane_of(pre[7 + k], plane));
        P.rp[k] = sqd::ord2f(lane_of(suf[1 + k], rl));
        P.rc[k] = sqd::ord2f(lane_of(suf[7 + k], rl));
    }
}

// choose_plane_wave for G = 64 / W groups of W (16 or 32) lanes at once,
// group g holding its bins in rows g W .. g W + nb - 1 (nb <= W): the scans
// and the minimum stay inside each group, whose plane is then read with
// per-lane shuffles. Lanes of groups without a range (gv false) compute
// nothing that is used.
template <int NBC>
__device__ __forceinline__ void choose_plane_groups_t(const uint32_t (*bins)[kBinWords], uint32_t nb, uint32_t sa_x2,
                                                      uint32_t W, bool gv, SahPlan& P) {
    const uint32_t lane = __lane_id(), g = lane / W, gl = lane % W, row = g * W;
    uint32_t pre[kBinWords], suf[kBinWords];
#pragma unroll
    for (uint32_t f = 0; f < (uint32_t)kBinWords; f++) {
        pre[f] = gl < nb ? bins[row + gl][f] : bin_identity(f);
        suf[f] = gl < nb ? bins[row + nb - 1 - gl][f] : bin_identity(f);
    }
    wave_scan_bins2<NBC>(pre, suf);
    const bool valid_lane = gv && gl + 2 <= nb;
    const uint32_t src = row + (valid_lane ? nb - 2 - gl : gl);
    uint32_t rgt[7];
#pragma unroll
    for (int f = 0; f < 7; f++) rgt[f] = (uint32_t)__shfl((int)suf[f], (int)src, 64);
    float lp[6], rp[6];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        lp[k] = sqd::ord2f(pre[1 + k]);
        rp[k] = sqd::ord2f(rgt[1 + k]);
    }
    const float score = __fmaf_rn(box_sa(lp, lp + 3, sa_x2), (float)(int)pre[0],
                                  box_sa(rp, rp + 3, sa_x2) * (float)(int)rgt[0]);
    const bool valid = valid_lane && score < 3.402823466e+38f;
    float best = valid ? score : 3.402823466e+38f;
    for (uint32_t off = W >> 1; off > 0; off >>= 1) best = fminf(best, __shfl_xor(best, (int)off, 64));
    const uint64_t win = (__ballot(valid && score == best) >> row) & (W == 64 ? ~0ull : (1ull << W) - 1ull);
    const uint32_t plane = win ? 63u - (uint32_t)__clzll((long long)win) : 0u;
    const uint32_t pl = row + plane, pr = row + (nb - 2 - plane);
    P.plane = plane;
    P.m = (uint32_t)__shfl((int)pre[0], (int)pl, 64);
#pragma unroll
    for (int k = 0; k < 6; k++) {
        P.lp[k] = sqd::ord2f((uint32_t)__shfl((int)pre[1 + k], (int)pl, 64));
        P.lc[k] = sqd::ord2f((uint32_t)__shfl((int)pre[7 + k], (int)pl, 64));
        P.rp[k] = sqd::ord2f((uint32_t)__shfl((int)suf[1 + k], (int)pr, 64));
        P.rc[k] = sqd::ord2f((uint32_t)__shfl((int)suf[7 + k], (int)pr, 64));
    }
}

// Dispatch on the bin count rounded up to a power of two (2 .. 64).
__device__ __forceinline__ void choose_plane_wave(const uint32_t (*bins)[kBinWords], uint32_t nb, uint32_t sa_x2,
                                                  uint32_t lane, SahPlan& P) {
    if (nb <= 2) choose_plane_wave_t<2>(bins, nb, sa_x2, lane, P);
    else if (nb <= 4) choose_plane_wave_t<4>(bins, nb, sa_x2, lane, P);
    else if (nb <= 8) choose_plane_wave_t<8>(bins, nb, sa_x2, lane, P);
    else if (nb <= 16) choose_plane_wave_t<16>(bins, nb, sa_x2, lane, P);
    else if (nb <= 32) choose_plane_wave_t<32>(bins, nb, sa_x2, lane, P);
    else choose_plane_wave_t<64>(bins, nb, sa_x2, lane, P);
}

// Groups hold at most W <= 32 bins.
__device__ __forceinline__ void choose_plane_groups(const uint32_t (*bins)[kBinWords], uint32_t nb, uint32_t sa_x2,
                                                    uint32_t W, bool gv, SahPlan& P) {
    if (nb <= 2) choose_plane_groups_t<2>(bins, nb, sa_x2, W, gv, P);
    else if (nb <= 4) choose_plane_groups_t<4>(bins, nb, sa_x2, W, gv, P);
    else if (nb <= 8) choose_plane_groups_t<8>(bins, nb, sa_x2, W, gv, P);
    else if (nb <= 16) choose_plane_groups_t<16>(bins, nb, sa_x2, W, gv, P);
    else choose_plane_groups_t<32>(bins, nb, sa_x2, W, gv, P);
}
// A level's small tasks (sub_max + 1 .. kBigTask primitives), one workgroup
// each: the four waves bin interleaved quarters of the range into their own
// LDS bins (bin_range, which records each position's bin), wave 0 chooses
// the plane on the folded bins, and the partition ranks positions
// block-wide in position order (stable: left ranks; Hoare: the r-th
// misplaced right position swaps with the r-th misplaced left position from
// the end). Each thread partitions the positions it binned. One wave per
// task measured slower at every scene size (a 2048-primitive range is then
// a chain of eight dependent binning steps on one wave). A look-ahead level
// reads its task count from the device (at most small_cap are queued),
// clears the counter of the level after next and records its depth.
__global__ __launch_bounds__(kSahThreads) void k_sah_group(SahArgs a, uint32_t ntasks) {
    __shared__ BinsLds s_bins;
    __shared__ SahPlan s_plan;
    __shared__ uint32_t s_w[2][2][kSahWaves];  // per-wave counts, double-buffered
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (a.small_n) {
        ntasks = *a.small_n;
        ntasks = ntasks < a.small_cap ? ntasks : a.small_cap;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            *a.small_zero = 0u;
            if (ntasks) atomicMax(&a.counters[5], a.child_level);
        }
    }
    if (blockIdx.x >= ntasks) return;  // whole workgroup
    const SahTask t = a.tasks[blockIdx.x];
    const uint32_t b = t.b, e = t.e;
    const int axis = choose_axis(t.cb);
    const float k1 = bin_k1(a.nb, t.cb, axis);
    bins_clear(s_bins, a.nb);
    __syncthreads();
    bin_range(a, b, e, axis, k1, sel3(t.cb, axis), s_bins, true);
    __syncthreads();
    bins_merge_waves(s_bins, a.nb);
    __syncthreads();
    if (threadIdx.x < 64) {
        SahPlan P;
        choose_plane_wave(s_bins[0], a.nb, a.sa_x2, threadIdx.x, P);
        if (threadIdx.x == 0) s_plan = P;
    }
    __syncthreads();
    const SahPlan P = s_plan;
    const uint32_t plane = P.plane, m = P.m;
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t run0 = 0, run1 = 0, it = 0;
    for (uint32_t base = b; base < e; base += 4 * kSahThreads) {
        uint32_t idv[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint32_t i = base + u * kSahThreads + threadIdx.x, ic = i < e ? i : b;
            idv[u] = a.ids_in[ic];
            bv[u] = a.bin_of[ic];
        }
#pragma unroll
        for (int u = 0; u < 4; u++, it++) {
            const uint32_t i = base + u * kSahThreads + threadIdx.x;
            const bool valid = i < e;
            // stable: q0 = left; Hoare: q0 = misplaced right, q1 = misplaced left
            bool q0, q1 = false;
            if (a.part == 1) {
                q0 = valid && bv[u] <= plane;
            } else {
                const bool right = valid && bv[u] > plane;
                q0 = right && i < b + m;
                q1 = valid && !right && i >= b + m;
            }
            const uint64_t m0 = __ballot(q0), m1 = __ballot(q1);
            uint32_t(*W)[kSahWaves] = s_w[it & 1];
            if (lane == 0) {
                W[0][wave] = (uint32_t)__popcll(m0);
                W[1][wave] = (uint32_t)__popcll(m1);
            }
            __syncthreads();
            uint32_t pre0 = 0, pre1 = 0, tot0 = 0, tot1 = 0;
#pragma unroll
            for (int w = 0; w < kSahWaves; w++) {
                const uint32_t c0 = W[0][w], c1 = W[1][w];
                pre0 += (uint32_t)w < wave ? c0 : 0u;
                pre1 += (uint32_t)w < wave ? c1 : 0u;
                tot0 += c0;
                tot1 += c1;
            }
            if (a.part == 1) {
                if (valid) {
                    const uint32_t rank_l = run0 + pre0 + (uint32_t)__popcll(m0 & lt);
                    const uint32_t pos = q0 ? b + rank_l : b + m + ((i - b) - rank_l);
                    if (pos < e) a.ids_out[pos] = idv[u];
                }
            } else {
                if (q0) {
                    const uint32_t r = run0 + pre0 + (uint32_t)__popcll(m0 & lt);
                    if (b + r < e) a.scratch[b + r] = i;
                } else if (q1) {
                    const uint32_t r = run1 + pre1 + (uint32_t)__popcll(m1 & lt);
                    if (b + m + r < e) a.scratch[b + m + r] = i;
                } else if (valid) {
                    a.ids_out[i] = idv[u];
                }
            }
            run0 += tot0;
            run1 += tot1;
        }
    }
    if (a.part == 0) {
        __threadfence_block();
        __syncthreads();  // every misplaced position recorded
        const uint32_t k = run0;  // == run1
        for (uint32_t r = threadIdx.x; r < k; r += kSahThreads) {
            const uint32_t pr = a.scratch[b + r], pl = a.scratch[b + m + (k - 1 - r)];
            if (pr - b >= e - b || pl - b >= e - b) continue;
            a.ids_out[pr] = a.ids_in[pl];
            a.ids_out[pl] = a.ids_in[pr];
        }
    }
    __threadfence_block();
    __syncthreads();  // the range's ids, for the leaves
    emit_children(a, t, P, threadIdx.x, kSahThreads);
}

// First look-ahead level: its task count n in counters[8], the next two
// counters clear.
__global__ void k_sah_ahead_init(uint32_t* counters, uint32_t n) {
    if (threadIdx.x < 3) counters[8 + threadIdx.x] = threadIdx.x == 0 ? n : 0u;
}

// ---- subtree tasks: one wave builds the whole subtree of a range of at
// most kSubtree primitives in LDS ----
// The range's ids and boxes are loaded once; the splits permute a list of
// local slots (Hoare in place by the same pairing, stable through a staging
// list), so every split sees the order the level kernels would give it, and
// node slots follow from counts alone. The kernel never reads back what it
// stores to global memory, so its synchronisation orders LDS only (node
// stores stay in flight across splits). Pending ranges sit on an LDS stack
// (the larger child is pushed first); the ranges on it are disjoint and each
// holds more than 8 primitives, so it never holds more than kSubtree / 9.
// A step splits the range on top of the stack with the whole wave or, when
// it and the ranges below it are small, several at once: 2 ranges of at most
// 128 primitives on 32-lane groups (at most 32 bins), 4 of at most 64 on
// 16-lane groups (at most 16 bins), each lane taking up to four positions. Each group bins, chooses its plane (the
// DPP scans stay inside 16-lane rows and 32-lane halves) and partitions its
// own range; the splits are independent, so the tree is the same.
struct SubEntry {
    uint32_t lo, hi, free, self, level;
    float cb[6];
};
constexpr int kSubStackCap = (int)(kSubtree / 9) + 1;
// Ranges of up to kGroupSpan W primitives share a step on W-lane groups.
constexpr uint32_t kGroupSpan = 4;

__global__ __launch_bounds__(64) void k_sah_subtree(SahArgs a, uint32_t nsub) {
    __shared__ float s_box[kSubtree][6];
    __shared__ uint16_t s_ord[kSubtree];
    __shared__ uint16_t s_tmp[kSubtree];
    __shared__ uint8_t s_bin[kSubtree];
    __shared__ uint32_t s_bins[kMaxBins][kBinWords];
    __shared__ SubEntry s_stack[kSubStackCap];
    const uint32_t lane = threadIdx.x;
    if (blockIdx.x >= nsub) return;
    const SahTask T = a.sub[blockIdx.x];
    const uint32_t n = T.e - T.b, nb = a.nb;
    if (n > kSubtree) return;  // never queued (push_task), kept as a guard of the LDS arrays
    const uint32_t* ids = a.ids_buf[T.chunk0 & 1];
    const uint32_t pb = scene_base(a, T.b);  // the whole subtree lies in one scene
    {
        // all loads of the range first (at most kSubtree / 64 per lane)
        uint32_t idv[kSubtree / 64];
#pragma unroll
        for (uint32_t u = 0; u < kSubtree / 64; u++) {
            const uint32_t j = u * 64 + lane;
            idv[u] = ids[T.b + (j < n ? j : 0u)];
        }
        float mn[kSubtree / 64][3], mx[kSubtree / 64][3];
#pragma unroll
        for (uint32_t u = 0; u < kSubtree / 64; u++) prim_box(a.boxes, idv[u], mn[u], mx[u]);
#pragma unroll
        for (uint32_t u = 0; u < kSubtree / 64; u++) {
            const uint32_t j = u * 64 + lane;
            if (j >= n) continue;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                s_box[j][k] = mn[u][k];
                s_box[j][3 + k] = mx[u][k];
            }
            s_ord[j] = (uint16_t)j;
        }
    }
    if (lane == 0) {
        SubEntry E;
        E.lo = 0; E.hi = n; E.free = T.free; E.self = T.self; E.level = T.nchunks;
        for (int k = 0; k < 6; k++) E.cb[k] = T.cb[k];
        s_stack[0] = E;
    }
    uint32_t sp = 1, deepest = 0;
    sqd::wave_lds_sync();
    while (sp > 0) {
        // group width: the smallest of 16 / 32 / 64 lanes that holds the top
        // range and its bins; then as many ranges as groups, from the top
        const uint32_t top_cnt = s_stack[sp - 1].hi - s_stack[sp - 1].lo;
        const uint32_t W = (top_cnt <= kGroupSpan * 16 && nb <= 16) ? 16u : (top_cnt <= kGroupSpan * 32 && nb <= 32) ? 32u : 64u;
        uint32_t G = 1, iters = (top_cnt + W - 1) / W;
        while (G < 64 / W && G < sp) {
            const SubEntry& x = s_stack[sp - 1 - G];
            const uint32_t c = x.hi - x.lo;
            if (c > kGroupSpan * W) break;
            iters = (c + W - 1) / W > iters ? (c + W - 1) / W : iters;
            G++;
        }
        const uint32_t g = lane / W, gl = lane % W;
        const bool gv = g < G;
        const SubEntry E = s_stack[sp - 1 - (gv ? g : 0u)];
        sp -= G;
        deepest = gv && E.level + 1 > deepest ? E.level + 1 : deepest;
        const uint32_t lo = E.lo, hi = E.hi, cnt = hi - lo;
        const uint32_t row = g * W;  // this group's bins: rows row .. row + nb - 1
        const int axis = choose_axis(E.cb);
        const float k1 = bin_k1(nb, E.cb, axis);
        const float cmin = sel3(E.cb, axis);
        const uint64_t gmask = W == 64 ? ~0ull : ((1ull << W) - 1ull) << row;
        const uint64_t lt = ((1ull << lane) - 1ull) & gmask;
        if (gv)
            for (uint32_t i = gl; i < nb * kBinWords; i += W)
                s_bins[row + i / kBinWords][i % kBinWords] = bin_identity(i % kBinWords);
        sqd::wave_lds_sync();
        for (uint32_t it = 0; it < iters; it++) {
            const uint32_t j = lo + it * W + gl;
            const bool act = gv && it * W + gl < cnt;
            float mn[3], mx[3], c[3];
            uint32_t bi = 0, w[kBinWords];
            if (act) {
                const uint32_t sl = s_ord[j];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    mn[k] = s_box[sl][k];
                    mx[k] = s_box[sl][3 + k];
                    c[k] = (mn[k] + mx[k]) * 0.5f;
                }
                int bb = (int)(k1 * (sel3(c, axis) - cmin));
                bi = (uint32_t)(bb < 0 ? 0 : (bb >= (int)nb ? (int)nb - 1 : bb));
                s_bin[j] = (uint8_t)bi;
                prim_words(mn, mx, c, w);
            }
            bin_add(s_bins, act, row + bi, w);
        }
        sqd::wave_lds_sync();
        SahPlan P;
        if (W == 64) {
            choose_plane_wave(s_bins, nb, a.sa_x2, lane, P);
        } else {
            choose_plane_groups(s_bins, nb, a.sa_x2, W, gv, P);
        }
        const uint32_t plane = P.plane, m = P.m;
        if (a.part == 1) {
            uint32_t run_l = 0;
            for (uint32_t it = 0; it < iters; it++) {
                const uint32_t j = lo + it * W + gl;
                const bool valid = gv && it * W + gl < cnt;
                const bool isl = valid && s_bin[j] <= plane;
                const uint64_t ml = __ballot(isl) & gmask;
                if (valid) {
                    const uint32_t rank_l = run_l + (uint32_t)__popcll(ml & lt);
                    const uint32_t pos = isl ? rank_l : m + ((j - lo) - rank_l);
                    if (pos < cnt) s_tmp[lo + pos] = s_ord[j];
                }
                run_l += (uint32_t)__popcll(ml);
            }
            sqd::wave_lds_sync();
            if (gv)
                for (uint32_t j = lo + gl; j < hi; j += W) s_ord[j] = s_tmp[j];
        } else {
            uint32_t run_r = 0, run_l = 0;
            for (uint32_t it = 0; it < iters; it++) {
                const uint32_t j = lo + it * W + gl;
                const bool valid = gv && it * W + gl < cnt;
                const bool right = valid && s_bin[j] > plane;
                const bool mis_r = valid && right && j < lo + m;
                const bool mis_l = valid && !right && j >= lo + m;
                const uint64_t mr = __ballot(mis_r) & gmask, ml = __ballot(mis_l) & gmask;
                if (mis_r) {
                    const uint32_t r = run_r + (uint32_t)__popcll(mr & lt);
                    if (r < cnt) s_tmp[lo + r] = (uint16_t)j;
                } else if (mis_l) {
                    const uint32_t r = run_l + (uint32_t)__popcll(ml & lt);
                    if (m + r < cnt) s_tmp[lo + m + r] = (uint16_t)j;
                }
                run_r += (uint32_t)__popcll(mr);
                run_l += (uint32_t)__popcll(ml);
            }
            sqd::wave_lds_sync();
            const uint32_t k = gv ? run_r : 0u;  // == run_l
            for (uint32_t r = gl; r < k; r += W) {
                const uint32_t pr = s_tmp[lo + r], pl = s_tmp[lo + m + (k - 1 - r)];
                if (pr - lo >= cnt || pl - lo >= cnt) continue;
                const uint16_t x = s_ord[pr], y = s_ord[pl];
                s_ord[pr] = y;
                s_ord[pl] = x;
            }
        }
        sqd::wave_lds_sync();

        // children: the slots of emit_children
        const uint32_t L = m, R = cnt - m, f = E.free;
        uint32_t left_index = f, fi = f + (L != 1);
        uint32_t right_index = fi;
        fi += (R != 1);
        uint32_t left_free = fi, right_free = fi + (L - 1) * 2 + (L == 1);
        if (a.layout == SQ_LAYOUT_FIXED && R == 1 && L != 1) {
            right_index = f + 1;
            left_free = f + 2;
            right_free = f + 1;
        }
        const bool left_leaf = L <= 8 || box_sa(P.lc, P.lc + 3, a.sa_x2) <= 0.0f;
        const bool right_leaf = R <= 8 || box_sa(P.rc, P.rc + 3, a.sa_x2) <= 0.0f;
        if (gv)
            for (uint32_t i = gl; i < cnt; i += W) {
                const bool is_left = i < L;
                if (is_left ? !left_leaf : !right_leaf) continue;
                const uint32_t slot = is_left ? left_free + i : right_free + (i - L);
                if (slot >= a.cap) continue;
                const uint32_t sl = s_ord[lo + i];
                sqd::Box3 bx{{s_box[sl][0], s_box[sl][1], s_box[sl][2]}, {s_box[sl][3], s_box[sl][4], s_box[sl][5]}};
                // local slot sl is position T.b + sl of the id buffer (not kept in LDS)
                sqd::store_node(a.nodes, slot, bx, ids[T.b + sl] - pb, 1, 1);
            }
        if (gv && gl == 0) {
            if (L != 1) write_inner_box(a, left_index, P.lp, P.lp + 3, left_free, left_leaf ? L : 2, pb);
            if (R != 1) write_inner_box(a, right_index, P.rp, P.rp + 3, right_free, right_leaf ? R : 2, pb);
        }
        // children to the stack, groups in order, each group's larger child
        // first (entries built in registers and written to computed
        // positions: an array indexed by a running count would live in
        // scratch memory)
        const bool push_l = gv && !left_leaf, push_r = gv && !right_leaf;
        const uint32_t nc = (push_l ? 1u : 0u) + (push_r ? 1u : 0u);
        const uint64_t lead = __ballot(gv && gl == 0);
        const uint64_t one = __ballot(gv && gl == 0 && nc >= 1), two = __ballot(gv && gl == 0 && nc == 2);
        const uint64_t before = (1ull << row) - 1ull;
        const uint32_t base = sp + (uint32_t)__popcll(one & before) + (uint32_t)__popcll(two & before);
        const uint32_t total = (uint32_t)__popcll(one) + (uint32_t)__popcll(two);
        (void)lead;
        if (sp + total > (uint32_t)kSubStackCap) {  // cannot happen: disjoint ranges of more than 8
            if (lane == 0) atomicOr(&a.counters[4], 2u);
            break;
        }
        if (gv && gl == 0 && nc) {
            const bool left_first = L >= R;
            SubEntry cl, cr;
            cl.lo = lo; cl.hi = lo + L; cl.free = left_free; cl.self = left_index; cl.level = E.level + 1;
            cr.lo = lo + L; cr.hi = hi; cr.free = right_free; cr.self = right_index; cr.level = E.level + 1;
            for (int k = 0; k < 6; k++) {
                cl.cb[k] = P.lc[k];
                cr.cb[k] = P.rc[k];
            }
            const uint32_t pos_l = base + (left_first ? 0u : (push_r ? 1u : 0u));
            const uint32_t pos_r = base + (left_first ? (push_l ? 1u : 0u) : 0u);
            if (push_l) s_stack[pos_l] = cl;
            if (push_r) s_stack[pos_r] = cr;
        }
        sp += total;
        sqd::wave_lds_sync();
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = (uint32_t)__shfl_xor((int)deepest, off, 64);
        deepest = o > deepest ? o : deepest;
    }
    if (lane == 0) atomicMax(&a.counters[5], deepest);
}

// ---- large tasks: the same steps spread over one workgroup per chunk ----
// Chunks of a level are numbered task after task in queue order; a block
// finds its task by binary search over the tasks' first chunks.
__device__ __forceinline__ uint32_t big_task_of(const SahArgs& a, uint32_t chunk) {
    uint32_t lo = 0, hi = a.nbig;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.big[mid].chunk0 <= chunk) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void chunk_range(const SahTask& t, uint32_t g, uint32_t& lo, uint32_t& hi) {
    const uint64_t cnt = t.e - t.b, per = (cnt + t.nchunks - 1) / t.nchunks;
    const uint64_t l = (uint64_t)g * per, h = l + per;
    lo = t.b + (uint32_t)(l < cnt ? l : cnt);
    hi = t.b + (uint32_t)(h < cnt ? h : cnt);
}

// Bins one chunk; writes the chunk's bins and each position's bin.
__global__ __launch_bounds__(kSahThreads) void k_big_bin(SahArgs a) {
    __shared__ BinsLds s_bins;
    const uint32_t chunk = blockIdx.x;
    const SahTask t = a.big[big_task_of(a, chunk)];
    uint32_t lo, hi;
    chunk_range(t, chunk - t.chunk0, lo, hi);
    const int axis = choose_axis(t.cb);
    const float k1 = bin_k1(a.nb, t.cb, axis);
    bins_clear(s_bins, a.nb);
    __syncthreads();
    bin_range(a, lo, hi, axis, k1, sel3(t.cb, axis), s_bins, true);
    __syncthreads();
    bins_merge_waves(s_bins, a.nb);
    __syncthreads();
    const uint32_t words = a.nb * kBinWords;
    for (uint32_t w = threadIdx.x; w < words; w += kSahThreads)
        a.part_bins[(size_t)chunk * words + w] = s_bins[0][w / kBinWords][w % kBinWords];
}

// Folds a task's chunk bins and chooses its plane. The block's threads split
// the fold: thread (part p, word w) folds word w over chunks p, p + P, ...
// (P = threads / words parts; a chunk's words are adjacent, so a part's
// loads of one chunk are coalesced), sixteen chunks' words fetched before
// any is folded; the parts are folded in LDS. The root of a 1M-primitive
// scene has 488 chunks: 87 us with one thread per word and eight loads in
// flight. LDS atomics from every (chunk, word) pair would serialise on the
// task's few hundred words.
constexpr int kPlaneThreads = 1024;

__global__ __launch_bounds__(kPlaneThreads) void k_big_plane(SahArgs a) {
    __shared__ uint32_t s_red[kMaxBins][kBinWords];
    __shared__ uint32_t s_part[kPlaneThreads];
    const SahTask t = a.big[blockIdx.x];
    const uint32_t words = a.nb * kBinWords;  // at most kMaxBins * kBinWords = 832
    const uint32_t parts = kPlaneThreads / words;
    const uint32_t w = threadIdx.x % words, p = threadIdx.x / words;
    const uint32_t* src = a.part_bins + (size_t)t.chunk0 * words;
    if (p < parts) {
        const uint32_t f = w % kBinWords;
        const uint32_t op = f == 0 ? 0u : ((f - 1) % 6 < 3 ? 1u : 2u);  // sum, min, max
        uint32_t acc = bin_identity(f);
        uint32_t c = p;
        for (; c + 15 * parts < t.nchunks; c += 16 * parts) {
            uint32_t v[16];
#pragma unroll
            for (int k = 0; k < 16; k++) v[k] = src[(size_t)(c + k * parts) * words + w];
#pragma unroll
            for (int k = 0; k < 16; k++) acc = op == 0 ? acc + v[k] : (op == 1 ? min(acc, v[k]) : max(acc, v[k]));
        }
        for (; c < t.nchunks; c += parts) {
            const uint32_t v = src[(size_t)c * words + w];
            acc = op == 0 ? acc + v : (op == 1 ? min(acc, v) : max(acc, v));
        }
        s_part[p * words + w] = acc;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < words; i += kPlaneThreads) {
        const uint32_t f = i % kBinWords;
        uint32_t acc = s_part[i];
        for (uint32_t q = 1; q < parts; q++) acc = bin_combine(f, acc, s_part[q * words + i]);
        s_red[i / kBinWords][f] = acc;
    }
    __syncthreads();
    __shared__ SahPlan s_plan;
    __shared__ uint32_t s_sum[2][kPlaneThreads / 64];
    __shared__ uint32_t s_mid;
    if (threadIdx.x < 64) {
        SahPlan P;
        choose_plane_wave(s_red, a.nb, a.sa_x2, threadIdx.x, P);
        if (threadIdx.x == 0) {
            a.plans[blockIdx.x] = P;
            s_plan = P;
            s_mid = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    // Each chunk's partition counts for the scatter and swap kernels, from
    // its bins' counts: stable, its left count; Hoare, its misplaced right
    // (right-side bins before the midpoint) and misplaced left counts. The
    // one chunk that holds the midpoint counts its positions.
    const uint32_t plane = s_plan.plane, mid = t.b + s_plan.m;
    for (uint32_t g = threadIdx.x; g < t.nchunks; g += kPlaneThreads) {
        uint32_t lo, hi;
        chunk_range(t, g, lo, hi);
        const uint32_t* cb = src + (size_t)g * words;  // bin k's count at cb[k * kBinWords]
        uint32_t left = 0, all = 0;
        for (uint32_t k0 = 0; k0 < a.nb; k0 += 16) {
            uint32_t v[16];
#pragma unroll
            for (uint32_t k = 0; k < 16; k++) v[k] = k0 + k < a.nb ? cb[(k0 + k) * kBinWords] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < 16; k++) {
                all += v[k];
                left += k0 + k <= plane ? v[k] : 0u;
            }
        }
        uint32_t c0 = left, c1 = 0;
        if (a.part == 0) {
            if (hi <= mid) {
                c0 = all - left;
            } else if (lo >= mid) {
                c0 = 0;
                c1 = left;
            } else {
                s_mid = g;  // the one chunk with lo < mid < hi
                continue;
            }
        }
        a.chunk_cnt[2 * (t.chunk0 + g)] = c0;
        a.chunk_cnt[2 * (t.chunk0 + g) + 1] = c1;
    }
    __syncthreads();
    const uint32_t gm = s_mid;
    if (a.part == 0 && gm != 0xFFFFFFFFu) {
        uint32_t lo, hi;
        chunk_range(t, gm, lo, hi);
        uint32_t c0 = 0, c1 = 0;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += kPlaneThreads) {
            const bool right = a.bin_of[i] > plane;
            c0 += (right && i < mid) ? 1u : 0u;
            c1 += (!right && i >= mid) ? 1u : 0u;
        }
        for (int off = 32; off > 0; off >>= 1) {
            c0 += __shfl_xor(c0, off, 64);
            c1 += __shfl_xor(c1, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            s_sum[0][threadIdx.x >> 6] = c0;
            s_sum[1][threadIdx.x >> 6] = c1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t0 = 0, t1 = 0;
            for (int w = 0; w < kPlaneThreads / 64; w++) {
                t0 += s_sum[0][w];
                t1 += s_sum[1][w];
            }
            a.chunk_cnt[2 * (t.chunk0 + gm)] = t0;
            a.chunk_cnt[2 * (t.chunk0 + gm) + 1] = t1;
        }
    }
}

// Per chunk: the scatter of the single-workgroup partition, its running
// ranks starting at the counts of the task's earlier chunks.
__global__ __launch_bounds__(kSahThreads) void k_big_scatter(SahArgs a) {
    __shared__ uint32_t s_tmp[kSahWaves];
    const uint32_t chunk = blockIdx.x, ti = big_task_of(a, chunk);
    const SahTask t = a.big[ti];
    const uint32_t g = chunk - t.chunk0;
    const uint32_t plane = a.plans[ti].plane, m = a.plans[ti].m, b = t.b, e = t.e;
    uint32_t lo, hi;
    chunk_range(t, g, lo, hi);
    uint32_t p0 = 0, p1 = 0;
    for (uint32_t j = threadIdx.x; j < g; j += kSahThreads) {
        p0 += a.chunk_cnt[2 * (t.chunk0 + j)];
        p1 += a.chunk_cnt[2 * (t.chunk0 + j) + 1];
    }
    p0 = block_sum(p0, s_tmp);
    p1 = block_sum(p1, s_tmp);
    if (a.part == 1) {
        uint32_t run_l = p0;
        for (uint32_t base = lo; base < hi; base += kSahThreads) {
            const uint32_t i = base + threadIdx.x;
            const bool valid = i < hi;
            const uint32_t id = valid ? a.ids_in[i] : 0u;
            const bool right = valid && a.bin_of[i] > plane;
            uint32_t isl = (valid && !right) ? 1u : 0u, tot;
            const uint32_t pre = block_scan(isl, s_tmp, &tot);
            if (valid) {
                const uint32_t rank_l = run_l + pre;
                const uint32_t pos = isl ? b + rank_l : b + m + ((i - b) - rank_l);
                if (pos < e) a.ids_out[pos] = id;
            }
            run_l += tot;
        }
    } else {
        uint32_t run_r = p0, run_l = p1;
        for (uint32_t base = lo; base < hi; base += kSahThreads) {
            const uint32_t i = base + threadIdx.x;
            const bool valid = i < hi;
            const uint32_t id = valid ? a.ids_in[i] : 0u;
            const bool right = valid && a.bin_of[i] > plane;
            const bool mis_r = valid && right && i < b + m;
            const bool mis_l = valid && !right && i >= b + m;
            const uint32_t packed = (mis_r ? 1u : 0u) | (mis_l ? 0x10000u : 0u);
            uint32_t tot;
            const uint32_t pre = block_scan(packed, s_tmp, &tot);
            if (mis_r) {
                const uint32_t r = run_r + (pre & 0xFFFFu);
                if (b + r < e) a.scratch[b + r] = i;
            } else if (mis_l) {
                const uint32_t r = run_l + (pre >> 16);
                if (b + m + r < e) a.scratch[b + m + r] = i;
            } else if (valid) {
                a.ids_out[i] = id;
            }
            run_r += tot & 0xFFFFu;
            run_l += tot >> 16;
        }
    }
}

// Hoare: the r-th misplaced right element and the r-th misplaced left element
// from the end exchange places; chunk g takes pairs [g*per, (g+1)*per).
__global__ __launch_bounds__(kSahThreads) void k_big_swap(SahArgs a) {
    __shared__ uint32_t s_tmp[kSahWaves];
    const uint32_t chunk = blockIdx.x, ti = big_task_of(a, chunk);
    const SahTask t = a.big[ti];
    const uint32_t g = chunk - t.chunk0, b = t.b, m = a.plans[ti].m;
    uint32_t k = 0;
    for (uint32_t j = threadIdx.x; j < t.nchunks; j += kSahThreads) k += a.chunk_cnt[2 * (t.chunk0 + j)];
    k = block_sum(k, s_tmp);
    const uint32_t cnt = t.e - t.b, per = (uint32_t)(((uint64_t)cnt + t.nchunks - 1) / t.nchunks);
    const uint64_t r0 = (uint64_t)g * per, r1 = r0 + per < k ? r0 + per : k;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += kSahThreads) {
        const uint32_t pr = a.scratch[b + r], pl = a.scratch[b + m + (k - 1 - r)];
        if (pr - b >= cnt || pl - b >= cnt) continue;
        a.ids_out[pr] = a.ids_in[pl];
        a.ids_out[pl] = a.ids_in[pr];
    }
}

__global__ __launch_bounds__(kSahThreads) void k_big_emit(SahArgs a) {
    __shared__ SahPlan s_plan;
    const uint32_t* s = (const uint32_t*)&a.plans[blockIdx.x];
    uint32_t* d = (uint32_t*)&s_plan;
    for (uint32_t i = threadIdx.x; i < sizeof(SahPlan) / 4; i += kSahThreads) d[i] = s[i];
    __syncthreads();
    emit_children(a, a.big[blockIdx.x], s_plan, threadIdx.x, kSahThreads);
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
    __shared__ uint32_t s_red[12][4];
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; k++) {
            s_red[k][wave] = lo[k];
            s_red[3 + k][wave] = hi[k];
            s_red[6 + k][wave] = lo[3 + k];
            s_red[9 + k][wave] = hi[3 + k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int k = threadIdx.x;
        const bool is_min = (k / 3) % 2 == 0;
        uint32_t r = s_red[k][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            const uint32_t x = s_red[k][w];
            r = is_min ? (x < r ? x : r) : (x > r ? x : r);
        }
        if (is_min) atomicMin(&red[k], r);
        else atomicMax(&red[k], r);
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
__global__ void k_sah_root(SahArgs a, uint32_t n, const uint32_t* red) {
    float pmn[3], pmx[3], cb[6];
    decode_box(red, pmn, pmx);
    decode_box(red + 6, cb, cb + 3);
    const uint32_t ri = 2 * n - 1;
    if (n <= 8 || box_sa(cb, cb + 3, a.sa_x2) <= 0.0f) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) write_leaf_box(a, i, a.ids_in[i], 0);
        if (threadIdx.x == 0) write_inner_box(a, ri, pmn, pmx, 0, n, 0);
        return;
    }
    if (threadIdx.x == 0) {
        write_inner_box(a, ri, pmn, pmx, 0, 2, 0);
        SahTask t;
        t.b = 0; t.e = n; t.free = 0; t.self = ri;
        for (int k = 0; k < 6; k++) t.cb[k] = cb[k];
        push_task(a, t);
    }
}

// ---- batched binned SAH (recursive builder over many scenes) ----
// Per-primitive boxes and ids as k_sah_setup, and each scene's primitive and
// centroid boxes: one workgroup per 4096-primitive tile reduces its part of
// every scene it overlaps (red: 12 ordered-int words per scene).
constexpr uint32_t kSbTile = 4096;

__global__ void k_batch_sah_red_init(uint32_t* red, uint32_t nscenes, uint32_t* counters) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 12 * nscenes; i += gridDim.x * blockDim.x)
        red[i] = (((i % 12) / 3) % 2 == 0) ? 0xFFFFFFFFu : 0u;
    if (blockIdx.x == 0 && threadIdx.x < 8) counters[threadIdx.x] = 0;
}


__global__ __launch_bounds__(256) void k_batch_sah_setup(const sq_triangle* __restrict__ tris,
                                                         const uint32_t* __restrict__ off, uint32_t nscenes,
                                                         uint32_t total, float4* __restrict__ boxes,
                                                         uint32_t* __restrict__ ids, uint32_t* __restrict__ red) {
    __shared__ uint32_t s_red[12][4];
    const uint32_t t0 = blockIdx.x * kSbTile;
    const uint32_t t1 = (total - t0) < kSbTile ? total : t0 + kSbTile;
    const int wave = threadIdx.x >> 6;
    for (uint32_t s = seg_of_pos(off, nscenes, t0); s < nscenes && off[s] < t1; s++) {
        const uint32_t b = off[s] > t0 ? off[s] : t0, e = off[s + 1] < t1 ? off[s + 1] : t1;
        uint32_t lo[6], hi[6];
        for (int k = 0; k < 6; k++) { lo[k] = 0xFFFFFFFFu; hi[k] = 0u; }
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
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
        for (int o = 32; o > 0; o >>= 1)
            for (int k = 0; k < 6; k++) {
                uint32_t x = __shfl_xor(lo[k], o, 64), y = __shfl_xor(hi[k], o, 64);
                lo[k] = x < lo[k] ? x : lo[k];
                hi[k] = y > hi[k] ? y : hi[k];
            }
        if ((threadIdx.x & 63) == 0) {
            for (int k = 0; k < 3; k++) {
                s_red[k][wave] = lo[k];
                s_red[3 + k][wave] = hi[k];
                s_red[6 + k][wave] = lo[3 + k];
                s_red[9 + k][wave] = hi[3 + k];
            }
        }
        __syncthreads();
        if (threadIdx.x < 12) {
            const int k = threadIdx.x;
            const bool is_min = (k / 3) % 2 == 0;
            uint32_t r = s_red[k][0];
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
                const uint32_t x = s_red[k][w];
                r = is_min ? (x < r ? x : r) : (x > r ? x : r);
            }
            if (is_min) atomicMin(&red[12 * s + k], r);
            else atomicMax(&red[12 * s + k], r);
        }
        __syncthreads();
    }
}

// One workgroup per scene: the root of build_hierarchy (builder.cpp:1587-1612)
// with the scene's slots and positions, as k_sah_root does for one scene.
__global__ void k_batch_sah_root(SahArgs a, const uint32_t* red) {
    const uint32_t s = blockIdx.x, base = a.seg_off[s], n = a.seg_off[s + 1] - base;
    float pmn[3], pmx[3], cb[6];
    decode_box(red + 12 * s, pmn, pmx);
    decode_box(red + 12 * s + 6, cb, cb + 3);
    const uint32_t nb = 2 * base, ri = nb + 2 * n - 1;
    if (n <= 8 || box_sa(cb, cb + 3, a.sa_x2) <= 0.0f) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) write_leaf_box(a, nb + i, a.ids_in[base + i], base);
        if (threadIdx.x == 0) write_inner_box(a, ri, pmn, pmx, nb, n, base);
        return;
    }
    if (threadIdx.x == 0) {
        write_inner_box(a, ri, pmn, pmx, nb, 2, base);
        SahTask t;
        t.b = base; t.e = base + n; t.free = nb; t.self = ri;
        for (int k = 0; k < 6; k++) t.cb[k] = cb[k];
        push_task(a, t);
    }
}

// ---- grid builder ----
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 10; i++) r |= ((x >> i) & 1u) << (3 * i);
    return r;
}

// Cell of each primitive in Morton cell order (compact_1_by_2 inverse) and the
// per-cell bins (bin_centroids_to_grid_sse, builder.cpp:1755-1839). Bins are
// gathered in LDS per workgroup and merged into the global cell bins once per
// workgroup (exact: counts add, ordered-integer min/max commute).
constexpr uint32_t kMaxCells = 512;
__global__ __launch_bounds__(256) void k_grid_cells(const float4* __restrict__ boxes, uint32_t n, uint32_t dim,
                                                    const uint32_t* __restrict__ red, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals, uint32_t* __restrict__ cellbins) {
    __shared__ uint32_t s_cells[kMaxCells * kBinWords];
    const uint32_t cells = dim * dim * dim;
    for (uint32_t i = threadIdx.x; i < cells * kBinWords; i += blockDim.x) s_cells[i] = bin_identity(i % kBinWords);
    __syncthreads();
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
        uint32_t* C = s_cells + cell * kBinWords;
        atomicAdd(&C[0], 1u);
        for (int k = 0; k < 3; k++) {
            atomicMin(&C[1 + k], sqd::f2ord(mn[k]));
            atomicMax(&C[4 + k], sqd::f2ord(mx[k]));
            atomicMin(&C[7 + k], sqd::f2ord(c[k]));
            atomicMax(&C[10 + k], sqd::f2ord(c[k]));
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cells * kBinWords; i += blockDim.x) {
        const uint32_t f = i % kBinWords, v = s_cells[i];
        if (s_cells[i - f] == 0) continue;  // empty in this workgroup
        if (f == 0) atomicAdd(&cellbins[i], v);
        else if ((f - 1) % 6 < 3) atomicMin(&cellbins[i], v);
        else atomicMax(&cellbins[i], v);
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
                                                   const uint32_t* cellbins, const uint32_t* red) {
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
            for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) write_leaf_box(a, free + j, a.ids_in[begin + j], 0);
        if (threadIdx.x == 0) {
            const float* bmn = cell_tight ? p : pmn;
            const float* bmx = cell_tight ? p + 3 : pmx;
            write_inner_box(a, cells - 1 + i, bmn, bmx, free, leaf ? cnt : 2, 0);
            if (!leaf) {
                SahTask t;
                t.b = begin; t.e = begin + cnt; t.free = free; t.self = cells - 1 + i;
                for (int k = 0; k < 6; k++) t.cb[k] = cb[k];
                push_task(a, t);
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

// Queue and workspace capacities of a build over n primitives.
struct SahCaps {
    size_t qcap, bigcap, chunkcap;
};

// Work sizes for scenes of about `scene` primitives: a smaller scene has
// fewer ranges a level, so it takes smaller chunks and hands ranges to the
// subtree kernel later, to keep more workgroups busy. The tree does not
// depend on them. Measured (16 bins; 32 at 1M): 69,632 triangles 0.95 ms with
// 2048 / 256, 0.70-0.74 ms with these (512 / 64); 1M triangles 1.95 ms either
// way (these give 2048 / 256 there; 512 / 64 would take 2.18 ms).
static void sah_sizes(uint32_t scene, SahArgs& a) {
    uint32_t sub = 64, ch = 512;
    while (sub < kSubtree && sub * 2 <= scene / 1024) sub *= 2;
    while (ch < kChunkMin && ch * 2 <= scene / 128) ch *= 2;
    a.sub_max = sub;
    a.chunk_min = ch;
}

// Sizes the context's workspace for n primitives (cells: grid cells, 0 for
// the recursive builder) and points the arguments at it; a.chunk_min must be
// set (sah_sizes).
static int sah_prepare(sq_ctx* c, uint32_t n, uint32_t cells, uint32_t nb, SahArgs& a, SahCaps& k) {
    // A task that is split has at least 9 primitives (a large one more than
    // kBigTask), and the tasks of one level own disjoint ranges.
    k.qcap = n / 9 + 2 + cells;
    k.bigcap = n / kBigTask + 2 + cells;
    k.chunkcap = n / a.chunk_min + k.bigcap + 1;  // sum of ceil(count / chunk_min) over a level
    a.small_cap = (uint32_t)k.qcap;
    a.big_cap = (uint32_t)k.bigcap;
    a.sub_cap = (uint32_t)k.qcap;
    SQ_TRY(c->aux[3].ensure((size_t)n * 32));             // boxes
    SQ_TRY(c->vals[0].ensure((size_t)n * 4));              // ids ping
    SQ_TRY(c->vals[1].ensure((size_t)n * 4));              // ids pong
    SQ_TRY(c->keys[0].ensure((size_t)n * 4));
    SQ_TRY(c->keys[1].ensure((size_t)n * 4));
    SQ_TRY(c->scratch.ensure((size_t)n * 4 + 256));
    SQ_TRY(c->tasks[0].ensure(k.qcap * sizeof(SahTask)));
    SQ_TRY(c->tasks[1].ensure(k.qcap * sizeof(SahTask)));
    SQ_TRY(c->tasks[2].ensure(k.qcap * sizeof(SahTask)));
    SQ_TRY(c->sah[0].ensure(k.bigcap * sizeof(SahTask)));
    SQ_TRY(c->sah[1].ensure(k.bigcap * sizeof(SahTask)));
    SQ_TRY(c->sah[2].ensure(k.bigcap * sizeof(SahPlan)));
    SQ_TRY(c->sah[3].ensure(k.chunkcap * nb * kBinWords * 4));
    SQ_TRY(c->sah[4].ensure(k.chunkcap * 8));
    SQ_TRY(c->sah[5].ensure((size_t)n + 256));
    SQ_TRY(c->counters.ensure(64));
    SQ_TRY(c->aux[2].ensure((size_t)cells * kBinWords * 4 + 64));
    a.boxes = c->aux[3].as<float4>();
    a.counters = c->counters.as<uint32_t>();
    a.small_ctr = a.counters;
    a.small_n = nullptr;
    a.small_zero = nullptr;
    a.scratch = c->scratch.as<uint32_t>();
    a.plans = c->sah[2].as<SahPlan>();
    a.part_bins = c->sah[3].as<uint32_t>();
    a.chunk_cnt = c->sah[4].as<uint32_t>();
    a.bin_of = c->sah[5].as<uint8_t>();
    a.next = c->tasks[0].as<SahTask>();
    a.next_big = c->sah[0].as<SahTask>();
    a.sub = c->tasks[2].as<SahTask>();
    a.ids_buf[0] = c->vals[0].as<uint32_t>();
    a.ids_buf[1] = c->vals[1].as<uint32_t>();
    a.child_level = 0;
    return SQ_OK;
}

// The level loop over the queued tasks (ids in buffer `in`), then the
// LDS subtrees; *levels: split levels of the deepest path.
// Posts the level's counters to the context's pinned mailbox: words 0..7,
// then (after a system-scope fence) the sequence number in word 8; and
// zeroes the next level's queue counters (words 0, 2, 3) for the launches
// the host makes after reading them.
__global__ void k_sah_post(uint32_t* __restrict__ counters, uint32_t* mail, uint32_t seq) {
    const uint32_t v = threadIdx.x < kMailWords ? counters[threadIdx.x] : 0u;
    if (threadIdx.x < kMailWords) mail[threadIdx.x] = v;
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(mail + kMailWords, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    if (threadIdx.x == 0 || threadIdx.x == 2 || threadIdx.x == 3) counters[threadIdx.x] = 0;
}

// Waits for the mailbox to carry `seq` and copies its counters to h.
// The host polls the pinned word; every few hundred polls it asks whether
// the stream is idle, which ends the wait (with an error if the post never
// arrived) instead of spinning on a stream that has failed or finished.
static int wait_mail(sq_ctx* c, uint32_t seq, uint32_t* h) {
    for (uint32_t spin = 1;; spin++) {
        if (__atomic_load_n(&c->h_mail[kMailWords], __ATOMIC_ACQUIRE) == seq) break;
        if ((spin & 255u) == 0) {
            const hipError_t e = hipStreamQuery(c->stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(&c->h_mail[kMailWords], __ATOMIC_ACQUIRE) == seq) break;
                SQ_RET(SQ_ERR_HIP, "sah: a level's task counts were not posted");
            }
            if (e != hipErrorNotReady) SQ_CHECK_HIP(e);
        }
    }
    for (uint32_t i = 0; i < kMailWords; i++) h[i] = c->h_mail[i];
    return SQ_OK;
}

static int post_and_wait(sq_ctx* c, uint32_t* counters, uint32_t* h) {
    const uint32_t seq = ++c->mail_seq;
    hipLaunchKernelGGL(k_sah_post, dim3(1), dim3(64), 0, c->stream, counters, c->h_mail, seq);
    SQ_CHECK_HIP(hipGetLastError());
    return wait_mail(c, seq, h);
}

// Levels below the last large task hold small tasks only, so they are
// enqueued kAheadLevels at a time without the host seeing their counts:
// level j reads its count from counters[8 + j % 3], queues its children's
// into 8 + (j + 1) % 3 and clears 8 + (j + 2) % 3 (read by level j - 1,
// counted into by level j + 1, both ordered by the stream); a level's grid
// covers at most twice the tasks of the level before (its workgroups past
// the level's count leave at once). The host checks the counts after each
// batch. On entry `nsmall` tasks of level `level` are
// queued in c->tasks[q] with their ids in buffer `in`.
static int sah_ahead_levels(sq_ctx* c, SahArgs a, int in, int q, uint32_t level, uint32_t nsmall,
                            const SahCaps& k, uint32_t* h) {
    uint32_t* counters = a.counters;
    hipLaunchKernelGGL(k_sah_ahead_init, dim3(1), dim3(64), 0, c->stream, counters, nsmall);
    uint64_t bound = nsmall;
    for (uint32_t j = 0;; ) {
        for (uint32_t i = 0; i < kAheadLevels; i++, j++) {
            a.tasks = c->tasks[q].as<SahTask>();
            a.next = c->tasks[q ^ 1].as<SahTask>();
            a.ids_in = c->vals[in].as<uint32_t>();
            a.ids_out = c->vals[in ^ 1].as<uint32_t>();
            a.out_buf = (uint32_t)(in ^ 1);
            a.child_level = level + 1;
            a.small_n = counters + 8 + j % 3;
            a.small_ctr = counters + 8 + (j + 1) % 3;
            a.small_zero = counters + 8 + (j + 2) % 3;
            uint64_t groups = bound < k.qcap ? bound : k.qcap;
            if (groups < 1) groups = 1;
            hipLaunchKernelGGL(k_sah_group, dim3((uint32_t)groups), dim3(kSahThreads), 0, c->stream, a, 0u);
            bound = bound * 2 < k.qcap ? bound * 2 : k.qcap;
            q ^= 1;
            in ^= 1;
            level++;
        }
        SQ_CHECK_HIP(hipGetLastError());
        SQ_TRY(post_and_wait(c, counters, h));
        if (h[4]) SQ_RET(SQ_ERR_OVERFLOW, "sah: task queue overflow at level %u", level);
        if (h[8 + j % 3] == 0) break;  // the next level has no task
        if (level > 4 * 1024 * 1024) SQ_RET(SQ_ERR_OVERFLOW, "sah: runaway recursion");
    }
    return SQ_OK;
}

static int sah_levels(sq_ctx* c, SahArgs a, int in, const SahCaps& k, uint32_t* levels_out) {
    uint32_t* counters = a.counters;
    SQ_TRY(pinned_ensure(c, kMailWords * 4));
    uint32_t* h = (uint32_t*)c->pinned;
    if (!c->h_mail) SQ_RET(SQ_ERR_HIP, "sah: the context has no level mailbox");
    uint32_t level = 0;
    int q = 0;
    for (;;) {
        // the level's task counts, posted by the device to pinned memory
        SQ_TRY(post_and_wait(c, counters, h));
        if (h[4]) SQ_RET(SQ_ERR_OVERFLOW, "sah: task queue overflow at level %u", level);
        const uint32_t nsmall = h[0];
        const uint64_t packed = (uint64_t)h[2] | ((uint64_t)h[3] << 32);
        const uint32_t nbig = (uint32_t)(packed >> kBigShift);
        const uint64_t nchunks = packed & ((1ull << kBigShift) - 1);
        if (!nsmall && !nbig) break;
        if (nsmall > k.qcap || nbig > k.bigcap || nchunks > k.chunkcap)
            SQ_RET(SQ_ERR_OVERFLOW, "sah: level %u holds %u + %u tasks, %llu chunks", level, nsmall, nbig,
                   (unsigned long long)nchunks);
        if (!nbig) {
            SQ_TRY(sah_ahead_levels(c, a, in, q, level, nsmall, k, h));
            break;
        }
        a.tasks = c->tasks[q].as<SahTask>();
        a.next = c->tasks[q ^ 1].as<SahTask>();
        a.big = c->sah[q].as<SahTask>();
        a.next_big = c->sah[q ^ 1].as<SahTask>();
        a.nbig = nbig;
        a.ids_in = c->vals[in].as<uint32_t>();
        a.ids_out = c->vals[in ^ 1].as<uint32_t>();
        a.out_buf = (uint32_t)(in ^ 1);
        a.child_level = level + 1;
        if (nsmall) hipLaunchKernelGGL(k_sah_group, dim3(nsmall), dim3(kSahThreads), 0, c->stream, a, nsmall);
        if (nbig) {
            const dim3 gc((uint32_t)nchunks), gt(nbig), bt(kSahThreads);
            hipLaunchKernelGGL(k_big_bin, gc, bt, 0, c->stream, a);
            hipLaunchKernelGGL(k_big_plane, gt, dim3(kPlaneThreads), 0, c->stream, a);
            hipLaunchKernelGGL(k_big_scatter, gc, bt, 0, c->stream, a);
            if (a.part == 0) hipLaunchKernelGGL(k_big_swap, gc, bt, 0, c->stream, a);
            hipLaunchKernelGGL(k_big_emit, gt, bt, 0, c->stream, a);
        }
        SQ_CHECK_HIP(hipGetLastError());
        q ^= 1;
        in ^= 1;
        level++;
        if (level > 4 * 1024 * 1024) SQ_RET(SQ_ERR_OVERFLOW, "sah: runaway recursion");
    }
    const uint32_t nsub = h[6];
    // counters[5]: 1 + the deepest level the look-ahead levels split
    uint32_t levels = level > h[5] ? level : h[5];
    if (nsub) {
        if (nsub > k.qcap) SQ_RET(SQ_ERR_OVERFLOW, "sah: %u subtree tasks", nsub);
        hipLaunchKernelGGL(k_sah_subtree, dim3(nsub), dim3(64), 0, c->stream, a, nsub);
        SQ_CHECK_HIP(hipGetLastError());
        SQ_CHECK_HIP(hipMemcpyAsync(h, counters, 32, hipMemcpyDeviceToHost, c->stream));
        SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
        if (h[4]) SQ_RET(SQ_ERR_OVERFLOW, "sah: subtree stack overflow");
        levels = h[5] > levels ? h[5] : levels;
    }
    *levels_out = levels;
    return SQ_OK;
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
    SahCaps k;
    sah_sizes(n, a);
    SQ_TRY(sah_prepare(c, n, grid ? cells : 0u, nb, a, k));
    SQ_TRY(c->aux[0].ensure(64 * 4));
    uint32_t* red = c->aux[0].as<uint32_t>();
    uint32_t* counters = a.counters;
    float4* boxes = (float4*)a.boxes;

    phase_begin(c, grid ? PH_GRID : PH_SAH);
    hipLaunchKernelGGL(k_sah_red_init, dim3(1), dim3(64), 0, c->stream, red, counters);
    uint32_t blocks = (n + 255) / 256, capb = (uint32_t)c->num_cus * 8;
    if (blocks > capb) blocks = capb;
    hipLaunchKernelGGL(k_sah_setup, dim3(blocks), dim3(256), 0, c->stream, tris, n, boxes, c->vals[0].as<uint32_t>(), red);
    int in = 0;
    if (grid) {
        uint32_t* cellbins = c->aux[2].as<uint32_t>();
        hipLaunchKernelGGL(k_grid_bins_init, dim3(1), dim3(256), 0, c->stream, cellbins, cells);
        const uint32_t cblocks = blocks < (uint32_t)c->num_cus * 2 ? blocks : (uint32_t)c->num_cus * 2;
        hipLaunchKernelGGL(k_grid_cells, dim3(cblocks), dim3(256), 0, c->stream, boxes, n, dim, red,
                           c->keys[0].as<uint32_t>(), c->vals[0].as<uint32_t>(), cellbins);
        int slot = 0;
        uint32_t kbits = 0;
        while ((1u << kbits) < cells) kbits++;
        SQ_TRY(radix_sort(c, n, kbits ? kbits : 1, 8, false, 0, &slot));
        in = slot;
        a.ids_in = c->vals[in].as<uint32_t>();
        a.out_buf = (uint32_t)in;
        hipLaunchKernelGGL(k_grid_init, dim3(1), dim3(256), 0, c->stream, a, cells,
                           kind == SQ_BUILDER_GRID ? 0u : 1u, cellbins, red);
        *root = 0;
    } else {
        a.ids_in = c->vals[0].as<uint32_t>();
        a.out_buf = 0;
        hipLaunchKernelGGL(k_sah_root, dim3(1), dim3(256), 0, c->stream, a, n, red);
        *root = 2 * n - 1;
    }
    SQ_CHECK_HIP(hipGetLastError());
    uint32_t levels = 0;
    SQ_TRY(sah_levels(c, a, in, k, &levels));
    phase_end(c, grid ? PH_GRID : PH_SAH);
    *depth = levels + 2 + (grid ? 7 : 0);
    return SQ_OK;
}

// Binned SAH (the recursive builder, nb bins) over the scenes of a batch:
// scene s holds primitives [off[s], off[s+1]) of `tris` and its tree goes to
// node slots [2 off[s], 2 off[s+1]) of d_nodes, root at 2 off[s+1] - 1
// (scene-local 2n - 1), child slots and leaf ids scene-local. Every scene's
// tasks share the level launches. *depth bounds every scene's tree depth.
int sah_build_batch(sq_ctx* c, const sq_triangle* tris, uint32_t total, const uint32_t* d_off, uint32_t nscenes,
                    uint32_t nb, uint32_t layout, sq_node* d_nodes, uint32_t* depth) {
    if (nb < 2 || nb > (uint32_t)kMaxBins) SQ_RET(SQ_ERR_INVALID, "num_bins %u outside 2..%d", nb, kMaxBins);
    SahArgs a{};
    a.nodes = d_nodes;
    a.cap = 2 * total;
    a.nb = nb;
    a.part = 0;
    a.sa_x2 = 1;
    a.layout = layout;
    a.seg_off = d_off;
    a.nseg = nscenes;
    SahCaps k;
    // a level of a batch holds the ranges of every scene: sized as one large scene
    sah_sizes(total, a);
    // every scene may add its root task and a few more to a level
    SQ_TRY(sah_prepare(c, total + 16 * nscenes, 0, nb, a, k));
    SQ_TRY(c->aux[0].ensure((size_t)12 * nscenes * 4 + 256));
    uint32_t* red = c->aux[0].as<uint32_t>();
    const uint32_t ib = (12 * nscenes + 255) / 256;
    hipLaunchKernelGGL(k_batch_sah_red_init, dim3(ib), dim3(256), 0, c->stream, red, nscenes, a.counters);
    const uint32_t tiles = (total + kSbTile - 1) / kSbTile;
    hipLaunchKernelGGL(k_batch_sah_setup, dim3(tiles), dim3(256), 0, c->stream, tris, d_off, nscenes, total,
                       (float4*)a.boxes, c->vals[0].as<uint32_t>(), red);
    a.ids_in = c->vals[0].as<uint32_t>();
    a.out_buf = 0;
    hipLaunchKernelGGL(k_batch_sah_root, dim3(nscenes), dim3(256), 0, c->stream, a, red);
    SQ_CHECK_HIP(hipGetLastError());
    uint32_t levels = 0;
    SQ_TRY(sah_levels(c, a, 0, k, &levels));
    *depth = levels + 2;
    return SQ_OK;
}

int refit_tree(sq_ctx* c, sq_bvh* b) {
    (void)c; (void)b;
    SQ_RET(SQ_ERR_UNSUPPORTED, "refit is not available in this build");
}

}  // namespace sq
