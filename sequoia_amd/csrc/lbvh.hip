// lbvh.hip — top-down Morton/LBVH node emission and bottom-up box refit.
//
// Reproduces recurse_morton / build_hierarchy_morton (builder.cpp:2281-2409,
// 2474-2508) slot for slot: root at 0 with child 1; a range of fewer than 8
// primitives becomes a fan of single-triangle leaves; otherwise it splits
// after find_split_morton's index (all-equal range -> midpoint n>>1) and the
// children go to the reference's slots
//   left  node f,            children from f + new_nodes
//   right node f + (L != 1), children from f + new_nodes + 2(L-1) + (L == 1)
// including the right-singleton quirk (layout REFERENCE) or with the
// singleton given the second pair slot (layout FIXED).
//
// A range of c >= 2 primitives owns the 2(c-1) slots from its free index; a
// fan of m leaves uses m of its 2(m-1) and leaves the rest unused, and slot
// 2n-1 of a scene's 2n is never used. With refit on, the emission writes
// every used slot whole and zeroes every unused one, so the node array needs
// no fill beforehand (without refit the links-only stores rely on one).
//
// Boxes (no reference counterpart, the reference writes none): leaf = the
// triangle's bounds; interior = union of its child list in list order.
//
// Three kernels, no cross-workgroup waiting, any number of scenes per launch
// (scene s: sorted range [seg_begin, seg_begin + n), nodes at node_base,
// leaf child = triangle index - tri_base):
//   top    one workgroup per scene splits every range larger than T level
//          by level, one range per thread (the split from a table of the
//          codes' top bits, else an 8-ary search on the sorted codes), and
//          queues ranges of 8..T primitives,
//   small  persistent workgroups take queued ranges of every scene: codes in
//          LDS, the reference's binary search per thread, level by level,
//          then refit bottom-up,
//   finish one workgroup per scene refits the nodes written by `top`,
//          deepest first, and records the scene's depth.
#include "internal.h"

namespace sq {

constexpr int kTopThreads = 1024;
constexpr int kSmallThreads = 512;
constexpr uint32_t kDefaultThreshold = 1024;  // 2048 takes half the LDS occupancy (see k_lbvh_small)
constexpr uint32_t kMaxThreshold = 2048;
constexpr int kMaxLevels = 160;

template <typename K>
__device__ __forceinline__ int clz_k(K x) {
    if (sizeof(K) == 8) return __clzll((long long)x);
    return __clz((int)x);
}

struct EmitArgs {
    const sq_triangle* tris;  // the batch's triangles (global ids)
    const uint32_t* idx;      // sorted global triangle ids
    sq_node* nodes;           // the scene's Node1[] (scene-local slots)
    uint32_t layout;
    uint32_t refit;
    uint32_t cap;             // node capacity of the scene: writes at or beyond it are dropped
    uint32_t tri_base;        // first global triangle id of the scene
};

__device__ __forceinline__ void write_leaf(const EmitArgs& a, uint32_t slot, uint32_t gtri, sqd::Box3* box) {
    if (slot >= a.cap) return;
    const uint32_t local = gtri - a.tri_base;
    if (a.refit) {
        float v[9];
        sqd::load_tri(a.tris, gtri, v);
        sqd::Box3 b = sqd::tri_box(v);
        sqd::store_node(a.nodes, slot, b, local, 1, 1);
        if (box) *box = b;
    } else {
        sqd::store_node_links(a.nodes, slot, local, 1, 1);
    }
}

// A fan of m < 8 single-triangle leaves at slots f.. for sorted positions
// b..b+m-1, and the union of their boxes. All index and triangle fetches are
// issued before any result is used (one thread writes the whole fan, so
// fetching leaf by leaf would wait for each in turn).
__device__ __forceinline__ void write_fan(const EmitArgs& a, uint32_t f, uint32_t b, uint32_t m, sqd::Box3* u) {
    // the fan's unused slots f+m .. f+2m-3
    for (uint32_t i = f + m; i < f + 2 * m - 2; i++)
        if (i < a.cap) sqd::store_zero_node(a.nodes, i);
    uint32_t gt[7];
#pragma unroll
    for (uint32_t i = 0; i < 7; i++) gt[i] = i < m ? a.idx[b + i] : 0u;
    if (!a.refit) {
#pragma unroll
        for (uint32_t i = 0; i < 7; i++)
            if (i < m) write_leaf(a, f + i, gt[i], nullptr);
        return;
    }
    float v[7][9];
#pragma unroll
    for (uint32_t i = 0; i < 7; i++)
        if (i < m) sqd::load_tri(a.tris, gt[i], v[i]);
#pragma unroll
    for (uint32_t i = 0; i < 7; i++) {
        if (i >= m) break;
        const sqd::Box3 lb = sqd::tri_box(v[i]);
        if (f + i < a.cap) sqd::store_node(a.nodes, f + i, lb, gt[i] - a.tri_base, 1, 1);
        if (i == 0) *u = lb; else sqd::box_grow(*u, lb);
    }
}

// Writes the two children of range [b, e) split after `split`; hands the
// child ranges that still need splitting (count >= 8) to the callback.
template <typename Push>
__device__ __forceinline__ void emit_children(const EmitArgs& a, uint32_t b, uint32_t e, uint32_t f, uint32_t split, Push push) {
    const uint32_t count = e - b;
    const uint32_t L = split + 1, R = count - L;
    const uint32_t nn = (L != 1) + (R != 1);
    uint32_t lfree = f + nn;
    uint32_t rfree = f + nn + (L - 1) * 2 + (L == 1);
    const uint32_t lslot = f;
    const uint32_t rslot = f + (L != 1);
    if (a.layout == SQ_LAYOUT_FIXED && R == 1 && L != 1) {
        lfree = f + 2;
        rfree = f + 1;
    }
    // left child
    if (L == 1) {
        write_leaf(a, lfree, a.idx[b], nullptr);
    } else if (L < 8) {
        sqd::Box3 u;
        write_fan(a, lfree, b, L, &u);
        if (lslot < a.cap) {
            if (a.refit) sqd::store_node(a.nodes, lslot, u, lfree, L, 0);
            else sqd::store_node_links(a.nodes, lslot, lfree, L, 0);
        }
    } else {
        if (lslot < a.cap) sqd::store_node_links(a.nodes, lslot, lfree, 2, 0);
        push(b, b + L, lfree, lslot);
    }
    // right child
    if (R == 1) {
        write_leaf(a, rfree, a.idx[b + L], nullptr);
    } else if (R < 8) {
        sqd::Box3 u;
        write_fan(a, rfree, b + L, R, &u);
        if (rslot < a.cap) {
            if (a.refit) sqd::store_node(a.nodes, rslot, u, rfree, R, 0);
            else sqd::store_node_links(a.nodes, rslot, rfree, R, 0);
        }
    } else {
        if (rslot < a.cap) sqd::store_node_links(a.nodes, rslot, rfree, 2, 0);
        push(b + L, e, rfree, rslot);
    }
}

// Union of a node's child list, in list order (interior refit).
__device__ __forceinline__ void refit_node(sq_node* nodes, uint32_t self, uint32_t cap) {
    if (self >= cap) return;
    const uint32_t* w = (const uint32_t*)(nodes + self);
    uint32_t cnt = w[3] & 0xFFFFu, child = w[7];
    if (cnt == 0 || child >= cap || cnt > cap - child) return;
    sqd::Box3 u = sqd::load_node_box(nodes, child);
    for (uint32_t j = 1; j < cnt; j++) sqd::box_grow(u, sqd::load_node_box(nodes, child + j));
    sqd::store_node_box(nodes, self, u);
}

// Top bits of a sorted code array: table[p] = the first position whose code
// has top bits (code >> pshift) >= p, for p < 2^(key bits - pshift). Each
// position that starts a run of equal top bits fills the entries from just
// past the previous run's top bits up to its own (the last position also
// fills the entries past it with n), so every entry is written once.
struct PrefixTable {
    const uint32_t* table;  // null: no table
    uint32_t pshift;
    uint32_t entries;
};

// A wave fills its lanes' entry ranges together, 64 entries per store, so a
// long range of empty top bits (the inside of a closed surface) costs no one
// lane thousands of stores.
// The same launch zeroes the emission counters (`zero`, `zero_words`) and
// writes the build's one scene descriptor, which would otherwise take a
// launch of their own.
template <typename K>
__global__ void k_prefix_table(const K* __restrict__ codes, uint32_t n, uint32_t pshift, uint32_t entries,
                               uint32_t* __restrict__ table, uint32_t* __restrict__ zero, uint32_t zero_words,
                               SceneDesc* desc, SceneDesc value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < zero_words; i += gridDim.x * blockDim.x) zero[i] = 0;
    if (desc && blockIdx.x == 0 && threadIdx.x == 0) *desc = value;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u), wstride = gridDim.x * blockDim.x;
    for (uint32_t wb = wave0; wb <= n; wb += wstride) {  // wave-uniform loop over positions 0..n
        const uint32_t i = wb + lane;
        uint64_t lo = 1, hi = 0;  // empty
        if (i <= n) {
            const uint64_t cur = i < n ? (uint64_t)(codes[i] >> pshift) : (uint64_t)entries;
            const uint64_t prev = i > 0 ? (uint64_t)(codes[i - 1] >> pshift) : ~0ull;
            if (cur != prev) {
                lo = prev + 1;  // ~0 + 1 == 0 for the first position
                hi = cur < entries ? cur : (uint64_t)entries - 1;
            }
        }
        uint64_t pending = __ballot(lo <= hi);
        while (pending) {
            const int l = __ffsll((long long)pending) - 1;
            pending &= pending - 1;
            const uint64_t plo = __shfl(lo, l, 64), phi = __shfl(hi, l, 64);
            const uint32_t v = __shfl(i, l, 64);
            for (uint64_t p = plo + lane; p <= phi; p += 64) table[p] = v;
        }
    }
}

// find_split_morton (builder.cpp:2281-2314), literal, on LDS codes.
template <typename K>
__device__ uint32_t find_split_ref(const K* codes, uint32_t count) {
    const K first = codes[0], lastc = codes[count - 1];
    if (first == lastc) return count >> 1;
    const int common = clz_k<K>(first ^ lastc);
    int split = 0, last = (int)count - 1, step = last;
    do {
        step = (step + 1) >> 1;
        int ns = split + step;
        if (ns < last) {
            if (clz_k<K>(first ^ codes[ns]) > common) split = ns;
        }
    } while (step > 1);
    return (uint32_t)split;
}

struct LbvhArgs {
    const sq_triangle* tris;
    const uint32_t* idx;
    const void* codes;
    sq_node* nodes;
    const SceneDesc* scenes;
    uint32_t layout, refit, threshold;
    Task* qa;                  // per-scene regions at SceneDesc::q_base
    Task* qb;
    Task* small;               // ranges of 8..T primitives of every scene
    uint32_t* top_list;        // per-scene regions at SceneDesc::top_base
    uint32_t* level_off;       // per-scene tables at SceneDesc::lvl_base
    uint32_t* scene_levels;    // per scene: number of top levels
    uint32_t* scene_small;     // per scene: most levels of one small range
    uint32_t* depth_out;       // per scene: tree depth (root = 1), may be null
    uint32_t* counters;        // [0] small-task count, [1] max small-range levels, [2] error flag
    uint32_t small_cap;        // entries of `small`
    PrefixTable pt;            // one-scene builds: split positions of the top bits
    uint32_t* host_rec;        // one-scene builds: pinned host record written by k_lbvh_finish
                               // (words 0-2 counters, 8 depth, 16 top levels), else null
};

__device__ __forceinline__ EmitArgs scene_emit(const LbvhArgs& a, const SceneDesc& d) {
    return EmitArgs{a.tris, a.idx, a.nodes + d.node_base, a.layout, a.refit, 2 * d.n, d.tri_base};
}

__device__ __forceinline__ void push_small(const LbvhArgs& a, const Task& t) {
    const uint32_t k = atomicAdd(&a.counters[0], 1u);
    if (k < a.small_cap) a.small[k] = t;
    else atomicOr(&a.counters[2], 1u);
}

// find_split_morton by one thread: the table when the highest differing bit
// lies within its top bits, else an 8-ary search (seven independent loads per
// step). The predicate "shares more leading bits with the first code than
// the last does" holds on a prefix of the sorted range, so the boundary the
// search finds is the reference's binary-search result.
template <typename K>
__device__ uint32_t thread_find_split(const K* __restrict__ codes, uint32_t b, uint32_t e, const PrefixTable& pt) {
    const uint32_t count = e - b;
    const K first = codes[b], last = codes[e - 1];
    if (first == last) return count >> 1;
    const int cp = clz_k<K>(first ^ last);
    const int h = (int)(8 * sizeof(K)) - 1 - cp;
    if (pt.table && h >= (int)pt.pshift) {
        const K x = ((first >> h) | (K)1) << h;
        const uint64_t p = (uint64_t)(x >> pt.pshift);
        if (p < pt.entries) {
            const uint32_t lb = pt.table[p];
            if (lb > b && lb < e) return lb - 1 - b;
        }
    }
    uint32_t lo = 0, hi = count - 1;  // the predicate holds at lo and fails at hi
    while (hi - lo > 1) {
        const uint32_t span = hi - lo;
        uint32_t pr[7];
        K c[7];
#pragma unroll
        for (int j = 0; j < 7; j++) {
            pr[j] = lo + (uint32_t)(((uint64_t)(j + 1) * span) >> 3);
            c[j] = codes[b + pr[j]];
        }
        uint32_t nlo = lo, nhi = hi;
#pragma unroll
        for (int j = 0; j < 7; j++) {
            if (pr[j] <= lo || pr[j] >= hi) continue;
            if (clz_k<K>(first ^ c[j]) > cp) nlo = pr[j] > nlo ? pr[j] : nlo;
            else nhi = pr[j] < nhi ? pr[j] : nhi;
        }
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

// Top phase with one range per thread: one workgroup per scene splits the
// ranges larger than T level by level, a level's ranges spread over the
// workgroup's threads (up to kTopThreads at once), the queue of the next
// level in LDS (in the scene's global queue when a level may hold more than
// kLaneQ ranges). The small ranges a chunk of threads produces take their
// places in the small queue with one global atomic per chunk. Slots, writes
// and the top list (level by level, in queue order) follow the reference's
// recursion; only the order in which a level's ranges are split differs, and
// node slots do not depend on it.
constexpr uint32_t kLaneQ = 2048;

template <typename K>
__global__ __launch_bounds__(kTopThreads) void k_lbvh_top_lanes(LbvhArgs a) {
    __shared__ Task s_q[2][kLaneQ];
    __shared__ uint32_t s_ncur, s_nnext, s_nsmall, s_small_base;
    const uint32_t scene = blockIdx.x;
    const SceneDesc d = a.scenes[scene];
    const EmitArgs e = scene_emit(a, d);
    const K* codes = (const K*)a.codes;
    const uint32_t qcap = d.n / a.threshold + 4;  // ranges > T in one level are disjoint
    const bool lds_q = qcap <= kLaneQ;
    Task* const q0 = lds_q ? s_q[0] : a.qa + d.q_base;
    Task* const q1 = lds_q ? s_q[1] : a.qb + d.q_base;
    uint32_t* top_list = a.top_list + d.top_base;
    uint32_t* level_off = a.level_off + d.lvl_base;
    if (threadIdx.x == 0) {
        s_ncur = 0; s_nnext = 0; s_nsmall = 0; s_small_base = 0;
        // the scene's unused slots: n+1 .. 2n-1 behind a root fan, else 2n-1
        for (uint32_t i = d.n < 8 ? d.n + 1 : 2 * d.n - 1; i < 2 * d.n; i++) sqd::store_zero_node(e.nodes, i);
        if (d.n < 8) {
            sqd::Box3 u, lb;
            for (uint32_t i = 0; i < d.n; i++) {
                write_leaf(e, 1 + i, a.idx[d.seg_begin + i], &lb);
                if (i == 0) u = lb; else sqd::box_grow(u, lb);
            }
            if (d.n > 0) {
                if (a.refit) sqd::store_node(e.nodes, 0, u, 1, d.n, 0);
                else sqd::store_node_links(e.nodes, 0, 1, d.n, 0);
            }
        } else {
            sqd::store_node_links(e.nodes, 0, 1, 2, 0);
            Task t{d.seg_begin, d.seg_begin + d.n, 1, 0, scene};
            if (d.n > a.threshold) { q0[0] = t; s_ncur = 1; }
            else push_small(a, t);
        }
    }
    __syncthreads();
    uint32_t level = 0, top_off = 0, cur = 0;
    while (true) {
        const uint32_t ncur = s_ncur;
        if (ncur == 0 || level >= (uint32_t)kMaxLevels) break;
        if (threadIdx.x == 0) level_off[level] = top_off;
        for (uint32_t c0 = 0; c0 < ncur; c0 += kTopThreads) {
            const uint32_t t = c0 + threadIdx.x;
            Task sm0{}, sm1{};
            uint32_t nsm = 0, li = 0;
            if (t < ncur) {
                const Task tk = (cur ? q1 : q0)[t];
                const uint32_t split = thread_find_split<K>(codes, tk.begin, tk.end, a.pt);
                Task* const nxt = cur ? q0 : q1;
                emit_children(e, tk.begin, tk.end, tk.free, split,
                              [&](uint32_t b, uint32_t en, uint32_t f, uint32_t self) {
                                  const Task c{b, en, f, self, scene};
                                  if (en - b > a.threshold) {
                                      const uint32_t k = atomicAdd(&s_nnext, 1u);
                                      if (k < qcap) nxt[k] = c;
                                      else atomicOr(&a.counters[2], 1u);
                                  } else {
                                      if (nsm == 0) sm0 = c; else sm1 = c;
                                      nsm++;
                                  }
                              });
                if (top_off + t < d.top_cap) top_list[top_off + t] = tk.self;
                else atomicOr(&a.counters[2], 1u);
                if (nsm) li = atomicAdd(&s_nsmall, nsm);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                s_small_base = s_nsmall ? atomicAdd(&a.counters[0], s_nsmall) : 0u;
                s_nsmall = 0;
            }
            __syncthreads();
            const uint32_t k = s_small_base + li;
            if (nsm >= 1) {
                if (k < a.small_cap) a.small[k] = sm0;
                else atomicOr(&a.counters[2], 1u);
            }
            if (nsm >= 2) {
                if (k + 1 < a.small_cap) a.small[k + 1] = sm1;
                else atomicOr(&a.counters[2], 1u);
            }
        }
        // every push of this level is in (the chunk barriers ordered them)
        if (threadIdx.x == 0) {
            s_ncur = s_nnext;
            s_nnext = 0;
        }
        top_off += ncur;
        cur ^= 1;
        level++;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (level < kMaxLevels) level_off[level] = top_off;
        a.scene_levels[scene] = level;
    }
}

// A split task inside one small range, in LDS: begin/end relative to the
// range (16 bits each), first free slot, own slot.
struct SmallTask {
    uint32_t be;
    uint32_t free;
    uint32_t self;
};

// What each node slot of a small range holds, relative to the range's first
// child slot (a range of c primitives places all its nodes and leaves in the
// 2(c-1) slots from its root's `free`): a leaf (sorted position), a fan node
// (first child slot, count), or a split node (written from the task list).
constexpr uint32_t kSlotLeaf = 1u << 31, kSlotFan = 1u << 30, kSlotSplit = 1u << 29;

// Persistent: one workgroup finishes one queued range at a time, entirely in
// LDS until its nodes are written:
//   1. codes in LDS; level by level each task is split with the reference's
//      binary search and its children recorded (slot tags, new tasks);
//   2. every leaf's triangle is fetched at once and its node written; the
//      leaf boxes go to an LDS box per slot (in the space the codes held);
//   3. fan nodes, then the split nodes deepest level first, take the union of
//      their child slots' LDS boxes in slot order, and are written; slots
//      that hold nothing (the unused tail of each fan) are zeroed.
// Only phase 2 reads global memory after the codes, and every node is
// written once, whole.
//
// Capacity: a range of c <= T primitives has at most c - 6 split tasks
// (tasks are ranges of >= 8; leaves of the task tree are disjoint, and every
// task with a single task child peels off primitives no other task holds),
// so T entries always suffice; the level count is at most the number of
// code bits plus log2(T). Both are still checked, and a violation raises the
// launch's error flag instead of writing out of bounds.
template <typename K, uint32_t TM, int NT>
__global__ __launch_bounds__(NT) void k_lbvh_small(LbvhArgs a) {
    constexpr int kSlots = 2 * TM;
    // the codes (emission) and the slot boxes (phases 2-3) share one area
    constexpr size_t kCodeBytes = sizeof(K) * TM, kBoxBytes = sizeof(sqd::Box3) * kSlots;
    __shared__ __align__(16) unsigned char s_area[kCodeBytes > kBoxBytes ? kCodeBytes : kBoxBytes];
    __shared__ uint32_t s_slot[kSlots];
    __shared__ uint16_t s_leafslot[TM];  // sorted position -> its leaf's slot (range-relative)
    __shared__ SmallTask s_tasks[TM];
    __shared__ uint32_t s_lvl[kMaxLevels + 1];
    __shared__ uint32_t s_total, s_err;
    K* s_codes = (K*)s_area;
    sqd::Box3* s_box = (sqd::Box3*)s_area;
    const K* codes = (const K*)a.codes;
    const uint32_t ntasks = a.counters[0];
    uint32_t max_lev = 0;
    for (uint32_t ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const Task root = a.small[ti];
        const SceneDesc d = a.scenes[root.scene];
        const EmitArgs e = scene_emit(a, d);
        const uint32_t base = root.begin, count = root.end - root.begin;
        if (count > TM || count < 8) {  // never queued so; flag and skip
            if (threadIdx.x == 0) atomicOr(&a.counters[2], 1u);
            continue;
        }
        const uint32_t sbase = root.free, nslots = 2 * (count - 1);
        // this thread's leaves are positions threadIdx.x + j * NT: their
        // codes and triangle ids are loaded together (branch-free: positions past
        // the range load the range's last entry)
        constexpr int kLeafPer = (TM + NT - 1) / NT;
        uint32_t lt[kLeafPer];
        K ck[kLeafPer];
#pragma unroll
        for (int j = 0; j < kLeafPer; j++) {
            const uint32_t i = threadIdx.x + j * NT, ic = i < count ? i : count - 1;
            ck[j] = codes[base + ic];
            lt[j] = a.idx[base + ic];
        }
#pragma unroll
        for (int j = 0; j < kLeafPer; j++)
            if (threadIdx.x + j * NT < count) s_codes[threadIdx.x + j * NT] = ck[j];
        for (uint32_t i = threadIdx.x; i < nslots; i += NT) s_slot[i] = 0;
        if (threadIdx.x == 0) {
            s_tasks[0] = SmallTask{count << 16, root.free, root.self};
            s_total = 1;
            s_err = 0;
        }
        sqd::lds_barrier();
        // 1. emission: slot tags and split tasks, level by level
        uint32_t lb = 0, le = 1, nlev = 0;
        while (lb < le) {
            if (threadIdx.x == 0) {
                if (nlev < (uint32_t)kMaxLevels) s_lvl[nlev] = lb;
                else s_err = 1;
            }
            for (uint32_t i = lb + threadIdx.x; i < le; i += NT) {
                const SmallTask tk = s_tasks[i];
                const uint32_t tb = tk.be & 0xFFFFu, te = tk.be >> 16, f = tk.free;
                const uint32_t cnt = te - tb, L = find_split_ref<K>(s_codes + tb, cnt) + 1, R = cnt - L;
                const uint32_t nn = (L != 1) + (R != 1);
                uint32_t lfree = f + nn, rfree = f + nn + (L - 1) * 2 + (L == 1);
                const uint32_t lslot = f, rslot = f + (L != 1);
                if (e.layout == SQ_LAYOUT_FIXED && R == 1 && L != 1) {
                    lfree = f + 2;
                    rfree = f + 1;
                }
                const uint32_t side_b[2] = {tb, tb + L}, side_n[2] = {L, R};
                const uint32_t side_free[2] = {lfree, rfree}, side_slot[2] = {lslot, rslot};
                for (int sd = 0; sd < 2; sd++) {
                    const uint32_t cb = side_b[sd], cn = side_n[sd], cf = side_free[sd] - sbase;
                    const uint32_t cs = side_slot[sd] - sbase;
                    if (cn < 8) {
                        // a leaf (cn == 1) or a fan of cn leaves at cf..
                        if (cf + cn > nslots || (cn > 1 && cs >= nslots)) { s_err = 1; continue; }
                        for (uint32_t j = 0; j < cn; j++) {
                            s_slot[cf + j] = kSlotLeaf | (cb + j);
                            s_leafslot[cb + j] = (uint16_t)(cf + j);
                        }
                        if (cn > 1) s_slot[cs] = kSlotFan | (cn << 16) | cf;
                    } else {
                        const uint32_t k = atomicAdd(&s_total, 1u);
                        if (k < TM) s_tasks[k] = SmallTask{cb | ((cb + cn) << 16), side_free[sd], side_slot[sd]};
                        else s_err = 1;
                        if (cs < nslots) s_slot[cs] = kSlotSplit;
                    }
                }
            }
            sqd::lds_barrier();
            lb = le;
            le = s_total < TM ? s_total : TM;
            nlev++;
            const bool stop = s_err != 0;
            sqd::lds_barrier();
            if (stop) break;
        }
        if (s_err) {
            if (threadIdx.x == 0) atomicOr(&a.counters[2], 1u);
            sqd::lds_barrier();
            continue;
        }
        if (threadIdx.x == 0) s_lvl[nlev] = le;
        max_lev = nlev > max_lev ? nlev : max_lev;
        if (threadIdx.x == 0) atomicMax(&a.scene_small[root.scene], nlev);
        // 2. leaves: every triangle of the range fetched at once (the codes are no
        // longer needed)
        float lv[kLeafPer][9];
        if (a.refit) {
#pragma unroll
            for (int j = 0; j < kLeafPer; j++) sqd::load_tri(a.tris, lt[j], lv[j]);
        }
#pragma unroll
        for (int j = 0; j < kLeafPer; j++) {
            const uint32_t i = threadIdx.x + j * NT;
            if (i >= count) break;
            const uint32_t sl = s_leafslot[i], slot = sbase + sl;
            if (a.refit) {
                const sqd::Box3 b = sqd::tri_box(lv[j]);
                s_box[sl] = b;
                if (slot < e.cap) sqd::store_node(e.nodes, slot, b, lt[j] - e.tri_base, 1, 1);
            } else if (slot < e.cap) {
                sqd::store_node_links(e.nodes, slot, lt[j] - e.tri_base, 1, 1);
            }
        }
        sqd::lds_barrier();
        // 3. fans, then split nodes deepest first: unions of child slots in slot order
        for (uint32_t sl = threadIdx.x; sl < nslots; sl += NT) {
            const uint32_t tag = s_slot[sl];
            if (tag == 0 && sbase + sl < e.cap) sqd::store_zero_node(e.nodes, sbase + sl);  // unused slot
            if (!(tag & kSlotFan)) continue;
            const uint32_t c0 = tag & 0xFFFFu, cn = (tag >> 16) & 0x3FFFu, slot = sbase + sl;
            if (c0 + cn > nslots) continue;  // never tagged so (checked at emission)
            if (a.refit) {
                sqd::Box3 u = s_box[c0];
                for (uint32_t j = 1; j < cn; j++) sqd::box_grow(u, s_box[c0 + j]);
                s_box[sl] = u;
                if (slot < e.cap) sqd::store_node(e.nodes, slot, u, sbase + c0, cn, 0);
            } else if (slot < e.cap) {
                sqd::store_node_links(e.nodes, slot, sbase + c0, cn, 0);
            }
        }
        sqd::lds_barrier();
        for (int lv = (int)nlev - 1; lv >= 0; lv--) {
            for (uint32_t i = s_lvl[lv] + threadIdx.x; i < s_lvl[lv + 1]; i += NT) {
                const SmallTask tk = s_tasks[i];
                const uint32_t c0 = tk.free - sbase;
                if (c0 + 1 >= nslots || (i > 0 && tk.self - sbase >= nslots)) continue;  // never queued so
                if (a.refit) {
                    sqd::Box3 u = s_box[c0];
                    sqd::box_grow(u, s_box[c0 + 1]);
                    if (i > 0) s_box[tk.self - sbase] = u;  // the range's root lies outside its slots
                    if (tk.self < e.cap) sqd::store_node(e.nodes, tk.self, u, tk.free, 2, 0);
                } else if (tk.self < e.cap) {
                    sqd::store_node_links(e.nodes, tk.self, tk.free, 2, 0);
                }
            }
            sqd::lds_barrier();
        }
    }
    if (threadIdx.x == 0 && max_lev) atomicMax(&a.counters[1], max_lev);
}

// One workgroup per scene: refits the nodes the top phase wrote, deepest
// first, and records the scene's depth.
__global__ __launch_bounds__(kTopThreads) void k_lbvh_finish(LbvhArgs a) {
    const uint32_t scene = blockIdx.x;
    const SceneDesc d = a.scenes[scene];
    const uint32_t levels = a.scene_levels[scene];
    const bool failed = a.counters[2] != 0 || levels >= (uint32_t)kMaxLevels;
    // depth 0 marks a tree whose emission ran out of workspace (see counters[2])
    const uint32_t depth = failed ? 0u : levels + a.scene_small[scene] + 2;
    if (threadIdx.x == 0 && a.depth_out) a.depth_out[scene] = depth;
    if (threadIdx.x == 0 && a.host_rec && scene == 0) {
        // the host reads these once the stream is idle: no copy command for them
        a.host_rec[0] = a.counters[0];
        a.host_rec[1] = a.counters[1];
        a.host_rec[2] = a.counters[2];
        a.host_rec[8] = depth;
        a.host_rec[16] = levels;
        __threadfence_system();
    }
    if (!a.refit || failed) return;
    sq_node* nodes = a.nodes + d.node_base;
    const uint32_t cap = 2 * d.n;
    const uint32_t* top_list = a.top_list + d.top_base;
    const uint32_t* level_off = a.level_off + d.lvl_base;
    for (int lv = (int)levels - 1; lv >= 0; lv--) {
        uint32_t lo = level_off[lv], hi = level_off[lv + 1];
        for (uint32_t i = lo + threadIdx.x; i < hi; i += kTopThreads) refit_node(nodes, top_list[i], cap);
        __syncthreads();
    }
}

// Zeroes the emission counters; `desc` non-null: also writes the one scene
// descriptor of a plain build (no host-to-device copy for it).
__global__ void k_zero_counters(uint32_t* c, uint32_t n, SceneDesc* desc, SceneDesc value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c[i] = 0;
    if (desc && blockIdx.x == 0 && threadIdx.x == 0) *desc = value;
}

static uint32_t clamp_threshold(uint32_t threshold) {
    uint32_t T = threshold ? threshold : kDefaultThreshold;
    if (T < 8) T = 8;
    if (T > kMaxThreshold) T = kMaxThreshold;
    return T;
}

// Ranges split by the top phase: at most n/T + 1 per level (disjoint, > T
// each) over at most 3*21 code-bit levels plus log2(n) midpoint levels of
// all-equal ranges, so 96 levels' worth bounds a scene's top list.
constexpr uint32_t kTopListLevels = 96;
static uint32_t top_list_cap(uint32_t n, uint32_t T) { return (n / T + 1) * kTopListLevels; }

// Scene descriptors for a concatenation of scenes: scene s covers triangles
// (and sorted positions) [tri_base, tri_base + n) and nodes [2 tri_base,
// 2 (tri_base + n)); its queue, top-list and level-table regions are prefix
// sums of n/T + 4, top_list_cap(n, T) and kMaxLevels + 2 entries.
void lbvh_fill_descs(SceneDesc* d, const uint32_t* counts, uint32_t nscenes, uint32_t threshold) {
    const uint32_t T = clamp_threshold(threshold);
    uint32_t tri = 0, q = 0, top = 0;
    for (uint32_t s = 0; s < nscenes; s++) {
        const uint32_t n = counts[s];
        const uint32_t top_cap = top_list_cap(n, T);
        d[s] = SceneDesc{tri, n, tri, 2 * tri, q, top, s * (uint32_t)(kMaxLevels + 2), top_cap};
        tri += n;
        q += n / T + 4;
        top += top_cap;
    }
}

struct WsLayout {
    size_t qa, qb, small, top, lvl, counters, scene_levels, scene_small, zero_end, total;
    size_t small_cap;
};
static WsLayout ws_layout(uint32_t total, uint32_t nscenes, uint32_t T) {
    WsLayout w;
    size_t q = (size_t)total / T + 4 * (size_t)nscenes + 4;
    size_t sm = (size_t)total / 8 + nscenes + 4;  // small ranges are disjoint, >= 8 each
    size_t tp = ((size_t)total / T + nscenes) * kTopListLevels + 8;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    w.qa = take(q * sizeof(Task));
    w.qb = take(q * sizeof(Task));
    w.small = take(sm * sizeof(Task));
    w.small_cap = sm;
    w.top = take(tp * 4);
    w.lvl = take((size_t)nscenes * (kMaxLevels + 2) * 4);
    // counters, scene_levels and scene_small are contiguous (zeroed together)
    w.counters = off;
    off += 64;
    w.scene_levels = off;
    off += (size_t)nscenes * 4;
    w.scene_small = off;
    off += (size_t)nscenes * 4;
    w.zero_end = off;
    off = (off + 255) & ~(size_t)255;
    w.total = off;
    return w;
}

size_t lbvh_batch_workspace(uint32_t total, uint32_t nscenes, uint32_t threshold) {
    return ws_layout(total, nscenes, clamp_threshold(threshold)).total;
}

static int lbvh_launch(sq_ctx* c, hipStream_t stream, const sq_triangle* tris, uint32_t total, const void* d_codes,
                       const uint32_t* d_idx, bool key64, const SceneDesc* d_scenes, uint32_t nscenes, uint32_t layout,
                       bool refit, uint32_t T, sq_node* d_nodes, void* ws, uint32_t* d_depth, bool split_refit_phase,
                       const SceneDesc* write_desc = nullptr, PrefixTable pt = PrefixTable{},
                       uint32_t* table_out = nullptr, uint32_t* host_rec = nullptr) {
    WsLayout w = ws_layout(total, nscenes, T);
    char* base = (char*)ws;
    LbvhArgs a;
    a.tris = tris;
    a.idx = d_idx;
    a.codes = d_codes;
    a.nodes = d_nodes;
    a.scenes = d_scenes;
    a.layout = layout;
    a.refit = refit ? 1u : 0u;
    a.threshold = T;
    a.qa = (Task*)(base + w.qa);
    a.qb = (Task*)(base + w.qb);
    a.small = (Task*)(base + w.small);
    a.top_list = (uint32_t*)(base + w.top);
    a.level_off = (uint32_t*)(base + w.lvl);
    a.scene_levels = (uint32_t*)(base + w.scene_levels);
    a.scene_small = (uint32_t*)(base + w.scene_small);
    a.depth_out = d_depth;
    a.counters = (uint32_t*)(base + w.counters);
    a.small_cap = (uint32_t)w.small_cap;
    a.pt = pt;
    a.host_rec = host_rec;
    const uint32_t zero_words = (uint32_t)((w.zero_end - w.counters) / 4);
    SceneDesc* desc = write_desc ? const_cast<SceneDesc*>(d_scenes) : nullptr;
    const SceneDesc desc_value = write_desc ? *write_desc : SceneDesc{};
    if (table_out) {
        // one scene: the split table's launch also zeroes the counters
        uint32_t tb = (total + 1 + 255) / 256;
        const uint32_t cap = (uint32_t)c->num_cus * 8;
        if (tb > cap) tb = cap;
        if (key64)
            hipLaunchKernelGGL(k_prefix_table<uint64_t>, dim3(tb), dim3(256), 0, stream, (const uint64_t*)d_codes,
                               total, pt.pshift, pt.entries, table_out, a.counters, zero_words, desc, desc_value);
        else
            hipLaunchKernelGGL(k_prefix_table<uint32_t>, dim3(tb), dim3(256), 0, stream, (const uint32_t*)d_codes,
                               total, pt.pshift, pt.entries, table_out, a.counters, zero_words, desc, desc_value);
    } else {
        uint32_t zb = (zero_words + 255) / 256;
        zb = zb > 64 ? 64 : zb;
        hipLaunchKernelGGL(k_zero_counters, dim3(zb), dim3(256), 0, stream, a.counters, zero_words, desc, desc_value);
    }
    if (key64) hipLaunchKernelGGL(k_lbvh_top_lanes<uint64_t>, dim3(nscenes), dim3(kTopThreads), 0, stream, a);
    else hipLaunchKernelGGL(k_lbvh_top_lanes<uint32_t>, dim3(nscenes), dim3(kTopThreads), 0, stream, a);
    uint32_t max_tasks = (uint32_t)(total / 8 + nscenes);
    uint32_t grid = (uint32_t)c->num_cus * 8;
    if (grid > max_tasks) grid = max_tasks;
    if (grid == 0) grid = 1;
    // LDS sized for the threshold: up to 1024 primitives per range, two
    // workgroups fit a CU
    if (T <= 1024) {
        if (key64) hipLaunchKernelGGL((k_lbvh_small<uint64_t, 1024, kSmallThreads>), dim3(grid), dim3(kSmallThreads), 0, stream, a);
        else hipLaunchKernelGGL((k_lbvh_small<uint32_t, 1024, kSmallThreads>), dim3(grid), dim3(kSmallThreads), 0, stream, a);
    } else {
        if (key64) hipLaunchKernelGGL((k_lbvh_small<uint64_t, kMaxThreshold, kSmallThreads>), dim3(grid), dim3(kSmallThreads), 0, stream, a);
        else hipLaunchKernelGGL((k_lbvh_small<uint32_t, kMaxThreshold, kSmallThreads>), dim3(grid), dim3(kSmallThreads), 0, stream, a);
    }
    if (refit && split_refit_phase) { phase_end(c, PH_EMIT); phase_begin(c, PH_REFIT); }
    hipLaunchKernelGGL(k_lbvh_finish, dim3(nscenes), dim3(kTopThreads), 0, stream, a);
    if (refit && split_refit_phase) phase_end(c, PH_REFIT);  // the emit phase ended before the refit
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

int lbvh_emit_batch(sq_ctx* c, hipStream_t stream, const sq_triangle* tris, uint32_t total, const void* d_codes,
                    const uint32_t* d_idx, bool key64, const SceneDesc* d_scenes, uint32_t nscenes, uint32_t layout,
                    uint32_t threshold, sq_node* d_nodes, void* ws, uint32_t* d_depth_out) {
    return lbvh_launch(c, stream, tris, total, d_codes, d_idx, key64, d_scenes, nscenes, layout, true,
                       clamp_threshold(threshold), d_nodes, ws, d_depth_out, false);
}

constexpr uint32_t kRecUnwritten = 0xFFFFFFFFu;  // the error word is 0 or 1

static void mark_build_end(sq_ctx* c) {
    if (!c->build_end) return;
    (void)hipEventRecord(c->build_end, c->stream);
    c->build_end_recorded = true;
}

int lbvh_emit(sq_ctx* c, const sq_triangle* tris, uint32_t n, const void* d_codes, const uint32_t* d_idx, bool key64,
              uint32_t code_bits, uint32_t layout, bool refit, uint32_t threshold, sq_node* d_nodes, uint32_t* depth) {
    const uint32_t T = clamp_threshold(threshold);
    WsLayout w = ws_layout(n, 1, T);
    // top-bit split table: 15 top bits of the code (all of them for short codes)
    const uint32_t key_bits = code_bits;
    PrefixTable pt{};
    pt.pshift = key_bits > 15 ? key_bits - 15 : 0;
    pt.entries = 1u << (key_bits - pt.pshift);
    SQ_TRY(c->aux[1].ensure(w.total + 256 + (size_t)pt.entries * 4));
    SQ_TRY(c->aux[2].ensure(sizeof(SceneDesc)));
    SQ_TRY(pinned_ensure(c, 256));
    SceneDesc hd;
    lbvh_fill_descs(&hd, &n, 1, T);
    char* base = c->aux[1].as<char>();
    uint32_t* d_table = (uint32_t*)(base + w.total + 256);
    pt.table = d_table;
    // the depth goes to a spare counter word; k_lbvh_finish writes the counters
    // (words 0-2), the depth (word 8) and the scene's top levels (word 16) to
    // the pinned record h, whose error word starts at a value it never takes
    uint32_t* d_depth = (uint32_t*)(base + w.counters) + 8;
    uint32_t* h = (uint32_t*)c->pinned;
    h[2] = kRecUnwritten;
    SQ_TRY(lbvh_launch(c, c->stream, tris, n, d_codes, d_idx, key64, c->aux[2].as<SceneDesc>(), 1, layout, refit, T,
                       d_nodes, c->aux[1].p, d_depth, true, &hd, pt, d_table, h));
    mark_build_end(c);
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    c->last.top_retries = 0;
    const uint32_t nsmall = h[0], top_levels = h[16];
    if (h[2] == kRecUnwritten) SQ_RET(SQ_ERR_HIP, "lbvh: the emission's result record was not written");
    if (h[2]) SQ_RET(SQ_ERR_OVERFLOW, "lbvh: emission workspace overflow");
    if (top_levels >= (uint32_t)kMaxLevels) SQ_RET(SQ_ERR_OVERFLOW, "lbvh: more than %d top levels", kMaxLevels);
    c->last.top_levels = top_levels;
    c->last.small_tasks = nsmall;
    // levels of split ranges + fan/leaf level + root
    *depth = h[8];
    return SQ_OK;
}

}  // namespace sq
