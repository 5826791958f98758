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
// Boxes (no reference counterpart, the reference writes none): leaf = the
// triangle's bounds; interior = union of its child list in list order.
//
// Three kernels, no cross-workgroup waiting:
//   top    one workgroup splits every range larger than T level by level
//          (one wave per range, 65-ary search on the sorted codes) and
//          queues ranges of 8..T primitives,
//   small  one workgroup per queued range: codes in LDS, the reference's
//          binary search per thread, level by level, then refit bottom-up,
//   refit  one workgroup refits the nodes written by `top`, deepest first.
#include "internal.h"

namespace sq {

constexpr int kTopThreads = 1024;
constexpr int kSmallThreads = 256;
constexpr uint32_t kDefaultThreshold = 2048;
constexpr uint32_t kMaxThreshold = 2048;
constexpr int kMaxLevels = 160;

template <typename K>
__device__ __forceinline__ int clz_k(K x) {
    if (sizeof(K) == 8) return __clzll((long long)x);
    return __clz((int)x);
}

struct EmitArgs {
    const sq_triangle* tris;
    const uint32_t* idx;
    sq_node* nodes;
    uint32_t layout;
    uint32_t refit;
    uint32_t cap;  // node capacity: writes at or beyond it are dropped
};

__device__ __forceinline__ void write_leaf(const EmitArgs& a, uint32_t slot, uint32_t tri, sqd::Box3* box) {
    if (slot >= a.cap) return;
    if (a.refit) {
        float v[9];
        sqd::load_tri(a.tris, tri, v);
        sqd::Box3 b = sqd::tri_box(v);
        sqd::store_node(a.nodes, slot, b, tri, 1, 1);
        if (box) *box = b;
    } else {
        sqd::store_node_links(a.nodes, slot, tri, 1, 1);
    }
}

// Writes the two children of range [b, e) split after `split`; returns the
// child ranges that still need splitting (count >= 8) through the callback.
template <typename Push>
__device__ void emit_children(const EmitArgs& a, uint32_t b, uint32_t e, uint32_t f, uint32_t split, Push push) {
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
        sqd::Box3 u, lb;
        for (uint32_t i = 0; i < L; i++) {
            write_leaf(a, lfree + i, a.idx[b + i], &lb);
            if (i == 0) u = lb; else sqd::box_grow(u, lb);
        }
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
        sqd::Box3 u, lb;
        for (uint32_t i = 0; i < R; i++) {
            write_leaf(a, rfree + i, a.idx[b + L + i], &lb);
            if (i == 0) u = lb; else sqd::box_grow(u, lb);
        }
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

// find_split_morton as a 65-ary search by one wave: the largest s < count-1
// whose code shares more leading bits with the first code than the last does.
template <typename K>
__device__ uint32_t wave_find_split(const K* __restrict__ codes, uint32_t b, uint32_t e, int lane) {
    const uint32_t count = e - b;
    const K first = codes[b], last = codes[e - 1];
    if (first == last) return count >> 1;
    const int cp = clz_k<K>(first ^ last);
    uint32_t lo = 0, hi = count - 1;
    while (hi - lo > 1) {
        uint32_t span = hi - lo;
        uint32_t p = lo + (uint32_t)(((uint64_t)(lane + 1) * span) / 65u);
        bool valid = p > lo && p < hi;
        bool pred = valid && clz_k<K>(first ^ codes[b + p]) > cp;
        uint32_t nlo = pred ? p : lo;
        uint32_t nhi = (valid && !pred) ? p : hi;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            uint32_t x = __shfl_xor(nlo, off, 64), y = __shfl_xor(nhi, off, 64);
            nlo = x > nlo ? x : nlo;
            nhi = y < nhi ? y : nhi;
        }
        lo = nlo;
        hi = nhi;
    }
    return lo;
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

struct TopArgs {
    EmitArgs e;
    const void* codes;
    uint32_t n, threshold;
    Task* q[2];
    Task* small;
    uint32_t* counters;  // [0] small count, [1] top node count, [2] levels
    uint32_t* top_list;
    uint32_t* top_level_off;
};

template <typename K>
__global__ __launch_bounds__(kTopThreads) void k_lbvh_top(TopArgs a) {
    __shared__ uint32_t s_ncur, s_nnext, s_nsmall, s_ntop, s_cur;
    const K* codes = (const K*)a.codes;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = kTopThreads / 64;
    if (threadIdx.x == 0) {
        s_ncur = 0; s_nnext = 0; s_nsmall = 0; s_ntop = 0; s_cur = 0;
        if (a.n < 8) {
            sqd::Box3 u, lb;
            for (uint32_t i = 0; i < a.n; i++) {
                write_leaf(a.e, 1 + i, a.e.idx[i], &lb);
                if (i == 0) u = lb; else sqd::box_grow(u, lb);
            }
            if (a.e.refit) sqd::store_node(a.e.nodes, 0, u, 1, a.n, 0);
            else sqd::store_node_links(a.e.nodes, 0, 1, a.n, 0);
        } else {
            sqd::store_node_links(a.e.nodes, 0, 1, 2, 0);
            Task t{0, a.n, 1, 0};
            if (a.n > a.threshold) { a.q[0][0] = t; s_ncur = 1; }
            else { a.small[0] = t; s_nsmall = 1; }
        }
    }
    __syncthreads();
    uint32_t level = 0;
    while (true) {
        const uint32_t ncur = s_ncur;
        if (ncur == 0) break;
        Task* cur = a.q[s_cur];
        Task* nxt = a.q[s_cur ^ 1];
        if (threadIdx.x == 0 && level < kMaxLevels) a.top_level_off[level] = s_ntop;
        __syncthreads();
        for (uint32_t t = wave; t < ncur; t += nwaves) {
            Task tk = cur[t];
            uint32_t split = wave_find_split<K>(codes, tk.begin, tk.end, lane);
            if (lane == 0) {
                emit_children(a.e, tk.begin, tk.end, tk.free, split,
                              [&](uint32_t b, uint32_t e, uint32_t f, uint32_t self) {
                                  Task c{b, e, f, self};
                                  if (e - b > a.threshold) nxt[atomicAdd(&s_nnext, 1u)] = c;
                                  else a.small[atomicAdd(&s_nsmall, 1u)] = c;
                              });
                a.top_list[atomicAdd(&s_ntop, 1u)] = tk.self;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_ncur = s_nnext;
            s_nnext = 0;
            s_cur ^= 1;
        }
        level++;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (level < kMaxLevels) a.top_level_off[level] = s_ntop;
        a.counters[0] = s_nsmall;
        a.counters[1] = s_ntop;
        a.counters[2] = level;
    }
}

struct SmallArgs {
    EmitArgs e;
    const void* codes;
    const Task* small;
    uint32_t* counters;  // [3] max small depth
};

template <typename K>
__global__ __launch_bounds__(kSmallThreads) void k_lbvh_small(SmallArgs a) {
    __shared__ K s_codes[kMaxThreshold];
    __shared__ Task s_tasks[kMaxThreshold / 4 + 2];
    __shared__ uint32_t s_lvl[kMaxLevels + 1];
    __shared__ uint32_t s_total;
    const K* codes = (const K*)a.codes;
    const uint32_t ntasks = a.counters[0];
    uint32_t max_lev = 0;
    // Persistent loop: the number of queued ranges is only known on the device.
    for (uint32_t ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const Task root = a.small[ti];
        const uint32_t base = root.begin, count = root.end - root.begin;
        for (uint32_t i = threadIdx.x; i < count; i += kSmallThreads) s_codes[i] = codes[base + i];
        if (threadIdx.x == 0) {
            s_tasks[0] = root;
            s_total = 1;
        }
        __syncthreads();
        uint32_t lb = 0, le = 1, nlev = 0;
        while (lb < le) {
            if (threadIdx.x == 0 && nlev < (uint32_t)kMaxLevels) s_lvl[nlev] = lb;
            for (uint32_t i = lb + threadIdx.x; i < le; i += kSmallThreads) {
                Task tk = s_tasks[i];
                uint32_t split = find_split_ref<K>(s_codes + (tk.begin - base), tk.end - tk.begin);
                emit_children(a.e, tk.begin, tk.end, tk.free, split,
                              [&](uint32_t b, uint32_t e, uint32_t f, uint32_t self) {
                                  s_tasks[atomicAdd(&s_total, 1u)] = Task{b, e, f, self};
                              });
            }
            __syncthreads();
            lb = le;
            le = s_total;
            nlev++;
            __syncthreads();
        }
        if (nlev > (uint32_t)kMaxLevels) nlev = kMaxLevels;
        if (threadIdx.x == 0) s_lvl[nlev] = le;
        max_lev = nlev > max_lev ? nlev : max_lev;
        __syncthreads();
        if (a.e.refit) {
            for (int lv = (int)nlev - 1; lv >= 0; lv--) {
                for (uint32_t i = s_lvl[lv] + threadIdx.x; i < s_lvl[lv + 1]; i += kSmallThreads)
                    refit_node(a.e.nodes, s_tasks[i].self, a.e.cap);
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && max_lev) atomicMax(&a.counters[3], max_lev);
}

__global__ __launch_bounds__(kTopThreads) void k_lbvh_top_refit(sq_node* nodes, uint32_t cap, const uint32_t* top_list,
                                                                 const uint32_t* top_level_off,
                                                                 const uint32_t* counters) {
    if (counters[2] >= (uint32_t)kMaxLevels) return;
    const uint32_t levels = counters[2];
    for (int lv = (int)levels - 1; lv >= 0; lv--) {
        uint32_t lo = top_level_off[lv], hi = top_level_off[lv + 1];
        for (uint32_t i = lo + threadIdx.x; i < hi; i += kTopThreads) refit_node(nodes, top_list[i], cap);
        __syncthreads();
    }
}

__global__ void k_zero_counters(uint32_t* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

int lbvh_emit(sq_ctx* c, const sq_triangle* tris, uint32_t n, const void* d_codes, const uint32_t* d_idx, bool key64,
              uint32_t layout, bool refit, uint32_t threshold, sq_node* d_nodes, uint32_t* depth) {
    uint32_t T = threshold ? threshold : kDefaultThreshold;
    if (T < 8) T = 8;
    if (T > kMaxThreshold) T = kMaxThreshold;
    const size_t qcap = n / T + 4;
    const size_t scap = n / 8 + 4;
    const size_t tcap = 2 * (n / T) + 8;
    SQ_TRY(c->tasks[0].ensure(qcap * sizeof(Task)));
    SQ_TRY(c->tasks[1].ensure(qcap * sizeof(Task)));
    SQ_TRY(c->tasks[2].ensure(scap * sizeof(Task)));
    SQ_TRY(c->aux[1].ensure(tcap * 4));
    SQ_TRY(c->aux[2].ensure((kMaxLevels + 2) * 4));
    SQ_TRY(c->counters.ensure(64));
    uint32_t* counters = c->counters.as<uint32_t>();
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, c->stream, counters, 16);

    EmitArgs ea{tris, d_idx, d_nodes, layout, refit ? 1u : 0u, 2 * n};
    TopArgs ta;
    ta.e = ea;
    ta.codes = d_codes;
    ta.n = n;
    ta.threshold = T;
    ta.q[0] = c->tasks[0].as<Task>();
    ta.q[1] = c->tasks[1].as<Task>();
    ta.small = c->tasks[2].as<Task>();
    ta.counters = counters;
    ta.top_list = c->aux[1].as<uint32_t>();
    ta.top_level_off = c->aux[2].as<uint32_t>();
    if (key64) hipLaunchKernelGGL(k_lbvh_top<uint64_t>, dim3(1), dim3(kTopThreads), 0, c->stream, ta);
    else hipLaunchKernelGGL(k_lbvh_top<uint32_t>, dim3(1), dim3(kTopThreads), 0, c->stream, ta);
    SQ_CHECK_HIP(hipGetLastError());

    {
        SmallArgs sa{ea, d_codes, c->tasks[2].as<Task>(), counters};
        uint32_t max_tasks = (uint32_t)(n / 8 + 1);
        uint32_t grid = (uint32_t)c->num_cus * 8;
        if (grid > max_tasks) grid = max_tasks;
        if (key64) hipLaunchKernelGGL(k_lbvh_small<uint64_t>, dim3(grid), dim3(kSmallThreads), 0, c->stream, sa);
        else hipLaunchKernelGGL(k_lbvh_small<uint32_t>, dim3(grid), dim3(kSmallThreads), 0, c->stream, sa);
        SQ_CHECK_HIP(hipGetLastError());
    }
    if (refit) {
        phase_end(c, PH_EMIT);
        phase_begin(c, PH_REFIT);
        hipLaunchKernelGGL(k_lbvh_top_refit, dim3(1), dim3(kTopThreads), 0, c->stream, d_nodes, 2 * n, ta.top_list,
                           ta.top_level_off, counters);
        SQ_CHECK_HIP(hipGetLastError());
        phase_end(c, PH_REFIT);
        phase_begin(c, PH_EMIT);
    }
    SQ_TRY(pinned_ensure(c, 64));
    uint32_t* h = (uint32_t*)c->pinned;
    SQ_CHECK_HIP(hipMemcpyAsync(h, counters, 16, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    const uint32_t nsmall = h[0], top_levels = h[2];
    if (top_levels >= (uint32_t)kMaxLevels) SQ_RET(SQ_ERR_OVERFLOW, "lbvh: more than %d top levels", kMaxLevels);
    c->last.top_levels = top_levels;
    c->last.small_tasks = nsmall;
    // levels of split ranges + fan/leaf level + root
    *depth = top_levels + h[3] + 2;
    return SQ_OK;
}

}  // namespace sq
