// context.hip — contexts, device buffers, scenes, BVH objects, timing and the
// build dispatcher of the C ABI.
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

namespace sq {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int DevBuf::ensure(size_t n) {
    if (n <= bytes && p) return SQ_OK;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    size_t want = n < 256 ? 256 : n;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? SQ_ERR_NOMEM : SQ_ERR_HIP;
    }
    bytes = want;
    return SQ_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// Build phases are timed only on request (each event leaves the device idle
// for ~10 us between kernels); traces are always timed.
void phase_begin(sq_ctx* c, Phase p) {
    if (p != PH_TRACE && !c->phase_timing) return;
    (void)hipEventRecord(c->ev_begin[p], c->stream);
    c->phase_used[p] = true;
}
void phase_end(sq_ctx* c, Phase p) {
    if (p != PH_TRACE && !c->phase_timing) return;
    (void)hipEventRecord(c->ev_end[p], c->stream);
}
void phases_reset(sq_ctx* c) {
    for (int i = 0; i < PH_COUNT; i++) c->phase_used[i] = false;
    // calls on a ctx end with a sync, so no kernel of an earlier call can
    // still write the sort error word
    c->sort_err_pending = false;
    if (c->h_sort_err) *c->h_sort_err = 0;
}

int sort_error_word(sq_ctx* c, uint32_t** word) {
    if (!c->h_sort_err) {
        void* p = nullptr;
        SQ_CHECK_HIP(hipHostMalloc(&p, 64, hipHostMallocDefault));
        c->h_sort_err = (uint32_t*)p;
        *c->h_sort_err = 0;  // cleared again by phases_reset at the start of each call
    }
    c->sort_err_pending = true;
    *word = c->h_sort_err;
    return SQ_OK;
}

int phases_collect(sq_ctx* c) {
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (c->sort_err_pending) {
        c->sort_err_pending = false;
        if (*c->h_sort_err) SQ_RET(SQ_ERR_HIP, "radix sort: a look-back wait exceeded its spin limit");
    }
    float ms[PH_COUNT] = {};
    for (int i = 0; i < PH_COUNT; i++) {
        if (!c->phase_used[i]) continue;
        SQ_CHECK_HIP(hipEventElapsedTime(&ms[i], c->ev_begin[i], c->ev_end[i]));
    }
    sq_timings& t = c->last;
    if (c->phase_used[PH_BOUNDS]) t.bounds_ms = ms[PH_BOUNDS];
    if (c->phase_used[PH_MORTON]) t.morton_ms = ms[PH_MORTON];
    if (c->phase_used[PH_SORT]) t.sort_ms = ms[PH_SORT];
    if (c->phase_used[PH_EMIT]) t.emit_ms = ms[PH_EMIT];
    if (c->phase_used[PH_REFIT]) t.refit_ms = ms[PH_REFIT];
    if (c->phase_used[PH_SAH]) t.sah_ms = ms[PH_SAH];
    if (c->phase_used[PH_GRID]) t.grid_ms = ms[PH_GRID];
    if (c->phase_used[PH_TRACE]) t.trace_ms = ms[PH_TRACE];
    return SQ_OK;
}

int pinned_ensure(sq_ctx* c, size_t bytes) {
    if (c->pinned_bytes >= bytes) return SQ_OK;
    if (c->pinned) (void)hipHostFree(c->pinned);
    c->pinned = nullptr;
    c->pinned_bytes = 0;
    SQ_CHECK_HIP(hipHostMalloc(&c->pinned, bytes, hipHostMallocDefault));
    c->pinned_bytes = bytes;
    return SQ_OK;
}

__global__ void k_fill_u32(uint32_t* __restrict__ dst, uint32_t value, size_t words) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < words; i += stride) dst[i] = value;
}

int launch_fill_u32_on(hipStream_t stream, int num_cus, void* dst, uint32_t value, size_t words) {
    if (words == 0) return SQ_OK;
    size_t blocks = (words + 255) / 256;
    size_t cap = (size_t)num_cus * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, stream, (uint32_t*)dst, value, words);
    SQ_CHECK_HIP(hipGetLastError());
    return SQ_OK;
}

int launch_fill_u32(sq_ctx* c, void* dst, uint32_t value, size_t words) {
    return launch_fill_u32_on(c->stream, c->num_cus, dst, value, words);
}

__global__ void k_copy_u32x4(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace sq

using namespace sq;

extern "C" {

const char* sq_last_error(void) { return g_err; }
const char* sq_version(void) { return "sequoia-amd 0.1 (gfx950)"; }

int sq_ctx_create(int device, sq_ctx** out) {
    if (!out) SQ_RET(SQ_ERR_INVALID, "sq_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    SQ_CHECK_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) SQ_RET(SQ_ERR_INVALID, "sq_ctx_create: device %d of %d", device, ndev);
    SQ_CHECK_HIP(hipSetDevice(device));
    sq_ctx* c = new sq_ctx();
    c->device = device;
    if (const char* e = std::getenv("SEQUOIA_TEST_POISON_NODES")) c->poison_nodes = std::strtoul(e, nullptr, 10) != 0;
    if (const char* e = std::getenv("SEQUOIA_TEST_WIDE_FETCH")) c->wide_fetch = std::strtoul(e, nullptr, 10) != 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        SQ_RET(SQ_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    // Timing events release to device scope only: a system-scope release
    // writes the whole L2 back, which costs tens of microseconds and would be
    // charged to the kernel before the event.
    for (int i = 0; i < PH_COUNT; i++) {
        (void)hipEventCreateWithFlags(&c->ev_begin[i], hipEventReleaseToDevice);
        (void)hipEventCreateWithFlags(&c->ev_end[i], hipEventReleaseToDevice);
    }
    // Pinned status words (trace fault bits and work totals are read back
    // through them) and the allocation-free entry points' scratch address
    // on this device: both set up here, so later calls need not allocate.
    int rc = pinned_ensure(c, 4096);
    if (rc == SQ_OK) {
        // the bounds kernel's ticket starts at zero and every launch leaves it so
        e = hipMalloc((void**)&c->d_bounds_ticket, 256);
        if (e == hipSuccess) e = hipMemset(c->d_bounds_ticket, 0, 256);
        if (e != hipSuccess) {
            set_error("bounds ticket: %s", hipGetErrorString(e));
            rc = SQ_ERR_HIP;
        }
    }
    if (rc == SQ_OK) {
        e = hipHostMalloc((void**)&c->h_mail, 256, hipHostMallocCoherent);
        if (e == hipSuccess) std::memset(c->h_mail, 0, 256);
        else {
            c->h_mail = nullptr;
            set_error("level mailbox: %s", hipGetErrorString(e));
            rc = SQ_ERR_HIP;
        }
    }
    if (rc == SQ_OK) rc = noalloc_setup(device);
    if (rc != SQ_OK) {
        sq_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return SQ_OK;
}

void sq_ctx_destroy(sq_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto* b : {&c->keys[0], &c->keys[1], &c->vals[0], &c->vals[1], &c->hist, &c->scratch, &c->tasks[0],
                    &c->tasks[1], &c->tasks[2], &c->counters, &c->spill, &c->stats, &c->aux[0], &c->aux[1],
                    &c->aux[2], &c->aux[3], &c->sah[0], &c->sah[1], &c->sah[2], &c->sah[3], &c->sah[4], &c->sah[5],
                    &c->sort_os})
        b->release();
    if (c->h_sort_err) (void)hipHostFree(c->h_sort_err);
    if (c->h_mail) (void)hipHostFree(c->h_mail);
    for (int i = 0; i < PH_COUNT; i++) {
        (void)hipEventDestroy(c->ev_begin[i]);
        (void)hipEventDestroy(c->ev_end[i]);
    }
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->d_bounds_ticket) (void)hipFree(c->d_bounds_ticket);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int sq_ctx_sync(sq_ctx* c) {
    if (!c) SQ_RET(SQ_ERR_INVALID, "sq_ctx_sync: ctx is NULL");
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

void* sq_ctx_stream(sq_ctx* c) { return c ? (void*)c->stream : nullptr; }

int sq_ctx_set_phase_timing(sq_ctx* c, int enable) {
    if (!c) SQ_RET(SQ_ERR_INVALID, "sq_ctx_set_phase_timing: NULL context");
    c->phase_timing = enable != 0;
    return SQ_OK;
}

int sq_get_timings(sq_ctx* c, sq_timings* out) {
    if (!c || !out) SQ_RET(SQ_ERR_INVALID, "sq_get_timings: NULL argument");
    *out = c->last;
    return SQ_OK;
}

int sq_device_alloc(sq_ctx* c, size_t bytes, void** out) {
    if (!c || !out) SQ_RET(SQ_ERR_INVALID, "sq_device_alloc: NULL argument");
    *out = nullptr;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    SQ_CHECK_HIP(hipMalloc(out, bytes ? bytes : 1));
    return SQ_OK;
}

void sq_device_free(sq_ctx* c, void* p) {
    if (c) (void)hipSetDevice(c->device);
    if (p) (void)hipFree(p);
}

int sq_device_copy(sq_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src || (bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15))
        SQ_RET(SQ_ERR_INVALID, "sq_device_copy: pointers and size must be 16-byte aligned");
    size_t n16 = bytes / 16, blocks = (n16 + 255) / 256, cap = (size_t)c->num_cus * 16;
    if (blocks > cap) blocks = cap;
    if (blocks == 0) return SQ_OK;
    hipLaunchKernelGGL(k_copy_u32x4, dim3((uint32_t)blocks), dim3(256), 0, c->stream, (uint4*)dst, (const uint4*)src, n16);
    SQ_CHECK_HIP(hipGetLastError());
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

int sq_copy_to_host(sq_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) SQ_RET(SQ_ERR_INVALID, "sq_copy_to_host: NULL argument");
    SQ_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

int sq_copy_to_device(sq_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) SQ_RET(SQ_ERR_INVALID, "sq_copy_to_device: NULL argument");
    SQ_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

int sq_scene_upload(sq_ctx* c, const sq_triangle* host, uint32_t n, sq_scene** out) {
    if (!c || !host || !out || n == 0) SQ_RET(SQ_ERR_INVALID, "sq_scene_upload: bad argument");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    sq_scene* s = new sq_scene();
    s->ctx = c;
    s->n = n;
    s->owned = true;
    hipError_t e = hipMalloc((void**)&s->d_tris, sizeof(sq_triangle) * (size_t)n);
    if (e != hipSuccess) {
        delete s;
        SQ_RET(e == hipErrorOutOfMemory ? SQ_ERR_NOMEM : SQ_ERR_HIP, "scene alloc: %s", hipGetErrorString(e));
    }
    e = hipMemcpyAsync(s->d_tris, host, sizeof(sq_triangle) * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(s->d_tris);
        delete s;
        SQ_RET(SQ_ERR_HIP, "scene upload: %s", hipGetErrorString(e));
    }
    *out = s;
    return SQ_OK;
}

int sq_scene_wrap_device(sq_ctx* c, const sq_triangle* dev, uint32_t n, sq_scene** out) {
    if (!c || !dev || !out || n == 0) SQ_RET(SQ_ERR_INVALID, "sq_scene_wrap_device: bad argument");
    sq_scene* s = new sq_scene();
    s->ctx = c;
    s->d_tris = (sq_triangle*)dev;
    s->n = n;
    s->owned = false;
    *out = s;
    return SQ_OK;
}

void sq_scene_free(sq_scene* s) {
    if (!s) return;
    if (s->owned && s->d_tris) (void)hipFree(s->d_tris);
    delete s;
}

const sq_triangle* sq_scene_device_ptr(const sq_scene* s) { return s ? s->d_tris : nullptr; }
uint32_t sq_scene_count(const sq_scene* s) { return s ? s->n : 0; }

static int default_bits(sq_builder kind) {
    switch (kind) {
        case SQ_BUILDER_MORTON: return 10;
        case SQ_BUILDER_MORTON2: return 20;
        case SQ_BUILDER_MORTON63: return 21;
        default: return 0;
    }
}

static int build_morton(sq_ctx* c, const sq_scene* s, sq_builder kind, const sq_build_params* p, sq_bvh* b) {
    const uint32_t n = s->n;
    uint32_t bits = p && p->morton_bits_per_axis ? p->morton_bits_per_axis : (uint32_t)default_bits(kind);
    if (bits < 1 || bits > 21) SQ_RET(SQ_ERR_INVALID, "morton bits per axis %u outside 1..21", bits);
    uint32_t digit = p && p->radix_digit_bits ? p->radix_digit_bits : 8;
    if (digit < 1 || digit > 8) SQ_RET(SQ_ERR_INVALID, "radix digit bits %u outside 1..8", digit);
    uint32_t layout = p ? p->layout : SQ_LAYOUT_REFERENCE;
    bool refit = !(p && p->no_refit);
    bool key64 = bits > 10;
    size_t kb = key64 ? 8 : 4;
    SQ_TRY(c->keys[0].ensure(kb * n));
    SQ_TRY(c->keys[1].ensure(kb * n));
    SQ_TRY(c->vals[0].ensure(4ull * n));
    SQ_TRY(c->vals[1].ensure(4ull * n));
    SQ_TRY(c->aux[0].ensure(scene_bounds_bytes(c, n)));

    phase_begin(c, PH_BOUNDS);
    SQ_TRY(scene_bounds(c, s->d_tris, n, c->aux[0].as<uint32_t>()));
    phase_end(c, PH_BOUNDS);

    phase_begin(c, PH_MORTON);
    SQ_TRY(morton_codes(c, s->d_tris, n, bits, c->aux[0].as<uint32_t>(), c->keys[0].p, c->vals[0].as<uint32_t>(), key64));
    phase_end(c, PH_MORTON);

    phase_begin(c, PH_SORT);
    int slot = 0;
    SQ_TRY(radix_sort(c, n, 3 * bits, digit, key64, 0, &slot));
    phase_end(c, PH_SORT);

    phase_begin(c, PH_EMIT);
    uint32_t depth = 0;
    uint32_t thr = p && p->subtree_threshold ? p->subtree_threshold : 0;
    SQ_TRY(lbvh_emit(c, s->d_tris, n, c->keys[slot].p, c->vals[slot].as<uint32_t>(), key64, 3 * bits, layout, refit, thr,
                     b->d_nodes, &depth));
    if (!refit) phase_end(c, PH_EMIT);  // with refit, the emission ends the phase before the refit kernel
    b->root = 0;
    b->depth = depth;
    return SQ_OK;
}

// Allocates b's node array of `count` slots; odd_lists: child lists start at
// odd slots (shift the array by one slot so they start 64-byte aligned).
static hipError_t alloc_nodes(sq_bvh* b, size_t count, bool odd_lists) {
    hipError_t e = hipMalloc(&b->d_alloc, sizeof(sq_node) * (count + 1));
    if (e != hipSuccess) {
        b->d_alloc = nullptr;
        return e;
    }
    b->d_nodes = (sq_node*)b->d_alloc + (odd_lists ? 1 : 0);
    return hipSuccess;
}

int sq_build(sq_ctx* c, const sq_scene* s, sq_builder kind, const sq_build_params* p, sq_bvh** out) {
    if (!c || !s || !out) SQ_RET(SQ_ERR_INVALID, "sq_build: NULL argument");
    *out = nullptr;
    if (s->n == 0) SQ_RET(SQ_ERR_INVALID, "sq_build: empty scene");
    if ((size_t)s->n * 2 + 1024 >= (1ull << 31)) SQ_RET(SQ_ERR_INVALID, "sq_build: %u triangles is too many", s->n);
    SQ_CHECK_HIP(hipSetDevice(c->device));
    uint32_t dim = p && p->grid_dim ? p->grid_dim : 4;
    bool grid = kind == SQ_BUILDER_GRID || kind == SQ_BUILDER_GRID_SSE || kind == SQ_BUILDER_GRID_AVX;
    uint32_t cap = 2 * s->n + (grid ? 2 * dim * dim * dim : 0);
    sq_bvh* b = new sq_bvh();
    b->ctx = c;
    b->scene = s;
    b->capacity = cap;
    const bool morton = kind == SQ_BUILDER_MORTON || kind == SQ_BUILDER_MORTON2 || kind == SQ_BUILDER_MORTON63;
    hipError_t e = alloc_nodes(b, cap, morton);
    if (e != hipSuccess) {
        delete b;
        SQ_RET(e == hipErrorOutOfMemory ? SQ_ERR_NOMEM : SQ_ERR_HIP, "node alloc: %s", hipGetErrorString(e));
    }
    phases_reset(c);
    c->last = sq_timings{};
    hipEvent_t t0, t1;
    (void)hipEventCreateWithFlags(&t0, hipEventReleaseToDevice);
    (void)hipEventCreateWithFlags(&t1, hipEventReleaseToDevice);
    (void)hipEventRecord(t0, c->stream);
    c->build_end = t1;
    c->build_end_recorded = false;
    // Morton builds with refit write every slot of the array (unused ones as
    // zeros); the other builders and links-only Morton builds start from zeros
    const bool writes_all = morton && !(p && p->no_refit);
    int rc = writes_all ? (c->poison_nodes ? launch_fill_u32(c, b->d_nodes, 0xFFFFFFFFu, (size_t)cap * 8) : SQ_OK)
                        : launch_fill_u32(c, b->d_nodes, 0u, (size_t)cap * 8);
    if (rc == SQ_OK) {
        switch (kind) {
            case SQ_BUILDER_MORTON:
            case SQ_BUILDER_MORTON2:
            case SQ_BUILDER_MORTON63: rc = build_morton(c, s, kind, p, b); break;
            case SQ_BUILDER_RECURSIVE:
            case SQ_BUILDER_HORIZONTAL:
            case SQ_BUILDER_GRID:
            case SQ_BUILDER_GRID_SSE:
            case SQ_BUILDER_GRID_AVX: rc = sah_build(c, s->d_tris, s->n, p, kind, b->d_nodes, &b->root, &b->depth); break;
            default: rc = SQ_ERR_INVALID; set_error("sq_build: unknown builder %d", (int)kind); break;
        }
    }
    if (!c->build_end_recorded) (void)hipEventRecord(t1, c->stream);
    c->build_end = nullptr;
    if (rc == SQ_OK) rc = phases_collect(c);
    if (rc == SQ_OK) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t0, t1);
        c->last.total_build_ms = ms;
        c->last.tree_depth = b->depth;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rc != SQ_OK) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(b->d_alloc);
        delete b;
        return rc;
    }
    *out = b;
    return SQ_OK;
}

static uint32_t host_depth(const sq_node* nodes, uint32_t count, uint32_t root) {
    // Iterative depth of the reachable tree; cycles are cut by a visit bound.
    std::vector<std::pair<uint32_t, uint32_t>> st;
    st.push_back({root, 1});
    uint32_t depth = 0;
    size_t visits = 0, limit = (size_t)count * 4 + 16;
    while (!st.empty() && visits++ < limit) {
        auto [i, d] = st.back();
        st.pop_back();
        if (i >= count) continue;
        if (d > depth) depth = d;
        const sq_node& nd = nodes[i];
        if (!(nd.flags & SQ_NODE_LEAF))
            for (uint32_t j = 0; j < nd.count; j++) st.push_back({nd.child + j, d + 1});
    }
    return depth;
}

int sq_bvh_from_host(sq_ctx* c, const sq_scene* s, const sq_node* nodes, uint32_t count, uint32_t root, sq_bvh** out) {
    if (!c || !s || !nodes || !out || root >= count) SQ_RET(SQ_ERR_INVALID, "sq_bvh_from_host: bad argument");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    sq_bvh* b = new sq_bvh();
    b->ctx = c;
    b->scene = s;
    b->capacity = count;
    b->root = root;
    b->depth = host_depth(nodes, count, root);
    size_t odd = 0, lists = 0;
    for (uint32_t i = 0; i < count; i++)
        if (!(nodes[i].flags & SQ_NODE_LEAF) && nodes[i].count > 0) {
            lists++;
            odd += nodes[i].child & 1u;
        }
    hipError_t e = alloc_nodes(b, count, 2 * odd > lists);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_nodes, nodes, sizeof(sq_node) * (size_t)count, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        if (b->d_alloc) (void)hipFree(b->d_alloc);
        delete b;
        SQ_RET(SQ_ERR_HIP, "sq_bvh_from_host: %s", hipGetErrorString(e));
    }
    *out = b;
    return SQ_OK;
}

int sq_bvh_info(const sq_bvh* b, uint32_t* root, uint32_t* cap, uint32_t* depth) {
    if (!b) SQ_RET(SQ_ERR_INVALID, "sq_bvh_info: NULL bvh");
    if (root) *root = b->root;
    if (cap) *cap = b->capacity;
    if (depth) *depth = b->depth;
    return SQ_OK;
}

int sq_bvh_download(const sq_bvh* b, sq_node* host, uint32_t capacity) {
    if (!b || !host || capacity < b->capacity) SQ_RET(SQ_ERR_INVALID, "sq_bvh_download: bad argument");
    sq_ctx* c = b->ctx;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    SQ_CHECK_HIP(hipMemcpyAsync(host, b->d_nodes, sizeof(sq_node) * (size_t)b->capacity, hipMemcpyDeviceToHost,
                                c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

const sq_node* sq_bvh_device_nodes(const sq_bvh* b) { return b ? b->d_nodes : nullptr; }

void sq_bvh_free(sq_bvh* b) {
    if (!b || b->view) return;
    if (b->d_alloc) (void)hipFree(b->d_alloc);
    if (b->d_plucker) (void)hipFree(b->d_plucker);
    delete b;
}

int sq_morton_codes(sq_ctx* c, const sq_scene* s, uint32_t bits, uint64_t* host_codes) {
    if (!c || !s || !host_codes || bits < 1 || bits > 21) SQ_RET(SQ_ERR_INVALID, "sq_morton_codes: bad argument");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    const uint32_t n = s->n;
    SQ_TRY(c->keys[0].ensure(8ull * n));
    SQ_TRY(c->vals[0].ensure(4ull * n));
    SQ_TRY(c->aux[0].ensure(scene_bounds_bytes(c, n)));
    SQ_TRY(scene_bounds(c, s->d_tris, n, c->aux[0].as<uint32_t>()));
    SQ_TRY(morton_codes(c, s->d_tris, n, bits, c->aux[0].as<uint32_t>(), c->keys[0].p, c->vals[0].as<uint32_t>(), true));
    SQ_CHECK_HIP(hipMemcpyAsync(host_codes, c->keys[0].p, 8ull * n, hipMemcpyDeviceToHost, c->stream));
    SQ_CHECK_HIP(hipStreamSynchronize(c->stream));
    return SQ_OK;
}

int sq_radix_sort_pairs(sq_ctx* c, uint64_t* keys, uint32_t* vals, uint32_t n, uint32_t key_bits, uint32_t digit) {
    if (!c || !keys || !vals || key_bits < 1 || key_bits > 64 || digit < 1 || digit > 8)
        SQ_RET(SQ_ERR_INVALID, "sq_radix_sort_pairs: bad argument");
    if (n == 0) return SQ_OK;
    SQ_CHECK_HIP(hipSetDevice(c->device));
    bool key64 = key_bits > 32;
    size_t kb = key64 ? 8 : 4;
    SQ_TRY(c->keys[0].ensure(kb * n));
    SQ_TRY(c->keys[1].ensure(kb * n));
    SQ_TRY(c->vals[0].ensure(4ull * n));
    SQ_TRY(c->vals[1].ensure(4ull * n));
    std::vector<uint32_t> k32;
    if (key64) {
        SQ_CHECK_HIP(hipMemcpyAsync(c->keys[0].p, keys, 8ull * n, hipMemcpyHostToDevice, c->stream));
    } else {
        k32.resize(n);
        for (uint32_t i = 0; i < n; i++) k32[i] = (uint32_t)keys[i];
        SQ_CHECK_HIP(hipMemcpyAsync(c->keys[0].p, k32.data(), 4ull * n, hipMemcpyHostToDevice, c->stream));
    }
    SQ_CHECK_HIP(hipMemcpyAsync(c->vals[0].p, vals, 4ull * n, hipMemcpyHostToDevice, c->stream));
    int slot = 0;
    phases_reset(c);
    phase_begin(c, PH_SORT);
    SQ_TRY(radix_sort(c, n, key_bits, digit, key64, 0, &slot));
    phase_end(c, PH_SORT);
    if (key64) {
        SQ_CHECK_HIP(hipMemcpyAsync(keys, c->keys[slot].p, 8ull * n, hipMemcpyDeviceToHost, c->stream));
    } else {
        SQ_CHECK_HIP(hipMemcpyAsync(k32.data(), c->keys[slot].p, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    }
    SQ_CHECK_HIP(hipMemcpyAsync(vals, c->vals[slot].p, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    SQ_TRY(phases_collect(c));
    if (!key64)
        for (uint32_t i = 0; i < n; i++) keys[i] = k32[i];
    return SQ_OK;
}

int sq_refit(sq_ctx* c, sq_bvh* b) {
    if (!c || !b) SQ_RET(SQ_ERR_INVALID, "sq_refit: NULL argument");
    SQ_CHECK_HIP(hipSetDevice(c->device));
    phases_reset(c);
    phase_begin(c, PH_REFIT);
    SQ_TRY(refit_tree(c, b));
    phase_end(c, PH_REFIT);
    return phases_collect(c);
}

}  // extern "C"
