// server.hip — continuous-batching request loop over resident BVHs.
//
// One scheduler thread per server. Requests queue in arrival order; when a
// wave slot is free the scheduler asks the admission policy
// (server_policy.h) for the next wave: every queued request that fits the
// wave limits and the device-memory budget, evicting least-recently-used
// unreferenced scenes when room is needed. A wave runs on its slot's stream:
//   1. one host-to-device copy of the wave's inputs (new scenes' triangles,
//      ray lists, descriptor and job tables), staged in pinned memory,
//   2. the batched Morton pipeline over the new scenes (morton63),
//   3. one copy kernel moving each new scene's triangles and tree into its
//      resident block,
//   4. one persistent multi-job trace over every request of the wave,
//   5. one device-to-host copy of all outputs,
// then the scheduler copies each request's outputs to its destinations. Two
// slots alternate, so a wave is formed and uploaded while the previous one
// is still on the device. Waves never wait on each other on the device: a
// trace of a scene whose building wave is still in flight joins a later
// wave. All device memory comes from an arena of large chunks (no device
// frees while serving).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "server_policy.h"
#include "trace_common.h"

namespace sq {

struct CopyItem {
    const uint32_t* src;
    uint32_t* dst;
    uint64_t words;
};

// Copies every item of the list; each item is spread over the whole grid.
__global__ __launch_bounds__(256) void k_copy_list(const CopyItem* __restrict__ items, uint32_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t c = 0; c < n; c++) {
        const CopyItem it = items[c];
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < it.words; i += stride)
            it.dst[i] = it.src[i];
    }
}

}  // namespace sq

using namespace sq;

namespace {

using Clock = std::chrono::steady_clock;

double now_ms(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

constexpr uint32_t kMortonBits = 21;
constexpr uint64_t kSceneSlack = 512;

uint64_t scene_bytes(uint32_t n) { return Arena::round(36ull * n) + Arena::round(64ull * n) + kSceneSlack; }

// Upper bound of a request's transient device bytes in its wave (inputs,
// build workspace, outputs, tables).
uint64_t work_bytes(const sq_request& r, uint64_t rays, uint64_t pix) {
    uint64_t b = 8192;
    if (r.tris) b += 168ull * r.ntris + 65536;   // tris, keys, vals, temporary nodes, sort and emit workspace
    if (r.nrays) b += 32ull * r.nrays;           // ray list
    if (!(r.flags & SQ_REQ_DEVICE_OUT))          // outputs staged in the wave's block
        b += 8ull * rays + 4ull * (r.nrays ? rays : pix) + (r.nrays ? 0 : 3ull * pix) + 1024;
    return b;
}

struct Req {
    uint64_t id = 0;
    sq_request r{};
    uint64_t rays = 0, pix = 0;
    int state = SQ_REQ_QUEUED;
    int status = SQ_OK;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t wave = 0;
    // placement in its wave
    uint32_t scene_idx = 0;  // index among the wave's new scenes (build requests)
    uint32_t fault = 0;      // kFault* bits when a ray was abandoned
    size_t o_hit_id = 0, o_hit_t = 0, o_counts = 0, o_rgb = 0;  // offsets in the output region
};

struct Resident {
    uint64_t block = 0;  // arena address of [tris | nodes]
    sq_triangle* tris = nullptr;
    sq_node* nodes = nullptr;
    uint32_t n = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_done = nullptr;
    bool own_stream = false;
    char* pin_in = nullptr;
    size_t pin_in_bytes = 0;
    char* pin_out = nullptr;
    size_t pin_out_bytes = 0;
    uint32_t* spill = nullptr;
    bool busy = false;
    uint64_t wave = 0;
    std::vector<uint64_t> reqs;
    std::vector<uint64_t> keys;
    uint64_t work = 0;    // bytes the policy charged
    uint64_t block = 0;   // arena block of the wave
    size_t out_off = 0, out_bytes = 0;  // output region inside the block
    size_t fail_off = 0;                // per-request fault bits in the output region
    size_t dep_off = 0;                 // per-new-scene tree depth in the output region
    double t_launch = 0;                // host time the wave was enqueued
};

}  // namespace

struct sq_server {
    sq_ctx* ctx = nullptr;
    sq_server_config cfg{};
    std::unique_ptr<Arena> arena;
    std::unique_ptr<Admission> adm;
    std::vector<void*> chunks;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread thread;
    bool stop = false;
    std::unordered_map<uint64_t, Req> reqs;
    std::deque<PolicyRequest> pending;
    std::unordered_map<uint64_t, Resident> resident;
    std::vector<Slot> slots;
    TracePlan plan{};
    uint64_t next_id = 1, next_wave = 1;
    Clock::time_point t0;
    // statistics
    uint64_t submitted = 0, completed = 0, failed = 0, waves = 0, wave_reqs = 0, rays_traced = 0, tris_built = 0;
    double stage_in_ms = 0, stage_out_ms = 0, wave_ms = 0;  // host copy time, launch -> retire time
    std::vector<float> lat;
    double first_submit = -1, last_done = 0;
    uint64_t outstanding = 0;
    int fatal = SQ_OK;
};

namespace {

void finish(sq_server* s, Req& q, int status, double t) {
    q.state = status == SQ_OK ? SQ_REQ_DONE : SQ_REQ_FAILED;
    q.status = status;
    q.t_done = t;
    if (status == SQ_OK) {
        s->completed++;
        s->lat.push_back((float)(t - q.t_submit));
        s->rays_traced += q.rays;
        s->last_done = t;
    } else {
        s->failed++;
    }
    s->outstanding--;
}

int status_of(Verdict v) {
    switch (v) {
        case Verdict::Reject: return SQ_ERR_BUDGET;
        case Verdict::UnknownScene: return SQ_ERR_INVALID;
        case Verdict::Expired: return SQ_ERR_TIMEOUT;
        case Verdict::Cancelled: return SQ_ERR_CANCELLED;
        default: return SQ_OK;
    }
}

constexpr size_t kDefaultStage = (size_t)256 << 20;  // pinned input and output staging per slot

// exact: allocate `need` bytes (set-up); otherwise leave slack for later waves
int pinned_grow(char** p, size_t* have, size_t need, bool exact = false) {
    if (*have >= need) return SQ_OK;
    if (*p) (void)hipHostFree(*p);
    *p = nullptr;
    *have = 0;
    // generous slack: a pinned reallocation stalls the loop
    size_t sz = exact ? need : (need * 2 > ((size_t)64 << 20) ? need * 2 : ((size_t)64 << 20));
    SQ_CHECK_HIP(hipHostMalloc((void**)p, sz, hipHostMallocDefault));
    *have = sz;
    return SQ_OK;
}

void complete_wave(sq_server* s, Slot& sl, std::unique_lock<std::mutex>& lk);

// Called between pieces of a long input copy (lock released): retires the
// other slots' finished waves, so that a large upload does not hold back the
// completion of the wave already on the device.
void retire_others(sq_server* s, const Slot& self, std::unique_lock<std::mutex>& lk) {
    lk.lock();
    bool any = false;
    for (Slot& o : s->slots) {
        if (&o == &self || !o.busy) continue;
        if (hipEventQuery(o.ev_done) == hipSuccess) {
            complete_wave(s, o, lk);
            any = true;
        }
    }
    lk.unlock();
    if (any) s->cv_done.notify_all();
}

// Stages an input in pieces of kStagePiece bytes: each piece is copied into
// pinned memory and its host-to-device copy enqueued at once, so the DMA of
// one piece overlaps the host copy of the next; finished waves are retired
// between pieces.
constexpr size_t kStagePiece = 16u << 20;
int stage_copy(sq_server* s, const Slot& self, std::unique_lock<std::mutex>& lk, char* pinned, const void* src,
               size_t bytes, char* dev) {
    for (size_t off = 0; off < bytes; off += kStagePiece) {
        const size_t n = bytes - off < kStagePiece ? bytes - off : kStagePiece;
        par_memcpy(pinned + off, (const char*)src + off, n);
        SQ_CHECK_HIP(hipMemcpyAsync(dev + off, pinned + off, n, hipMemcpyHostToDevice, self.stream));
        if (off + n < bytes) retire_others(s, self, lk);
    }
    return SQ_OK;
}

// Lays out, stages and enqueues one wave on slot `sl`. Called with the lock
// held; returns an error only for device/runtime failures.
int launch_wave(sq_server* s, Slot& sl, const std::vector<uint64_t>& ids, uint64_t wave,
                std::unique_lock<std::mutex>& lk) {
    // ---- layout ----
    std::vector<Req*> rq;
    for (uint64_t id : ids) rq.push_back(&s->reqs[id]);
    std::vector<uint32_t> counts;
    uint64_t T = 0;
    for (Req* q : rq)
        if (q->r.tris) {
            q->scene_idx = (uint32_t)counts.size();
            counts.push_back(q->r.ntris);
            T += q->r.ntris;
        }
    const uint32_t S = (uint32_t)counts.size();
    size_t in_off = 0;
    auto take_in = [&](size_t bytes) { size_t o = in_off; in_off += Arena::round(bytes); return o; };
    const size_t o_tris = take_in(36 * T);
    size_t o_rays = in_off;
    uint64_t list_rays = 0;
    for (Req* q : rq)
        if (q->r.nrays) list_rays += q->r.nrays;
    in_off = o_rays + Arena::round(32 * list_rays);
    const size_t o_off = take_in(4ull * (S + 1));
    const size_t o_desc = take_in(sizeof(SceneDesc) * S);
    const size_t o_copy = take_in(sizeof(CopyItem) * 2 * S);
    const size_t o_jobs = take_in(sizeof(TraceArgs) * rq.size());
    const size_t o_jt = take_in(4ull * (rq.size() + 1));
    const size_t in_bytes = in_off;
    // device-only workspace
    size_t off = in_bytes;
    auto take = [&](size_t bytes) { size_t o = off; off += Arena::round(bytes); return o; };
    size_t o_keys[2] = {0, 0}, o_vals[2] = {0, 0}, o_hist = 0, o_ws = 0, o_ord = 0, o_nodes = 0;
    const uint32_t digit = 8;
    if (S) {
        o_keys[0] = take(8 * T);
        o_keys[1] = take(8 * T);
        o_vals[0] = take(4 * T);
        o_vals[1] = take(4 * T);
        o_hist = take(sort_hist_bytes((uint32_t)T, digit));
        o_ws = take(lbvh_batch_workspace((uint32_t)T, S, 0));
        o_ord = take(24ull * S);
        o_nodes = take(64 * T);
    }
    // outputs: work counters first
    const size_t out_off = off;
    size_t oo = 0;
    auto take_out = [&](size_t bytes) { size_t o = oo; oo += Arena::round(bytes); return o; };
    const size_t o_work = take_out(64);
    const size_t o_fail = take_out(4 * rq.size());
    const size_t o_dep = take_out(4ull * (S ? S : 1));  // per new scene: tree depth, 0 = build failed
    for (Req* q : rq) {
        const bool cam = q->r.nrays == 0, dev_out = (q->r.flags & SQ_REQ_DEVICE_OUT) != 0;
        q->o_hit_id = (q->r.out.hit_id && !dev_out) ? take_out(4 * q->rays) : SIZE_MAX;
        q->o_hit_t = (q->r.out.hit_t && !dev_out) ? take_out(4 * q->rays) : SIZE_MAX;
        q->o_counts = (q->r.out.counts && !dev_out) ? take_out(4 * (cam ? q->pix : q->rays)) : SIZE_MAX;
        q->o_rgb = (cam && q->r.out.rgb && !dev_out) ? take_out(3 * q->pix) : SIZE_MAX;
    }
    const size_t out_bytes = oo;
    const size_t total = out_off + out_bytes;
    uint64_t block = 0;
    if (!s->arena->alloc(total, &block)) SQ_RET(SQ_ERR_NOMEM, "server: device pool exhausted (%zu B)", total);
    char* dev = (char*)block;
    // new scenes' resident blocks
    std::vector<Resident*> nres(S, nullptr);
    for (Req* q : rq)
        if (q->r.tris) {
            Resident& R = s->resident[q->r.key];
            R.n = q->r.ntris;
            const uint64_t tb = Arena::round(36ull * R.n);
            if (!s->arena->alloc(tb + Arena::round(64ull * R.n), &R.block)) {
                R.block = 0;
                s->arena->release(block);
                SQ_RET(SQ_ERR_NOMEM, "server: device pool exhausted (scene of %u triangles)", R.n);
            }
            R.tris = (sq_triangle*)R.block;
            R.nodes = (sq_node*)(R.block + tb);
            nres[q->scene_idx] = &R;
        }
    struct BlockGuard {
        sq_server* s;
        uint64_t block;
        bool armed = true;
        ~BlockGuard() {
            if (armed) s->arena->release(block);
        }
    } guard{s, block};
    // ---- pinned input staging ----
    SQ_TRY(pinned_grow(&sl.pin_in, &sl.pin_in_bytes, in_bytes + 256));
    SQ_TRY(pinned_grow(&sl.pin_out, &sl.pin_out_bytes, out_bytes + 256));
    char* h = sl.pin_in;
    std::vector<uint32_t> offs(S + 1, 0);
    for (uint32_t i = 0; i < S; i++) offs[i + 1] = offs[i] + counts[i];
    size_t list_pos = 0;
    std::vector<size_t> list_off(rq.size(), 0);
    for (size_t k = 0; k < rq.size(); k++)
        if (rq[k]->r.nrays) {
            list_off[k] = list_pos;
            list_pos += rq[k]->r.nrays;
        }
    // The requests of this wave and their host buffers are stable while the
    // lock is released (only the scheduler touches admitted requests).
    const double t_in = now_ms(s->t0);
    lk.unlock();
    double retire_ms = 0;  // time spent retiring other waves between pieces
    int stage_rc = SQ_OK;
    for (size_t k = 0; k < rq.size() && stage_rc == SQ_OK; k++) {
        const Req* q = rq[k];
        const double r0 = s->stage_out_ms;
        if (q->r.tris) {
            const size_t o = o_tris + 36ull * offs[q->scene_idx];
            stage_rc = stage_copy(s, sl, lk, h + o, q->r.tris, 36ull * q->r.ntris, dev + o);
        }
        if (q->r.nrays && stage_rc == SQ_OK) {
            const size_t o = o_rays + 32ull * list_off[k];
            stage_rc = stage_copy(s, sl, lk, h + o, q->r.rays, 32ull * q->r.nrays, dev + o);
        }
        retire_ms += s->stage_out_ms - r0;
    }
    lk.lock();
    SQ_TRY(stage_rc);
    s->stage_in_ms += now_ms(s->t0) - t_in - retire_ms;
    std::memcpy(h + o_off, offs.data(), 4ull * (S + 1));
    std::vector<SceneDesc> desc(S);
    if (S) lbvh_fill_descs(desc.data(), counts.data(), S, 0);
    std::memcpy(h + o_desc, desc.data(), sizeof(SceneDesc) * S);
    std::vector<CopyItem> copies;
    for (uint32_t i = 0; i < S; i++) {
        copies.push_back({(const uint32_t*)(dev + o_tris + 36ull * offs[i]), (uint32_t*)nres[i]->tris, 9ull * counts[i]});
        copies.push_back({(const uint32_t*)(dev + o_nodes + 64ull * offs[i]), (uint32_t*)nres[i]->nodes,
                          16ull * counts[i]});
    }
    std::memcpy(h + o_copy, copies.data(), sizeof(CopyItem) * copies.size());
    std::vector<TraceArgs> jobs(rq.size());
    std::vector<uint32_t> jt(rq.size() + 1, 0);
    uint64_t tiles = 0;
    char* outd = dev + out_off;
    // ray lists take tiles of 64 k rays, k = min(4, ceil(wave rays / lanes)),
    // as a single list's launch does: a wave too short to give every wave of
    // the launch a 256-ray tile cuts its lists finer
    uint64_t wave_rays = 0;
    for (Req* q : rq) wave_rays += q->rays;
    const uint64_t lanes = (uint64_t)s->plan.grid * s->plan.block;
    const uint64_t kt = lanes ? (wave_rays + lanes - 1) / lanes : 4;
    const uint32_t list_tile = 64u * (uint32_t)(kt < 1 ? 1 : (kt > 4 ? 4 : kt));
    for (size_t k = 0; k < rq.size(); k++) {
        Req* q = rq[k];
        const Resident& R = s->resident[q->r.key];
        TraceArgs& a = jobs[k];
        a = TraceArgs{};
        a.nodes = R.nodes;
        a.tris = R.tris;
        a.root = 0;
        a.node_count = 2 * R.n;
        a.tri_count = R.n;
        if (q->r.nrays) {
            a.camera = 0;
            a.rays = (const sq_ray*)(dev + o_rays + 32ull * list_off[k]);
            a.nrays = q->r.nrays;
            a.ray_tile = list_tile;
            a.num_tiles = (q->r.nrays + list_tile - 1) / list_tile;
        } else {
            const sq_camera& c = q->r.camera;
            a.camera = 1;
            a.ox = c.origin.x; a.oy = c.origin.y; a.oz = c.origin.z;
            a.tanh = c.tan_half_fov;
            a.w = c.width; a.h = c.height; a.spp = c.spp; a.jitter = c.jitter; a.seed = c.seed;
            a.tmin = c.tmin; a.tmax = c.tmax;
            // a wave of one camera request tiles it as a launch of that camera
            // alone would; waves of several jobs take 8x8 tiles, as batches do
            camera_tiling(c.width, c.height, c.spp, rq.size() == 1 && !s->plan.closest, a);
        }
        if (q->r.flags & SQ_REQ_DEVICE_OUT) {
            a.hit_id = q->r.out.hit_id;
            a.hit_t = q->r.out.hit_t;
            a.counts = q->r.out.counts;
            a.rgb = q->r.nrays ? nullptr : q->r.out.rgb;
        } else {
            a.hit_id = q->o_hit_id != SIZE_MAX ? (uint32_t*)(outd + q->o_hit_id) : nullptr;
            a.hit_t = q->o_hit_t != SIZE_MAX ? (float*)(outd + q->o_hit_t) : nullptr;
            a.counts = q->o_counts != SIZE_MAX ? (uint32_t*)(outd + q->o_counts) : nullptr;
            a.rgb = q->o_rgb != SIZE_MAX ? (uint8_t*)(outd + q->o_rgb) : nullptr;
        }
        jt[k] = (uint32_t)tiles;
        tiles += a.num_tiles;
    }
    jt[rq.size()] = (uint32_t)tiles;
    std::memcpy(h + o_jobs, jobs.data(), sizeof(TraceArgs) * jobs.size());
    std::memcpy(h + o_jt, jt.data(), 4ull * jt.size());

    // ---- device work on the slot stream ----
    // Every scene a trace of this wave reads was built by a retired wave or
    // by this wave (the policy defers the rest), so the wave depends only on
    // earlier work of its own stream.
    hipStream_t st = sl.stream;
    // the triangles and ray lists went up piece by piece; the tables now
    SQ_CHECK_HIP(hipMemcpyAsync(dev + o_off, h + o_off, in_bytes - o_off, hipMemcpyHostToDevice, st));
    if (S) {
        BatchBuild j;
        j.tris = (const sq_triangle*)(dev + o_tris);
        j.total = (uint32_t)T;
        j.nscenes = S;
        j.d_off = (const uint32_t*)(dev + o_off);
        j.d_desc = (const SceneDesc*)(dev + o_desc);
        j.d_ord = (uint32_t*)(dev + o_ord);
        j.keys[0] = dev + o_keys[0];
        j.keys[1] = dev + o_keys[1];
        j.vals[0] = (uint32_t*)(dev + o_vals[0]);
        j.vals[1] = (uint32_t*)(dev + o_vals[1]);
        j.hist = (uint32_t*)(dev + o_hist);
        j.ws = dev + o_ws;
        j.d_depth = (uint32_t*)(dev + out_off + o_dep);
        j.nodes = (sq_node*)(dev + o_nodes);
        j.bits = kMortonBits;
        j.digit = digit;
        j.layout = SQ_LAYOUT_REFERENCE;
        j.threshold = 0;
        SQ_TRY(enqueue_batch_build(s->ctx, st, j));
        hipLaunchKernelGGL(k_copy_list, dim3(1024), dim3(256), 0, st, (const CopyItem*)(dev + o_copy),
                           (uint32_t)copies.size());
        SQ_CHECK_HIP(hipGetLastError());
    }
    TraceArgs a{};
    a.slots = s->plan.slots;
    a.spill = sl.spill;
    a.spill_cap = s->plan.spill_cap;
    a.work = (uint32_t*)(outd + o_work);
    a.stats = nullptr;
    a.num_tiles = (uint32_t)tiles;
    a.jobs = (const TraceArgs*)(dev + o_jobs);
    a.job_tiles = (const uint32_t*)(dev + o_jt);
    a.njobs = (uint32_t)rq.size();
    a.job_fail = (uint32_t*)(outd + o_fail);
    SQ_TRY(launch_fill_u32_on(st, s->ctx->num_cus, a.job_fail, 0u, rq.size()));
    bool narrow = !s->ctx->wide_fetch;
    for (const TraceArgs& j : jobs) narrow = narrow && narrow_ok(j.node_count, j.tri_count);
    SQ_TRY(trace_launch_plan(st, s->plan, a, true, narrow));
    SQ_CHECK_HIP(hipMemcpyAsync(sl.pin_out, outd, out_bytes, hipMemcpyDeviceToHost, st));
    SQ_CHECK_HIP(hipEventRecord(sl.ev_done, st));
    guard.armed = false;
    sl.t_launch = now_ms(s->t0);
    sl.busy = true;
    sl.wave = wave;
    sl.reqs = ids;
    sl.keys.clear();
    for (Req* q : rq) sl.keys.push_back(q->r.key);
    sl.block = block;
    sl.out_off = out_off;
    sl.out_bytes = out_bytes;
    sl.fail_off = o_fail;
    sl.dep_off = o_dep;
    for (Req* q : rq) {
        q->state = SQ_REQ_RUNNING;
        q->wave = wave;
        if (q->r.tris) s->tris_built += q->r.ntris;
    }
    s->waves++;
    s->wave_reqs += rq.size();
    return SQ_OK;
}

void release_evicted(sq_server* s, const std::vector<uint64_t>& keys) {
    for (uint64_t k : keys) {
        auto it = s->resident.find(k);
        if (it == s->resident.end()) continue;
        if (it->second.block) s->arena->release(it->second.block);
        s->resident.erase(it);
    }
}

// Copies a finished wave's outputs to the requests' destinations.
void complete_wave(sq_server* s, Slot& sl, std::unique_lock<std::mutex>& lk) {
    const uint32_t* fail = (const uint32_t*)(sl.pin_out + sl.fail_off);
    const uint32_t* depth = (const uint32_t*)(sl.pin_out + sl.dep_off);
    // a tree whose emission overflowed its workspace fails every request of the wave using it
    std::vector<uint64_t> bad_keys;
    for (uint64_t id : sl.reqs) {
        const Req& q = s->reqs[id];
        if (q.r.tris && depth[q.scene_idx] == 0) bad_keys.push_back(q.r.key);
    }
    std::vector<Req*> rq;
    std::vector<int> status;
    for (size_t k = 0; k < sl.reqs.size(); k++) {
        Req& q = s->reqs[sl.reqs[k]];
        q.fault = fail[k];
        const bool bad_tree = std::find(bad_keys.begin(), bad_keys.end(), q.r.key) != bad_keys.end();
        rq.push_back(&q);
        status.push_back((q.fault || bad_tree) ? SQ_ERR_OVERFLOW : SQ_OK);
    }
    const double t_out = now_ms(s->t0);
    s->wave_ms += t_out - sl.t_launch;
    lk.unlock();
    for (size_t k = 0; k < rq.size(); k++) {
        const Req& q = *rq[k];
        const bool cam = q.r.nrays == 0;
        if (status[k] != SQ_OK || (q.r.flags & SQ_REQ_DEVICE_OUT)) continue;
        if (q.r.out.hit_id) par_memcpy(q.r.out.hit_id, sl.pin_out + q.o_hit_id, 4 * q.rays);
        if (q.r.out.hit_t) par_memcpy(q.r.out.hit_t, sl.pin_out + q.o_hit_t, 4 * q.rays);
        if (q.r.out.counts) par_memcpy(q.r.out.counts, sl.pin_out + q.o_counts, 4 * (cam ? q.pix : q.rays));
        if (cam && q.r.out.rgb) par_memcpy(q.r.out.rgb, sl.pin_out + q.o_rgb, 3 * q.pix);
    }
    lk.lock();
    const double t = now_ms(s->t0);
    s->stage_out_ms += t - t_out;
    for (size_t k = 0; k < rq.size(); k++) finish(s, *rq[k], status[k], t);
    s->adm->wave_done(sl.wave, sl.work, sl.keys);
    for (uint64_t k : bad_keys)
        if (s->adm->evict(k)) release_evicted(s, {k});
    s->arena->release(sl.block);
    sl.busy = false;
    sl.reqs.clear();
    sl.keys.clear();
}

void fail_decision(sq_server* s, uint64_t id, int status, double t) {
    auto it = s->reqs.find(id);
    if (it != s->reqs.end() && it->second.state == SQ_REQ_QUEUED) finish(s, it->second, status, t);
}

void scheduler(sq_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progressed = false;
        // retire finished waves, oldest first
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            hipError_t q = hipEventQuery(sl.ev_done);
            if (q == hipSuccess) {
                complete_wave(s, sl, lk);
                progressed = true;
            } else if (q != hipErrorNotReady) {
                s->fatal = SQ_ERR_HIP;
                set_error("server: wave failed: %s", hipGetErrorString(q));
                for (uint64_t id : sl.reqs) finish(s, s->reqs[id], SQ_ERR_HIP, now_ms(s->t0));
                sl.busy = false;
                progressed = true;
            }
        }
        // form a wave for a free slot
        Slot* free_slot = nullptr;
        for (Slot& sl : s->slots)
            if (!sl.busy) { free_slot = &sl; break; }
        if (free_slot && !s->pending.empty() && s->fatal == SQ_OK) {
            const double t = now_ms(s->t0);
            const uint64_t wave = s->next_wave++;
            std::vector<uint64_t> evicted;
            auto dec = s->adm->form_wave(s->pending, t, wave, &evicted);
            release_evicted(s, evicted);
            std::vector<uint64_t> admit;
            uint64_t work = 0;
            for (auto& d : dec) {
                if (d.v == Verdict::Admit) {
                    admit.push_back(d.id);
                    Req& q = s->reqs[d.id];
                    q.t_admit = t;
                    work += work_bytes(q.r, q.rays, q.pix);
                } else {
                    fail_decision(s, d.id, status_of(d.v), t);
                }
            }
            if (!admit.empty()) {
                free_slot->work = work;
                int rc = launch_wave(s, *free_slot, admit, wave, lk);
                if (rc != SQ_OK) {
                    // nothing of the wave was enqueued: fail its requests and
                    // drop the scenes it would have built
                    std::vector<uint64_t> keys, built;
                    for (uint64_t id : admit) {
                        Req& q = s->reqs[id];
                        keys.push_back(q.r.key);
                        if (q.r.tris) built.push_back(q.r.key);
                        finish(s, q, rc, now_ms(s->t0));
                    }
                    s->adm->wave_done(wave, work, keys);
                    for (uint64_t k : built) s->adm->evict(k);
                    release_evicted(s, built);
                    free_slot->busy = false;
                }
            }
            progressed = progressed || !dec.empty();
        }
        if (progressed) s->cv_done.notify_all();
        bool any_busy = false;
        for (Slot& sl : s->slots) any_busy = any_busy || sl.busy;
        if (s->stop && !any_busy) break;
        if (!progressed) {
            if (any_busy) s->cv_work.wait_for(lk, std::chrono::microseconds(50));
            else s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

}  // namespace

extern "C" {

int sq_server_create(int device, const sq_server_config* cfg, sq_server** out) {
    if (!out) SQ_RET(SQ_ERR_INVALID, "sq_server_create: NULL out");
    *out = nullptr;
    sq_server_config c{};
    if (cfg) c = *cfg;
    if (!c.budget_bytes) c.budget_bytes = 200ull * 1000 * 1000 * 1000;
    if (!c.max_wave_requests) c.max_wave_requests = 256;
    if (!c.max_wave_rays) c.max_wave_rays = 16u << 20;
    if (!c.max_wave_tris) c.max_wave_tris = 8u << 20;
    if (!c.inflight_waves) c.inflight_waves = 2;
    if (c.inflight_waves > 2) SQ_RET(SQ_ERR_INVALID, "sq_server_create: at most 2 in-flight waves");
    if (c.semantics > SQ_TRACE_CLOSEST) SQ_RET(SQ_ERR_INVALID, "sq_server_create: semantics %u", c.semantics);
    auto s = std::make_unique<sq_server>();
    SQ_TRY(sq_ctx_create(device, &s->ctx));
    s->cfg = c;
    s->t0 = Clock::now();
    sq_server* raw = s.get();
    s->arena = std::make_unique<Arena>(4ull << 30, [raw](uint64_t bytes, uint64_t* base) {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return false;
        raw->chunks.push_back(p);
        *base = (uint64_t)p;
        return true;
    });
    // device pool reserved up front (growing it while serving stalls the
    // wave that needs the room)
    const uint64_t pool0 = (c.reserved[1] ? (uint64_t)c.reserved[1] : 16ull) << 30;
    if (!s->arena->reserve(pool0 < c.budget_bytes ? pool0 : c.budget_bytes)) {
        sq_server_destroy(s.release());
        SQ_RET(SQ_ERR_NOMEM, "server: could not reserve the initial device pool");
    }
    WaveLimits lim;
    lim.max_requests = c.max_wave_requests;
    lim.max_rays = c.max_wave_rays;
    lim.max_tris = c.max_wave_tris;
    lim.cost_ratio = 2;
    lim.min_wave_cost = 2u << 20;
    lim.group_backlog = 32;
    s->adm = std::make_unique<Admission>(c.budget_bytes, lim);
    sq_trace_params tp{};
    tp.semantics = c.semantics;
    int rc = trace_plan(s->ctx, &tp, morton_depth_bound(kMortonBits, 0x7FFFFFFFu), false, &s->plan);
    s->slots.resize(c.inflight_waves);
    // slot 0 runs on the context's stream, slot 1 on one more stream
    for (size_t i = 0; i < s->slots.size(); i++) {
        Slot& sl = s->slots[i];
        if (rc != SQ_OK) break;
        if (i == 0) {
            sl.stream = s->ctx->stream;
        } else if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess) {
            sl.own_stream = true;
        } else {
            set_error("server: stream setup failed");
            rc = SQ_ERR_HIP;
            break;
        }
        if (hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void**)&sl.spill, s->plan.spill_words * 4) != hipSuccess) {
            set_error("server: event/spill setup failed");
            rc = SQ_ERR_HIP;
            break;
        }
        // pinned staging sized up front: growing it while serving pins pages
        // with the scheduler's lock held and stalls every request
        const size_t stage = c.reserved[0] ? (size_t)c.reserved[0] << 20 : kDefaultStage;
        if (pinned_grow(&sl.pin_in, &sl.pin_in_bytes, stage, true) != SQ_OK ||
            pinned_grow(&sl.pin_out, &sl.pin_out_bytes, stage, true) != SQ_OK) {
            rc = SQ_ERR_NOMEM;
            break;
        }
    }
    if (rc != SQ_OK) {
        sq_server_destroy(s.release());
        return rc;
    }
    s->thread = std::thread(scheduler, raw);
    *out = s.release();
    return SQ_OK;
}

void sq_server_destroy(sq_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const double t = now_ms(s->t0);
        for (auto& p : s->pending) fail_decision(s, p.id, SQ_ERR_CANCELLED, t);
        s->pending.clear();
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->thread.joinable()) s->thread.join();
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
        if (sl.own_stream) (void)hipStreamDestroy(sl.stream);
        if (sl.spill) (void)hipFree(sl.spill);
        if (sl.pin_in) (void)hipHostFree(sl.pin_in);
        if (sl.pin_out) (void)hipHostFree(sl.pin_out);
    }
    s->arena.reset();
    for (void* p : s->chunks) (void)hipFree(p);
    if (s->ctx) sq_ctx_destroy(s->ctx);
    delete s;
}

int sq_server_submit(sq_server* s, const sq_request* r, uint64_t* id) {
    if (!s || !r || !id) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: NULL argument");
    if (r->tris && r->ntris == 0) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: empty scene");
    if (r->tris && (uint64_t)r->ntris * 2 + 1024 >= (1ull << 31)) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: scene too large");
    if (r->nrays && !r->rays) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: nrays without rays");
    uint64_t rays = 0, pix = 0;
    if (r->nrays) {
        rays = r->nrays;
    } else {
        const sq_camera& c = r->camera;
        if (c.width == 0 || c.height == 0 || c.spp == 0) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: empty image");
        pix = (uint64_t)c.width * c.height;
        rays = pix * c.spp;
        if (((uint64_t)(c.width + 3) / 4) * ((c.height + 3) / 4) >= (1ull << 31))
            SQ_RET(SQ_ERR_INVALID, "sq_server_submit: image too large");
    }
    if (rays >= (1ull << 32)) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: too many rays");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal != SQ_OK) SQ_RET(s->fatal, "sq_server_submit: server failed earlier");
        if (s->stop) SQ_RET(SQ_ERR_INVALID, "sq_server_submit: server is shutting down");
        const double t = now_ms(s->t0);
        Req q;
        q.id = s->next_id++;
        q.r = *r;
        q.rays = rays;
        q.pix = pix;
        q.t_submit = t;
        PolicyRequest p;
        p.id = q.id;
        p.key = r->key;
        p.build = r->tris != nullptr;
        p.scene_bytes = p.build ? scene_bytes(r->ntris) : 0;
        p.work_bytes = work_bytes(*r, rays, pix);
        p.tris = p.build ? r->ntris : 0;
        p.rays = rays;
        p.submit_ms = t;
        p.deadline_ms = r->timeout_ms > 0 ? t + r->timeout_ms : 0;
        s->reqs[q.id] = q;
        s->pending.push_back(p);
        s->submitted++;
        s->outstanding++;
        if (s->first_submit < 0) s->first_submit = t;
        *id = q.id;
    }
    s->cv_work.notify_all();
    return SQ_OK;
}

int sq_server_wait(sq_server* s, uint64_t id, float timeout_ms, sq_request_status* st) {
    if (!s) SQ_RET(SQ_ERR_INVALID, "sq_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->reqs.count(id)) SQ_RET(SQ_ERR_INVALID, "sq_server_wait: unknown request %llu", (unsigned long long)id);
    // the map may rehash while the lock is released: look the request up each time
    auto done = [&] {
        const Req& q = s->reqs.at(id);
        return q.state == SQ_REQ_DONE || q.state == SQ_REQ_FAILED;
    };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else s->cv_done.wait_for(lk, std::chrono::microseconds((int64_t)(timeout_ms * 1000.0f)), done);
    const bool fin = done();
    const Req& q = s->reqs.at(id);
    if (st) {
        st->state = q.state;
        st->status = q.status;
        st->latency_ms = fin ? (float)(q.t_done - q.t_submit) : 0.f;
        st->queue_ms = q.t_admit > 0 ? (float)(q.t_admit - q.t_submit) : 0.f;
        st->wave = q.wave;
        st->fault = q.fault;
        st->reserved = 0;
    }
    if (!fin) SQ_RET(SQ_ERR_TIMEOUT, "sq_server_wait: request %llu still running", (unsigned long long)id);
    s->reqs.erase(id);  // a finished request is reported once
    return SQ_OK;
}

int sq_server_cancel(sq_server* s, uint64_t id) {
    if (!s) SQ_RET(SQ_ERR_INVALID, "sq_server_cancel: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second.state != SQ_REQ_QUEUED)
        SQ_RET(SQ_ERR_INVALID, "sq_server_cancel: request %llu is not queued", (unsigned long long)id);
    for (auto& p : s->pending)
        if (p.id == id) p.cancelled = true;
    s->cv_work.notify_all();
    return SQ_OK;
}

int sq_server_drain(sq_server* s) {
    if (!s) SQ_RET(SQ_ERR_INVALID, "sq_server_drain: NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_work.notify_all();
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    return s->fatal;
}

int sq_server_evict(sq_server* s, uint64_t key) {
    if (!s) SQ_RET(SQ_ERR_INVALID, "sq_server_evict: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->adm->evict(key)) SQ_RET(SQ_ERR_INVALID, "sq_server_evict: scene %llu unknown or in use",
                                    (unsigned long long)key);
    release_evicted(s, {key});
    return SQ_OK;
}

int sq_server_get_stats(sq_server* s, sq_server_stats* o) {
    if (!s || !o) SQ_RET(SQ_ERR_INVALID, "sq_server_get_stats: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *o = sq_server_stats{};
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->failed = s->failed;
    o->waves = s->waves;
    o->evictions = s->adm->evictions();
    o->resident_scenes = s->adm->resident_count();
    o->resident_bytes = s->adm->resident_bytes();
    o->inflight_bytes = s->adm->inflight_bytes();
    o->pool_reserved_bytes = s->arena->reserved();
    o->rays_traced = s->rays_traced;
    o->tris_built = s->tris_built;
    if (!s->lat.empty()) {
        std::vector<float> v = s->lat;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            size_t k = (size_t)std::ceil(p * v.size()) - 1;
            return v[std::min(k, v.size() - 1)];
        };
        o->p50_ms = pct(0.50);
        o->p99_ms = pct(0.99);
        double sum = 0;
        for (float x : v) sum += x;
        o->mean_ms = (float)(sum / v.size());
        o->max_ms = v.back();
    }
    o->span_s = s->first_submit >= 0 ? (float)((s->last_done - s->first_submit) / 1000.0) : 0.f;
    o->mean_wave_requests = s->waves ? (float)s->wave_reqs / s->waves : 0.f;
    o->stage_in_ms = (float)s->stage_in_ms;
    o->stage_out_ms = (float)s->stage_out_ms;
    o->wave_ms = (float)s->wave_ms;
    return SQ_OK;
}

int sq_server_reset_stats(sq_server* s) {
    if (!s) SQ_RET(SQ_ERR_INVALID, "sq_server_reset_stats: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->completed = s->failed = s->waves = s->wave_reqs = s->rays_traced = s->tris_built = 0;
    s->submitted = s->outstanding;
    s->lat.clear();
    s->stage_in_ms = s->stage_out_ms = s->wave_ms = 0;
    s->first_submit = -1;
    s->last_done = 0;
    return SQ_OK;
}

}  // extern "C"
