// server_policy.h — host-side bookkeeping of the serving loop: the device
// memory arena and the admission / residency policy. No HIP calls: chunk
// memory comes from a callback, so this logic is exercised on the CPU by the
// tests and driven by the GPU server in server.hip.
#pragma once
#include <cstdint>
#include <deque>
#include <functional>
#include <list>
#include <map>
#include <unordered_map>
#include <vector>

namespace sq {

// First-fit sub-allocator over large chunks obtained once from `grow` and
// kept until destruction (no device frees in steady state). Blocks are
// 256-byte aligned; freeing coalesces with free neighbours.
class Arena {
   public:
    using Grow = std::function<bool(uint64_t bytes, uint64_t* base)>;
    Arena(uint64_t chunk_bytes, Grow grow) : chunk_(chunk_bytes), grow_(std::move(grow)) {}

    static uint64_t round(uint64_t b) { return (b + 255) & ~uint64_t(255); }

    // Returns false when no chunk has room and a new chunk cannot be made.
    bool alloc(uint64_t bytes, uint64_t* addr) {
        const uint64_t need = round(bytes ? bytes : 1);
        for (auto& c : chunks_) {
            for (auto it = c.free.begin(); it != c.free.end(); ++it) {
                if (it->second >= need) {
                    const uint64_t off = it->first, left = it->second - need;
                    c.free.erase(it);
                    if (left) c.free[off + need] = left;
                    c.used[off] = need;
                    used_ += need;
                    *addr = c.base + off;
                    return true;
                }
            }
        }
        const uint64_t size = need > chunk_ ? need : chunk_;
        uint64_t base = 0;
        if (!grow_(size, &base)) return false;
        chunks_.push_back(Chunk{base, size, {}, {}});
        Chunk& c = chunks_.back();
        if (size > need) c.free[need] = size - need;
        c.used[0] = need;
        used_ += need;
        reserved_ += size;
        *addr = base;
        return true;
    }

    // Obtains empty chunks until `bytes` are reserved (set-up: later waves
    // then find room without growing). False if a chunk cannot be made.
    bool reserve(uint64_t bytes) {
        while (reserved_ < bytes) {
            uint64_t base = 0;
            if (!grow_(chunk_, &base)) return false;
            chunks_.push_back(Chunk{base, chunk_, {}, {}});
            chunks_.back().free[0] = chunk_;
            reserved_ += chunk_;
        }
        return true;
    }

    bool release(uint64_t addr) {
        for (auto& c : chunks_) {
            if (addr < c.base || addr >= c.base + c.size) continue;
            const uint64_t off = addr - c.base;
            auto u = c.used.find(off);
            if (u == c.used.end()) return false;
            uint64_t start = off, len = u->second;
            used_ -= len;
            c.used.erase(u);
            auto nx = c.free.lower_bound(start);
            if (nx != c.free.end() && nx->first == start + len) {
                len += nx->second;
                nx = c.free.erase(nx);
            }
            if (nx != c.free.begin()) {
                auto pv = std::prev(nx);
                if (pv->first + pv->second == start) {
                    start = pv->first;
                    len += pv->second;
                    c.free.erase(pv);
                }
            }
            c.free[start] = len;
            return true;
        }
        return false;
    }

    uint64_t used() const { return used_; }
    uint64_t reserved() const { return reserved_; }
    size_t chunks() const { return chunks_.size(); }
    template <typename F>
    void for_each_chunk(F f) const {
        for (auto& c : chunks_) f(c.base, c.size);
    }
    // Largest free block in any chunk.
    uint64_t largest_free() const {
        uint64_t m = 0;
        for (auto& c : chunks_)
            for (auto& f : c.free) m = f.second > m ? f.second : m;
        return m;
    }

   private:
    struct Chunk {
        uint64_t base, size;
        std::map<uint64_t, uint64_t> free;  // offset -> length
        std::map<uint64_t, uint64_t> used;  // offset -> length
    };
    uint64_t chunk_;
    Grow grow_;
    std::list<Chunk> chunks_;
    uint64_t used_ = 0, reserved_ = 0;
};

// Request as the policy sees it.
struct PolicyRequest {
    uint64_t id = 0;
    uint64_t key = 0;        // scene key
    bool build = false;      // build + trace (false: trace-only on a resident scene)
    uint64_t scene_bytes = 0;  // resident footprint of a scene it builds
    uint64_t work_bytes = 0;   // transient device bytes of its wave share
    uint64_t tris = 0, rays = 0;
    double submit_ms = 0, deadline_ms = 0;  // deadline 0 = none
    bool cancelled = false;
};

enum class Verdict { Admit, Wait, Reject, UnknownScene, Expired, Cancelled };

struct WaveLimits {
    uint32_t max_requests = 256;
    uint64_t max_rays = 64ull << 20;
    uint64_t max_tris = 32ull << 20;
    // Size grouping (cost_ratio 0: off). A request's cost is rays + 2 x
    // triangles built; once a wave holds a request, later ones join only while
    // the wave's cost stays within max(cost_ratio x the first one's cost,
    // min_wave_cost). A request that does not fit stays queued for a later
    // wave, so small requests are not held back behind a large one in its
    // wave and a large one runs in a wave of its own.
    uint32_t cost_ratio = 0;
    uint64_t min_wave_cost = 0;
    // Size grouping applies while at most this many requests are queued
    // (0: always). Past it the queue is a backlog, and waves take every
    // request that fits the limits: fewer, larger waves raise throughput when
    // latency is already set by the queue.
    uint32_t group_backlog = 0;
};

inline uint64_t request_cost(const PolicyRequest& r) { return r.rays + (r.build ? 2 * r.tris : 0); }

// Scene residency with LRU eviction of unreferenced scenes, and wave
// admission under a byte budget (resident scenes + in-flight waves).
class Admission {
   public:
    explicit Admission(uint64_t budget, WaveLimits lim = WaveLimits()) : budget_(budget), lim_(lim) {}

    struct Scene {
        uint64_t bytes = 0;
        uint32_t refs = 0;     // in-flight requests using it
        bool ready = false;    // built (false while its wave runs)
        uint64_t wave = 0;     // wave that built it
        std::list<uint64_t>::iterator lru;
    };

    bool resident(uint64_t key) const { return scenes_.count(key) != 0; }
    const Scene* scene(uint64_t key) const {
        auto it = scenes_.find(key);
        return it == scenes_.end() ? nullptr : &it->second;
    }
    uint64_t resident_bytes() const { return resident_; }
    uint64_t inflight_bytes() const { return inflight_; }
    uint64_t budget() const { return budget_; }
    size_t resident_count() const { return scenes_.size(); }
    uint64_t evictions() const { return evictions_; }

    // Forms one wave from the FIFO `pending` (ids in order). Admitted and
    // finished requests leave `pending`; the first request that must wait
    // for memory stops admission (FIFO, no overtaking of a waiting build).
    // Trace-only requests on a scene whose building wave is still in flight
    // stay queued (later requests may pass them).
    // `evicted` receives the keys of scenes dropped to make room.
    struct Decision {
        uint64_t id;
        Verdict v;
    };
    std::vector<Decision> form_wave(std::deque<PolicyRequest>& pending, double now_ms, uint64_t wave_id,
                                    std::vector<uint64_t>* evicted) {
        std::vector<Decision> out;
        uint32_t nreq = 0;
        uint64_t rays = 0, tris = 0, bytes = 0, cost = 0, cap = 0;
        const bool group = lim_.cost_ratio && (lim_.group_backlog == 0 || pending.size() <= lim_.group_backlog);
        for (auto it = pending.begin(); it != pending.end();) {
            PolicyRequest& r = *it;
            if (r.cancelled) { out.push_back({r.id, Verdict::Cancelled}); it = pending.erase(it); continue; }
            if (r.deadline_ms > 0 && now_ms > r.deadline_ms) {
                out.push_back({r.id, Verdict::Expired});
                it = pending.erase(it);
                continue;
            }
            const uint64_t need = r.work_bytes + (r.build ? r.scene_bytes : 0);
            if (need > budget_) { out.push_back({r.id, Verdict::Reject}); it = pending.erase(it); continue; }
            if (!r.build && !resident(r.key)) {
                out.push_back({r.id, Verdict::UnknownScene});
                it = pending.erase(it);
                continue;
            }
            if (!r.build) {
                // a scene still being built by a wave in flight is traced by a
                // later wave: waves never wait on each other on the device
                const Scene& sc = scenes_.at(r.key);
                if (!sc.ready && sc.wave != wave_id) {
                    ++it;
                    continue;
                }
            }
            if (nreq >= lim_.max_requests || (nreq && (rays + r.rays > lim_.max_rays || tris + r.tris > lim_.max_tris)))
                break;
            if (group && nreq && cost + request_cost(r) > cap) {
                ++it;  // another size class: a later wave
                continue;
            }
            if (resident_ + inflight_ + bytes + need > budget_) {
                make_room(resident_ + inflight_ + bytes + need - budget_, r.build ? r.key : ~0ull, evicted);
                if (resident_ + inflight_ + bytes + need > budget_) break;  // wait for memory
            }
            if (r.build) {
                auto old = scenes_.find(r.key);
                if (old != scenes_.end()) {
                    // rebuild under the same key: the old tree goes when unreferenced
                    if (old->second.refs) break;
                    drop(old->first);
                    if (evicted) evicted->push_back(r.key);
                }
                Scene s;
                s.bytes = r.scene_bytes;
                s.refs = 1;
                s.ready = false;
                s.wave = wave_id;
                lru_.push_front(r.key);
                s.lru = lru_.begin();
                scenes_[r.key] = s;
                resident_ += r.scene_bytes;
            } else {
                Scene& s = scenes_[r.key];
                s.refs++;
                lru_.erase(s.lru);
                lru_.push_front(r.key);
                s.lru = lru_.begin();
            }
            bytes += r.work_bytes;
            rays += r.rays;
            tris += r.build ? r.tris : 0;
            if (nreq == 0) {
                const uint64_t c = request_cost(r) * lim_.cost_ratio;
                cap = c > lim_.min_wave_cost ? c : lim_.min_wave_cost;
            }
            cost += request_cost(r);
            nreq++;
            out.push_back({r.id, Verdict::Admit});
            it = pending.erase(it);
        }
        inflight_ += bytes;
        return out;
    }

    // A wave finished: its transient bytes return and its requests release
    // their scenes; scenes it built become ready.
    void wave_done(uint64_t wave_id, uint64_t work_bytes, const std::vector<uint64_t>& keys) {
        inflight_ -= work_bytes;
        for (uint64_t k : keys) {
            auto it = scenes_.find(k);
            if (it == scenes_.end()) continue;
            if (it->second.refs) it->second.refs--;
            if (it->second.wave == wave_id) it->second.ready = true;
        }
    }

    // Drops an unreferenced scene; false if unknown or in use.
    bool evict(uint64_t key) {
        auto it = scenes_.find(key);
        if (it == scenes_.end() || it->second.refs) return false;
        drop(key);
        return true;
    }

   private:
    void drop(uint64_t key) {
        auto it = scenes_.find(key);
        resident_ -= it->second.bytes;
        lru_.erase(it->second.lru);
        scenes_.erase(it);
        evictions_++;
    }
    // Evicts least-recently-used unreferenced scenes until `need` bytes are
    // freed or none is left (never `keep`).
    void make_room(uint64_t need, uint64_t keep, std::vector<uint64_t>* evicted) {
        uint64_t freed = 0;
        const std::vector<uint64_t> order(lru_.rbegin(), lru_.rend());  // oldest first
        for (uint64_t k : order) {
            if (freed >= need) break;
            const Scene& s = scenes_[k];
            if (s.refs || k == keep) continue;
            freed += s.bytes;
            if (evicted) evicted->push_back(k);
            drop(k);
        }
    }

    uint64_t budget_;
    WaveLimits lim_;
    std::unordered_map<uint64_t, Scene> scenes_;
    std::list<uint64_t> lru_;  // front = most recent
    uint64_t resident_ = 0, inflight_ = 0, evictions_ = 0;
};

}  // namespace sq
