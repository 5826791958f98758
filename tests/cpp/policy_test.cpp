// CPU test of the serving loop's host-side policy (server_policy.h): the
// device-memory arena and admission / residency / eviction decisions.
// Built and run by tests/test_server_policy.py; exits non-zero on failure.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <vector>

#include "server_policy.h"

using namespace sq;

static int failures = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            failures++;                                            \
        }                                                          \
    } while (0)

static PolicyRequest build(uint64_t id, uint64_t key, uint64_t scene, uint64_t work, uint64_t tris = 1000,
                           uint64_t rays = 1000) {
    PolicyRequest r;
    r.id = id;
    r.key = key;
    r.build = true;
    r.scene_bytes = scene;
    r.work_bytes = work;
    r.tris = tris;
    r.rays = rays;
    return r;
}
static PolicyRequest trace(uint64_t id, uint64_t key, uint64_t work, uint64_t rays = 1000) {
    PolicyRequest r;
    r.id = id;
    r.key = key;
    r.work_bytes = work;
    r.rays = rays;
    return r;
}

static void test_arena() {
    uint64_t next = 1 << 20;
    int grows = 0;
    bool allow = true;
    Arena a(4096, [&](uint64_t bytes, uint64_t* base) {
        if (!allow) return false;
        *base = next;
        next += bytes + (1 << 20);
        grows++;
        return true;
    });
    uint64_t x, y, z, w;
    CHECK(a.alloc(100, &x));           // new chunk of 4096
    CHECK(grows == 1);
    CHECK(a.alloc(1000, &y));          // same chunk
    CHECK(grows == 1 && y == x + 256);
    CHECK(a.used() == 256 + 1024);
    CHECK(a.alloc(10000, &z));         // larger than a chunk: own chunk
    CHECK(grows == 2 && a.reserved() == 4096 + 10240);
    CHECK(a.release(y));
    CHECK(a.release(x));               // coalesces with y's block and the tail
    CHECK(a.largest_free() == 4096);
    CHECK(a.alloc(4096, &w) && w == x);  // whole first chunk again
    CHECK(!a.release(w + 256));        // not a block start
    allow = false;
    uint64_t v;
    CHECK(!a.alloc(1 << 20, &v));      // no room and no growth
    CHECK(a.release(z) && a.release(w));
    CHECK(a.used() == 0);
}

static void test_budget_and_queue() {
    Admission adm(10000);
    std::deque<PolicyRequest> q;
    q.push_back(build(1, 11, 3000, 1000));
    q.push_back(build(2, 12, 3000, 1000));
    q.push_back(build(3, 13, 3000, 1001));  // would exceed 10000: waits
    q.push_back(build(4, 14, 20000, 1));    // larger than the budget: rejected only when reached
    std::vector<uint64_t> ev;
    auto d = adm.form_wave(q, 0, 1, &ev);
    CHECK(d.size() == 2 && d[0].v == Verdict::Admit && d[1].v == Verdict::Admit);
    CHECK(q.size() == 2 && q.front().id == 3);
    CHECK(adm.resident_bytes() == 6000 && adm.inflight_bytes() == 2000);
    d = adm.form_wave(q, 0, 2, &ev);  // nothing evictable (both referenced)
    CHECK(d.empty() && q.size() == 2);
    adm.wave_done(1, 2000, {11, 12});
    CHECK(adm.inflight_bytes() == 0 && adm.scene(11)->ready && adm.scene(11)->refs == 0);
    d = adm.form_wave(q, 0, 3, &ev);  // evicts LRU scene 11, admits 3, rejects 4
    CHECK(d.size() == 2 && d[0].id == 3 && d[0].v == Verdict::Admit && d[1].id == 4 && d[1].v == Verdict::Reject);
    CHECK(ev.size() == 1 && ev[0] == 11 && !adm.resident(11) && adm.resident(12) && adm.resident(13));
    CHECK(adm.evictions() == 1);
}

static void test_trace_only_and_lru() {
    Admission adm(100000);
    std::deque<PolicyRequest> q;
    q.push_back(build(1, 1, 10000, 100));
    q.push_back(build(2, 2, 10000, 100));
    q.push_back(trace(3, 1, 100));    // same wave as its build
    q.push_back(trace(4, 99, 100));   // unknown scene
    std::vector<uint64_t> ev;
    auto d = adm.form_wave(q, 0, 1, &ev);
    CHECK(d.size() == 4 && d[2].v == Verdict::Admit && d[3].v == Verdict::UnknownScene);
    CHECK(adm.scene(1)->refs == 2);
    adm.wave_done(1, 300, {1, 2, 1});
    CHECK(adm.scene(1)->refs == 0 && adm.scene(2)->refs == 0);
    // touch scene 1 so scene 2 becomes least recently used
    q.push_back(trace(5, 1, 100));
    d = adm.form_wave(q, 0, 2, &ev);
    CHECK(d.size() == 1 && d[0].v == Verdict::Admit);
    // a build that needs room evicts 2 (LRU, unreferenced), never 1 (referenced)
    Admission small(25000);
    std::deque<PolicyRequest> q2;
    q2.push_back(build(1, 1, 10000, 100));
    q2.push_back(build(2, 2, 10000, 100));
    small.form_wave(q2, 0, 1, &ev);
    small.wave_done(1, 200, {1, 2});
    q2.push_back(trace(3, 1, 100));
    small.form_wave(q2, 0, 2, &ev);  // 1 referenced and most recent
    ev.clear();
    q2.push_back(build(4, 4, 10000, 100));
    auto d2 = small.form_wave(q2, 0, 3, &ev);
    CHECK(d2.size() == 1 && d2[0].v == Verdict::Admit && ev.size() == 1 && ev[0] == 2);
    CHECK(small.resident(1) && !small.resident(2) && small.resident(4));
    CHECK(!small.evict(1));  // in use
    small.wave_done(2, 100, {1});
    CHECK(small.evict(1) && !small.resident(1));
}

static void test_cancel_expire_limits() {
    WaveLimits lim;
    lim.max_requests = 2;
    lim.max_rays = 5000;
    Admission adm(1 << 30, lim);
    std::deque<PolicyRequest> q;
    PolicyRequest c = build(1, 1, 100, 100);
    c.cancelled = true;
    PolicyRequest e = build(2, 2, 100, 100);
    e.deadline_ms = 5.0;
    q.push_back(c);
    q.push_back(e);
    q.push_back(build(3, 3, 100, 100, 10, 3000));
    q.push_back(build(4, 4, 100, 100, 10, 3000));  // rays over the wave limit: next wave
    std::vector<uint64_t> ev;
    auto d = adm.form_wave(q, 10.0, 1, &ev);
    CHECK(d.size() == 3 && d[0].v == Verdict::Cancelled && d[1].v == Verdict::Expired && d[2].v == Verdict::Admit);
    CHECK(q.size() == 1 && q.front().id == 4);
    d = adm.form_wave(q, 10.0, 2, &ev);
    CHECK(d.size() == 1 && d[0].v == Verdict::Admit);
    // rebuild of a referenced key waits; once unreferenced it replaces the old scene
    q.push_back(build(5, 3, 200, 100));
    d = adm.form_wave(q, 10.0, 3, &ev);
    CHECK(d.empty());
    adm.wave_done(1, 100, {3});
    ev.clear();
    d = adm.form_wave(q, 10.0, 4, &ev);
    CHECK(d.size() == 1 && d[0].v == Verdict::Admit && adm.scene(3)->bytes == 200 && ev.size() == 1);
}

static void test_trace_waits_for_build_wave() {
    Admission adm(1 << 30);
    std::deque<PolicyRequest> q;
    q.push_back(build(1, 7, 100, 100));
    q.push_back(trace(2, 7, 100));  // same wave as the build: admitted
    std::vector<uint64_t> ev;
    auto d = adm.form_wave(q, 0, 1, &ev);
    CHECK(d.size() == 2 && d[1].v == Verdict::Admit);
    // while wave 1 is in flight, a trace of scene 7 waits for a later wave;
    // requests behind it still go
    q.push_back(trace(3, 7, 100));
    q.push_back(build(4, 8, 100, 100));
    d = adm.form_wave(q, 0, 2, &ev);
    CHECK(d.size() == 1 && d[0].id == 4 && d[0].v == Verdict::Admit);
    CHECK(q.size() == 1 && q.front().id == 3);
    adm.wave_done(1, 200, {7, 7});
    CHECK(adm.scene(7)->ready);
    d = adm.form_wave(q, 0, 3, &ev);
    CHECK(d.size() == 1 && d[0].id == 3 && d[0].v == Verdict::Admit);
}

static void test_size_grouping() {
    WaveLimits lim;
    lim.cost_ratio = 2;
    lim.min_wave_cost = 10000;
    Admission adm(1 << 30, lim);
    std::deque<PolicyRequest> q;
    q.push_back(build(1, 1, 100, 100, 1000, 2000));      // cost 4000: cap max(8000, 10000)
    q.push_back(build(2, 2, 100, 100, 1000000, 500000));  // cost 2.5M: another size class
    q.push_back(build(3, 3, 100, 100, 1000, 3000));       // joins: 9000
    q.push_back(build(4, 4, 100, 100, 1000, 3000));       // would make 14000: waits
    std::vector<uint64_t> ev;
    auto d = adm.form_wave(q, 0, 1, &ev);
    CHECK(d.size() == 2 && d[0].id == 1 && d[1].id == 3);
    CHECK(q.size() == 2 && q[0].id == 2 && q[1].id == 4);
    // the large request heads the queue: it forms a wave with what fits 2x its cost
    d = adm.form_wave(q, 0, 2, &ev);
    CHECK(d.size() == 2 && d[0].id == 2 && d[1].id == 4);
    // grouping off: one wave takes everything that fits the other limits
    Admission all(1 << 30);
    std::deque<PolicyRequest> q2;
    q2.push_back(build(1, 1, 100, 100, 1000, 2000));
    q2.push_back(build(2, 2, 100, 100, 1000000, 500000));
    d = all.form_wave(q2, 0, 1, &ev);
    CHECK(d.size() == 2 && q2.empty());
    // a backlog longer than group_backlog: grouping is off for that wave
    WaveLimits lb = lim;
    lb.group_backlog = 3;
    Admission busy(1 << 30, lb);
    std::deque<PolicyRequest> q3;
    q3.push_back(build(1, 1, 100, 100, 1000, 2000));
    q3.push_back(build(2, 2, 100, 100, 1000000, 500000));
    q3.push_back(build(3, 3, 100, 100, 1000, 3000));
    q3.push_back(build(4, 4, 100, 100, 1000, 3000));
    d = busy.form_wave(q3, 0, 1, &ev);
    CHECK(d.size() == 4 && q3.empty());
    // at or below it, grouping holds
    Admission calm(1 << 30, lb);
    std::deque<PolicyRequest> q4;
    q4.push_back(build(1, 1, 100, 100, 1000, 2000));
    q4.push_back(build(2, 2, 100, 100, 1000000, 500000));
    q4.push_back(build(3, 3, 100, 100, 1000, 3000));
    d = calm.form_wave(q4, 0, 1, &ev);
    CHECK(d.size() == 2 && d[0].id == 1 && d[1].id == 3);
}

int main() {
    test_arena();
    test_size_grouping();
    test_trace_waits_for_build_wave();
    test_budget_and_queue();
    test_trace_only_and_lru();
    test_cancel_expire_limits();
    if (failures) {
        std::printf("%d failure(s)\n", failures);
        return 1;
    }
    std::printf("policy tests passed\n");
    return 0;
}
